"""GPU parity of the dynamic scene (the rt-wave pattern, config C3 at test
size): per frame the heightfield is displaced on the device, the LBVH is
rebuilt from the previous build's permuted buffers, and the frame is
rendered, eagerly and as replays of one captured hipGraph whose device clock
advances the animation.

The displacement itself is outside the bit-exact contract (it feeds the
path); each frame's displaced vertices are read back and given to the oracle,
whose build and render must then match the GPU's framebuffer bit for bit.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _setup(ctx, n=97, res=(160, 96)):
    import rtx
    from rtx import scenes
    sd = scenes.config_c3(n=n, res=res)
    s = scenes.load(rtx.Scene(ctx), sd)
    s.set_waves(sd.waves)
    cam = scenes.camera_of(sd)
    return rtx, sd, s, cam


def test_animate_matches_host_time(ctx):
    # animate() at clock c equals an explicit displacement at t0 + dt*c
    # (float32 product, then sum), and advances the clock
    rtx, sd, s, _ = _setup(ctx, n=65, res=(32, 32))
    t0, dt = np.float32(0.25), np.float32(1.0 / 30.0)
    s.set_clock(5)
    s.animate(float(t0), float(dt))
    assert s.clock() == 6
    pos_a, nrm_a = s.vertices()
    from rtx import scenes
    s2 = scenes.load(rtx.Scene(ctx), sd)
    t = np.float32(t0 + np.float32(dt * np.float32(5)))
    s2.displace_waves(sd.waves, float(t))
    pos_b, nrm_b = s2.vertices()
    assert np.array_equal(pos_a, pos_b) and np.array_equal(nrm_a, nrm_b)
    # the displacement formula itself, against a float64 evaluation
    p = sd.pos.astype(np.float64)
    y = np.zeros(p.shape[0])
    for a, k, dx, dz, om, ph in sd.waves.astype(np.float64):
        y += a * np.sin(k * (dx * p[:, 0] + dz * p[:, 2]) - om * float(t) + ph)
    assert np.allclose(pos_a[:, 1], y, atol=2e-5)
    assert np.array_equal(pos_a[:, [0, 2]], sd.pos[:, [0, 2]])
    assert np.allclose(np.linalg.norm(nrm_a, axis=1), 1.0, atol=1e-6)
    s.destroy()
    s2.destroy()


def test_dynamic_frames_eager_and_graph_match_oracle(ctx):
    rtx, sd, s, cam = _setup(ctx)
    w, h = sd.width, sd.height
    R = rtx.Renderer(ctx, w, h, sd.max_bounces)
    R.set_animation(True, 0.0, 1.0 / 30.0)
    lights = O.lights_from(sd.lights)
    ocam = O.camera_from(cam)
    state = {"idx": sd.idx, "mm": sd.matmap}

    def check(tag):
        pos, nrm = s.vertices()
        b = O.Build(pos, state["idx"], state["mm"], 30, 1)
        oacc, ost, orays = O.render(b, nrm, sd.uv, sd.materials, lights, ocam, w, h, sd.max_bounces)
        assert np.array_equal(R.fb.accum(), oacc), f"{tag}: framebuffer differs"
        st = R.last_stats()
        assert st.primary_rays + st.secondary_rays == orays, f"{tag}: ray count differs"
        assert st.shadow_rays == ost["shadow"], f"{tag}: shadow ray count differs"
        state["idx"], state["mm"] = b.idx.reshape(-1, 3), b.matmap   # the next build's input order
        return pos

    R.frame_async(s, cam, rebuild=True)      # clock 0
    ctx.sync()
    p0 = check("eager frame 0")
    R.capture(s, cam, rebuild=True)
    seen = [p0]
    for k in range(3):                       # clocks 1, 2, 3
        R.replay()
        ctx.sync()
        assert s.clock() == k + 2
        seen.append(check(f"replay {k}"))
    R.frame(s, cam, rebuild=True)           # plain path, clock 4
    seen.append(check("plain frame"))
    # the surface really moves from frame to frame
    for a, b in zip(seen, seen[1:]):
        assert not np.array_equal(a[:, 1], b[:, 1])
    R.destroy()
    s.destroy()
