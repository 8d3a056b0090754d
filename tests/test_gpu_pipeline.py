"""GPU parity of pipelined frames (the reference's RendererT pattern: the
next frame's build runs on a second stream while this frame renders from a
copy of the last build).

A pipelined run must produce exactly the images of the sequential run, frame
for frame: its first frame builds before rendering, and every later frame
renders the build made during the frame before.  The sequential frames are
themselves checked against the oracle here (static scene) and in
test_gpu_dynamic.py (dynamic scene).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _frames(ctx, rtx, s, cam, w, h, bounces, n, pipelined, graph, animate=None, reset=None):
    R = rtx.Renderer(ctx, w, h, bounces)
    R.set_pipelined(pipelined)
    if animate is not None:
        R.set_animation(True, *animate)
    if reset is not None:
        reset()
    out = []
    R.frame_async(s, cam, rebuild=True)
    ctx.sync()
    out.append((R.fb.accum(), R.last_stats()))
    if graph:
        R.capture(s, cam, rebuild=True)
    for _ in range(n - 1):
        if graph:
            R.replay()
        else:
            R.frame_async(s, cam, rebuild=True)
        ctx.sync()
        out.append((R.fb.accum(), R.last_stats()))
    R.destroy()
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_pipelined_static_frames_match_oracle(ctx, graph):
    import rtx
    from rtx import scenes
    sd = scenes.config_c1()
    s = scenes.load(rtx.Scene(ctx), sd)
    w, h = 160, 120
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], w / h)
    got = _frames(ctx, rtx, s, cam, w, h, 2, 4, pipelined=True, graph=graph)
    # a static scene's builds all give the tree of the first rebuild from
    # the permuted order (the stable sort of sorted codes keeps the order)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    b = O.Build(sd.pos, b.idx.reshape(-1, 3), b.matmap, 30, 1)
    oacc, st, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam),
                               w, h, 2)
    for k, (acc, gst) in enumerate(got):
        assert np.array_equal(acc, oacc), f"pipelined frame {k} differs from the oracle"
        assert gst.primary_rays + gst.secondary_rays == orays and gst.shadow_rays == st["shadow"]
    s.destroy()


@pytest.mark.parametrize("graph", [False, True])
def test_pipelined_dynamic_frames_match_sequential(ctx, graph):
    import rtx
    from rtx import scenes
    sd = scenes.config_c3(n=97, res=(160, 96))

    def run(pipelined):
        s = scenes.load(rtx.Scene(ctx), sd)
        s.set_waves(sd.waves)
        cam = scenes.camera_of(sd)
        fr = _frames(ctx, rtx, s, cam, sd.width, sd.height, sd.max_bounces, 5, pipelined=pipelined, graph=graph,
                     animate=(0.0, 1.0 / 30.0), reset=lambda: s.set_clock(0))
        s.destroy()
        return fr

    seq = run(False)
    pip = run(True)
    for k, ((a, sa), (b, sb)) in enumerate(zip(seq, pip)):
        assert np.array_equal(a, b), f"pipelined frame {k} differs from sequential frame {k}"
        assert sa.secondary_rays == sb.secondary_rays and sa.shadow_rays == sb.shadow_rays
    # the surface moves, so consecutive frames differ
    assert not np.array_equal(seq[0][0], seq[1][0])


@pytest.mark.parametrize("depth", [1, 2, 3, 8])
def test_frames_in_flight_match_sequential_dynamic(ctx, depth):
    # the frame pipeline keeps `depth` frames in flight; every frame must be
    # the sequential run's (dynamic scene: each frame has its own geometry)
    import rtx
    from rtx import scenes
    sd = scenes.config_c3(n=97, res=(160, 96))
    cam = scenes.camera_of(sd)
    s = scenes.load(rtx.Scene(ctx), sd)       # both runs start from the uploaded order
    s.set_waves(sd.waves)
    nf = max(6, depth + 2)                    # every slot is reused at least once
    seq = _frames(ctx, rtx, s, cam, sd.width, sd.height, sd.max_bounces, nf, pipelined=False, graph=False,
                  animate=(0.0, 1.0 / 30.0), reset=lambda: s.set_clock(0))
    s.destroy()
    s = scenes.load(rtx.Scene(ctx), sd)
    s.set_waves(sd.waves)
    pl = rtx.Pipeline(ctx, s, sd.width, sd.height, sd.max_bounces, depth=depth)
    pl.set_animation(True, 0.0, 1.0 / 30.0)
    got = {}
    for k in range(nf):
        f = pl.frame(cam, rebuild=True)
        assert f == k
        if k >= depth - 1:          # the oldest frame in flight, before its slot is reused
            old = k - (depth - 1)
            pl.wait(old)
            got[old] = (pl.fb(old).accum(), pl.stats(old))
    for old in range(nf - (depth - 1), nf):
        pl.wait(old)
        got[old] = (pl.fb(old).accum(), pl.stats(old))
    pl.destroy()
    s.destroy()
    for k in range(nf):
        a, sa = seq[k]
        b, sb = got[k]
        assert np.array_equal(a, b), f"frame {k} (depth {depth}) differs from the sequential frame"
        assert sa.secondary_rays == sb.secondary_rays and sa.shadow_rays == sb.shadow_rays
