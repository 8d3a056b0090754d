"""GPU parity at the benchmark workloads' real sizes (C2-C5 below), beside
the smaller cases of the other GPU tests.

  C2: the headline frame -- 200 spheres (1,000,000 triangles), 1920x1080,
      2 point lights, Whitted depth 3, LBVH rebuilt every frame -- through
      the plain path and through frames in flight (3 at once, one of the
      depths the bench sweeps), every frame's accumulators bit-exact against the
      oracle's render of the same frame.
  C3: two animated frames of the 724x724 grid (1,045,458 triangles),
      1280x720, depth 2, rebuilt every frame, against the oracle on the
      read-back vertices.
  C4: the 64-scene batch (20M triangles, 256^2 to 1080p, depth 1-4) through
      the wavefront renderer: every frame equals its stand-alone frame bit for
      bit, and the two smallest frames equal the oracle.
  C5: the 12 resident scenes (20M triangles) served at 2000 requests/s:
      sampled requests' images in the mix equal the same requests rendered
      alone, and requests overlap on the device.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rtx():
    import rtx
    return rtx


@pytest.fixture(scope="module")
def c2_full(ctx):
    rtx = _rtx()
    from rtx import scenes
    sd = scenes.config_c2()
    assert sd.T == 1_000_000 and (sd.width, sd.height, sd.max_bounces) == (1920, 1080, 3)
    cam = scenes.camera_of(sd)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    oacc, ost, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam),
                                sd.width, sd.height, sd.max_bounces)
    return sd, cam, b, oacc, orays, int(ost["shadow"])


def test_c2_plain_frame_full_size(ctx, c2_full):
    rtx = _rtx()
    from rtx import scenes
    sd, cam, b, oacc, orays, oshadow = c2_full
    s = scenes.load(rtx.Scene(ctx), sd)
    R = rtx.Renderer(ctx, sd.width, sd.height, sd.max_bounces)
    try:
        for k in range(2):  # the second frame rebuilds from the permuted buffers
            st = R.frame(s, cam, rebuild=True)
            assert st.primary_rays + st.secondary_rays == orays, f"frame {k}: ray count"
            assert st.shadow_rays == oshadow, f"frame {k}: shadow ray count"
            assert np.array_equal(R.fb.accum(), oacc), f"frame {k}: accumulators differ from the oracle"
        assert np.array_equal(s.nodes()["a"], b.nodes["a"]), "BVH topology"
    finally:
        R.destroy()
        s.destroy()


def test_c2_frames_in_flight_full_size(ctx, c2_full):
    # the bench's headline mode: 3 frames in flight, each rebuilding the LBVH
    rtx = _rtx()
    from rtx import scenes
    sd, cam, b, oacc, orays, oshadow = c2_full
    s = scenes.load(rtx.Scene(ctx), sd)
    pl = rtx.Pipeline(ctx, s, sd.width, sd.height, sd.max_bounces, depth=3)
    try:
        frames = [pl.frame(cam, rebuild=True) for _ in range(5)]
        for f in frames[-3:]:
            pl.wait(f)
            st = pl.stats(f)
            assert st.primary_rays + st.secondary_rays == orays, f"frame {f}: ray count"
            assert np.array_equal(pl.fb(f).accum(), oacc), f"frame {f}: accumulators differ from the oracle"
    finally:
        pl.destroy()
        s.destroy()


def test_c3_animated_frames_full_size(ctx):
    rtx = _rtx()
    from rtx import scenes
    sd = scenes.config_c3()
    assert sd.T == 1_045_458 and (sd.width, sd.height, sd.max_bounces) == (1280, 720, 2)
    s = scenes.load(rtx.Scene(ctx), sd)
    s.set_waves(sd.waves)
    cam = scenes.camera_of(sd)
    R = rtx.Renderer(ctx, sd.width, sd.height, sd.max_bounces)
    R.set_animation(True, 0.0, 1.0 / 30.0)
    lights, ocam = O.lights_from(sd.lights), O.camera_from(cam)
    idx, mm = sd.idx, sd.matmap
    try:
        prev = None
        for k in range(2):
            R.frame_async(s, cam, rebuild=True)
            ctx.sync()
            pos, nrm = s.vertices()
            b = O.Build(pos, idx, mm, 30, 1)
            oacc, ost, orays = O.render(b, nrm, sd.uv, sd.materials, lights, ocam, sd.width, sd.height,
                                        sd.max_bounces)
            st = R.last_stats()
            assert st.primary_rays + st.secondary_rays == orays, f"frame {k}: ray count"
            assert np.array_equal(R.fb.accum(), oacc), f"frame {k}: accumulators differ from the oracle"
            idx, mm = b.idx.reshape(-1, 3), b.matmap
            if prev is not None:
                assert not np.array_equal(prev[:, 1], pos[:, 1]), "the surface moves between frames"
            prev = pos
    finally:
        R.destroy()
        s.destroy()


def test_c4_batch_full_size(ctx):
    rtx = _rtx()
    from rtx import scenes
    frames = scenes.config_c4()
    assert len(frames) == 64
    scs, cams = [], []
    for fr in frames:
        sc = scenes.load(rtx.Scene(ctx), fr.sd)
        sc.build(30, 1)
        scs.append(sc)
        cams.append(scenes.orbit_camera(fr.sd, fr.direction, fr.width, fr.height))
    need = sum(rtx.Wavefront.frame_bytes(f.width, f.height, f.bounces, f.ray_capacity) for f in frames)
    wf = rtx.Wavefront(ctx, arena_bytes=need + (64 << 20), max_frames=64)
    try:
        ids = [wf.add(sc, cam, f.width, f.height, f.bounces, tag=i, ray_capacity=f.ray_capacity)
               for i, (sc, cam, f) in enumerate(zip(scs, cams, frames))]
        wf.run()
        wf.sync()
        total = 0
        for i, (sc, cam, f) in enumerate(zip(scs, cams, frames)):
            R = rtx.Renderer(ctx, f.width, f.height, f.bounces, ray_capacity=f.ray_capacity)
            st = R.frame(sc, cam, rebuild=False)
            ws = wf.stats(ids[i])
            assert (ws.primary_rays, ws.secondary_rays, ws.shadow_rays) == \
                (st.primary_rays, st.secondary_rays, st.shadow_rays), f"frame {i}: ray counts"
            assert np.array_equal(wf.accum(ids[i]), R.fb.accum()), f"frame {i}: differs from its stand-alone frame"
            R.destroy()
            total += st.primary_rays + st.secondary_rays
        assert total > 50_000_000
        # the two cheapest frames against the oracle
        cost = [f.width * f.height * (1 << f.bounces) + f.sd.T // 10 for f in frames]
        for i in np.argsort(cost)[:2]:
            f = frames[i]
            b = O.Build(f.sd.pos, f.sd.idx, f.sd.matmap, 30, 1)
            oacc, _, orays = O.render(b, f.sd.nrm, f.sd.uv, f.sd.materials, O.lights_from(f.sd.lights),
                                      O.camera_from(cams[i]), f.width, f.height, f.bounces,
                                      cap=f.ray_capacity)
            assert np.array_equal(wf.accum(ids[i]), oacc), f"frame {i}: differs from the oracle"
    finally:
        wf.destroy()
        for sc in scs:
            sc.destroy()


def test_c5_served_requests_full_size(ctx):
    rtx = _rtx()
    from rtx import scenes
    sds = scenes.config_c5()
    assert len(sds) == 12 and sum(sd.T for sd in sds) > 19_000_000
    scs = []
    for sd in sds:
        sc = scenes.load(rtx.Scene(ctx), sd)
        sc.build(30, 1)
        scs.append(sc)
    srv = rtx.Server(ctx, workers=64, max_bounces=3, arena_bytes=16 << 30)
    try:
        ids = [srv.add_scene(sc) for sc in scs]
        reqs = scenes.c5_requests(sds, 2000.0, 300, 2000 * 7 + 1)
        sampled = list(range(0, 300, 19))
        start = rtx.now_ns() + 5_000_000
        for i, r in enumerate(reqs):
            r["scene"] = ids[r["scene"]]
            r["arrival_ns"] += start
            if i in sampled:
                r["out"] = np.zeros(r["width"] * r["height"] * 3, np.uint32)
        srv.submit(reqs)
        srv.drain(120_000)
        done = {c["tag"]: c for c in srv.poll(400)}
        assert sorted(done) == list(range(300)) and all(c["status"] == 0 for c in done.values())
        spans = sorted((c["start_ns"], c["done_ns"]) for c in done.values())
        assert sum(1 for a, b in zip(spans, spans[1:]) if b[0] < a[1]) > 50, "requests overlap on the device"
        for i in sampled:
            r = reqs[i]
            R = rtx.Renderer(ctx, r["width"], r["height"], r["bounces"])
            st = R.frame(scs[r["scene"]], r["camera"], rebuild=False)
            assert np.array_equal(r["out"], R.fb.accum().reshape(-1)), f"request {i} differs alone and in the mix"
            assert done[i]["rays"] == st.primary_rays + st.secondary_rays
            R.destroy()
    finally:
        srv.destroy()
        for sc in scs:
            sc.destroy()
