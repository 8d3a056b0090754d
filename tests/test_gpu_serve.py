"""GPU tests of the request server (resident scenes, continuous batching on
the wavefront renderer).

A request's image must be the same in the concurrent mix as rendered alone:
each served request returns its accumulators, which are compared bit for bit
with a stand-alone frame of the same scene, camera, size and depth (and one
of them with the oracle), and requests must overlap on the device.
Admission: scenes by the HBM budget; requests by the bytes they take from
the request arena, first come first served (a tiny arena serializes them;
a request larger than the arena is refused); invalid requests; and the plain
one-at-a-time baseline.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CLASSES = [(64, 64), (128, 96)]


@pytest.fixture(scope="module")
def scenes2(ctx):
    import rtx
    from rtx import scenes
    sd1 = scenes.config_c1()
    s1 = scenes.load(rtx.Scene(ctx), sd1)
    s1.build(30, 1)
    sd2 = scenes.config_c2(n_spheres=40, res=(128, 96), slices=24, stacks=16)
    s2 = scenes.load(rtx.Scene(ctx), sd2)
    s2.build(30, 1)
    yield [(sd1, s1), (sd2, s2)]
    s1.destroy()
    s2.destroy()


def _requests(rtx, n, seed, now):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        sc = int(rng.integers(0, 2))
        w, h = CLASSES[int(rng.integers(0, len(CLASSES)))]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        centre = np.array([0.0, 0.5, 0.0]) if sc == 0 else np.zeros(3)
        radius = 8.0 if sc == 0 else 30.0
        pos = centre + radius * d
        cam = rtx.camera_set(pos, centre, (0.0, 1.0, 0.0), np.pi / 4, w / h)
        reqs.append(dict(scene=sc, width=w, height=h, bounces=int(rng.integers(1, 4)), camera=cam,
                         arrival_ns=now + int(i * 200_000), tag=i, out=np.zeros(w * h * 3, np.uint32)))
    return reqs


def _alone(ctx, rtx, s, r):
    R = rtx.Renderer(ctx, r["width"], r["height"], r["bounces"])
    R.frame_async(s, r["camera"], rebuild=False)
    ctx.sync()
    acc = R.fb.accum().reshape(-1)
    st = R.last_stats()
    R.destroy()
    return acc, st


@pytest.mark.parametrize("plain,groups,step_pixels", [(False, 1, None), (False, 3, None), (True, 0, None),
                                                      (False, 2, 1), (False, 2, 20_000)])
def test_served_images_equal_stand_alone(ctx, scenes2, plain, groups, step_pixels):
    # step_pixels: new primary pixels a group's step takes at most (1: one new
    # request a step; the others wait at the head for the next step)
    import rtx
    old_cap = ctx.tuned("srv_step_pixels")
    if step_pixels is not None:
        ctx.tune("srv_step_pixels", step_pixels)
    try:
        srv = rtx.Server(ctx, CLASSES, workers=16, max_bounces=3, plain=plain, arena_bytes=1 << 30, groups=groups)
    finally:
        ctx.tune("srv_step_pixels", old_cap)
    ids = [srv.add_scene(s) for _, s in scenes2]
    assert ids == [0, 1]
    reqs = _requests(rtx, 24, 7, rtx.now_ns() + 2_000_000)
    srv.submit(reqs)
    srv.drain(60_000)
    done = {c["tag"]: c for c in srv.poll()}
    srv.destroy()
    assert sorted(done) == list(range(len(reqs)))
    for r in reqs:
        c = done[r["tag"]]
        assert c["status"] == 0
        assert c["arrival_ns"] == r["arrival_ns"] and c["start_ns"] >= c["arrival_ns"] and c["done_ns"] >= c["start_ns"]
        acc, st = _alone(ctx, rtx, scenes2[r["scene"]][1], r)
        assert np.array_equal(r["out"], acc), f"request {r['tag']} differs from its stand-alone frame"
        assert c["rays"] == st.primary_rays + st.secondary_rays
    spans = sorted((c["start_ns"], c["done_ns"]) for c in done.values())
    overlaps = sum(1 for a, b in zip(spans, spans[1:]) if b[0] < a[1])
    if plain:
        assert overlaps == 0, "the plain baseline serves one request at a time"
    else:
        assert overlaps >= 2, "requests in flight together must overlap on the device"
    # anchor one request on the oracle
    r = reqs[0]
    sd, _ = scenes2[r["scene"]]
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    oacc, _, _ = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(r["camera"]),
                          r["width"], r["height"], r["bounces"])
    assert np.array_equal(r["out"], oacc.reshape(-1))


def test_admission_and_invalid_requests(ctx, scenes2):
    import rtx
    arena = 64 << 20
    srv = rtx.Server(ctx, workers=2, max_bounces=2, arena_bytes=arena)
    wb, sb = srv.memory()
    assert wb >= arena and sb == 0
    srv.destroy()
    # a budget that holds the request arena and one scene, not two
    _, s1 = scenes2[0]
    _, s2 = scenes2[1]
    srv = rtx.Server(ctx, workers=2, max_bounces=2, arena_bytes=arena, hbm_budget=wb + 1)
    with pytest.raises(rtx.RtxError):
        srv.add_scene(s1)
    srv.destroy()
    srv = rtx.Server(ctx, workers=2, max_bounces=2, arena_bytes=arena)
    sid = srv.add_scene(s1)
    cam = rtx.camera_set((0.0, 4.0, -8.0), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), np.pi / 4, 1.0)
    now = rtx.now_ns()
    assert rtx.Wavefront.frame_bytes(2048, 2048, 2) > arena
    srv.submit([dict(scene=sid, width=50, height=50, bounces=1, camera=cam, arrival_ns=now, tag=1),     # any size
                dict(scene=7, width=64, height=64, bounces=1, camera=cam, arrival_ns=now, tag=2),      # no such scene
                dict(scene=sid, width=64, height=64, bounces=3, camera=cam, arrival_ns=now, tag=3),    # too deep
                dict(scene=sid, width=2048, height=2048, bounces=2, camera=cam, arrival_ns=now, tag=4),  # > arena
                dict(scene=sid, width=64, height=64, bounces=2, camera=cam, arrival_ns=now, tag=5)])
    srv.drain(30_000)
    st = {c["tag"]: c["status"] for c in srv.poll()}
    srv.destroy()
    assert st[1] == 0 and st[2] == -1 and st[3] == -1 and st[4] == -3 and st[5] == 0


def test_tiny_arena_queues_requests(ctx, scenes2):
    # room for one request at a time: the others wait first come first
    # served, start only after the one before has finished and released its
    # memory, and still equal their stand-alone frames
    import rtx
    one = rtx.Wavefront.frame_bytes(128, 96, 3)
    srv = rtx.Server(ctx, workers=8, max_bounces=3, arena_bytes=one + one // 2, groups=1)
    sid = [srv.add_scene(s) for _, s in scenes2]
    reqs = _requests(rtx, 6, 11, rtx.now_ns() + 1_000_000)
    for i, r in enumerate(reqs):
        r.update(width=128, height=96, bounces=3, arrival_ns=reqs[0]["arrival_ns"] + i,
                 out=np.zeros(128 * 96 * 3, np.uint32))
        r["camera"] = rtx.camera_set(*_cam_args(r["scene"]), np.pi / 4, 128 / 96)
        r["scene"] = sid[r["scene"]]
    srv.submit(reqs)
    srv.drain(60_000)
    done = sorted(srv.poll(), key=lambda c: c["start_ns"])
    srv.destroy()
    assert [c["status"] for c in done] == [0] * 6
    assert [c["tag"] for c in done] == list(range(6)), "admitted first come first served"
    for a, b in zip(done, done[1:]):
        assert b["start_ns"] >= a["done_ns"], "a request may start only once the one before has released its memory"
    for r in reqs:
        acc, _ = _alone(ctx, rtx, scenes2[r["scene"]][1], r)
        assert np.array_equal(r["out"], acc)


def _cam_args(sc):
    centre = np.array([0.0, 0.5, 0.0]) if sc == 0 else np.zeros(3)
    return (centre + np.array([3.0, 5.0, -8.0 if sc == 0 else -30.0]), centre, (0.0, 1.0, 0.0))
