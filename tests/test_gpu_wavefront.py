"""GPU tests of the wavefront renderer (rtx_wavefront_*): frames of several
scenes, sizes and depths rendered together, one launch per stage per step.

Bar: every frame's accumulators, RGBA8 and ray counts equal its stand-alone
frame's (rtx_render_frame, no rebuild: the reference's sequence,
src/rt/renderer.cpp:113-214) bit for bit, in every traversal form, whether
the frames start together or at different steps; one frame is also checked
against the oracle.  Admission: frames take rtx_wavefront_frame_bytes() of
the arena, a full arena refuses a frame (RTX_ENOMEM) until one is released,
and a frame larger than the arena is rejected (RTX_EINVAL).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rtx():
    import rtx
    return rtx


def _chain_scene(ctx):
    rtx = _rtx()
    from chain import chain_scene
    pos, idx = chain_scene()
    s = rtx.Scene(ctx)
    nrm = np.tile(np.array([[0.0, 0.0, -1.0]], np.float32), (len(pos), 1))
    s.set_geometry(pos, idx, nrm, None, np.arange(len(idx), dtype=np.int32) % 2)
    s.set_materials(rtx.make_materials([(0.8, 0.3, 0.2, 0.5, 0.5, 0.0), (0.2, 0.7, 0.3, 0.2, 0.0, 1.4)]))
    s.set_lights(rtx.make_lights((0.2, 0.2, 0.2), [dict(type=rtx.LIGHT_POINT, pos=(0.2, 0.3, -1.0))]))
    s.build(32, 1)
    cent = pos[idx[4:]].mean(axis=(0, 1))
    cam = rtx.camera_set(cent + np.array([0.004, 0.003, -0.02]), cent, (0, 1, 0), 0.35, 1.0)
    return s, cam


@pytest.fixture(scope="module")
def wf_scenes(ctx):
    rtx = _rtx()
    from rtx import scenes
    out = {}
    sd = scenes.config_c1()
    a = scenes.load(rtx.Scene(ctx), sd)
    a.build(30, 1)
    out["c1"] = (a, scenes.camera_of(sd), sd)
    sd2 = scenes.config_c2(n_spheres=24, res=(320, 200))
    b = scenes.load(rtx.Scene(ctx), sd2)
    b.build(30, 1)
    out["spheres"] = (b, scenes.camera_of(sd2), sd2)
    sd3 = scenes.config_c3(n=97, res=(160, 96))
    sd3.pos[:, 1] = scenes.displace_host(sd3, 0.7)[:, 1]
    c = scenes.load(rtx.Scene(ctx), sd3)
    c.build(30, 1)
    out["water"] = (c, scenes.camera_of(sd3), sd3)
    # C1's geometry, textured, with a cubemap and a spot light
    d = scenes.load(rtx.Scene(ctx), sd)
    rng = np.random.default_rng(3)
    mats = sd.materials.copy().view(np.int32)
    mats[7] = 0  # the ground samples texture 0
    d.set_materials(mats.view(np.uint8))
    d.set_textures([rng.integers(0, 256, (16, 24, 4), dtype=np.uint8)])
    d.set_cubemap([rng.integers(0, 256, (8, 8, 4), dtype=np.uint8) for _ in range(6)])
    d.set_lights(rtx.make_lights((0.1, 0.1, 0.1), [dict(type=rtx.LIGHT_SPOT, pos=(1.0, 6.0, -2.0),
                                                       dir=(-0.1690309, -0.9860133, 0.0), angle=0.93)]))
    d.build(30, 1)
    out["textured"] = (d, rtx.camera_set((0.0, 2.5, -6.0), (0.0, 1.2, 0.0), (0.0, 1.0, 0.0), sd.camera["fov"], 1.0),
                       sd)
    e, ecam = _chain_scene(ctx)
    out["chain"] = (e, ecam, None)
    yield out
    for s, _, _ in out.values():
        s.destroy()


# (scene, width, height, bounces, ray_capacity)
FRAMES = [("c1", 256, 256, 1, 0), ("spheres", 200, 150, 3, 0), ("water", 160, 96, 2, 0),
          ("textured", 97, 61, 2, 0), ("chain", 64, 64, 2, 0), ("spheres", 320, 200, 0, 0),
          ("c1", 33, 31, 3, 0), ("spheres", 128, 128, 2, 700), ("water", 64, 48, 1, 0), ("textured", 128, 96, 0, 0)]


def _stand_alone(ctx, s, cam, w, h, b, cap):
    rtx = _rtx()
    R = rtx.Renderer(ctx, w, h, b, ray_capacity=cap)
    try:
        st = R.frame(s, cam, rebuild=False)
        return R.fb.accum(), R.fb.rgba8(), (st.primary_rays, st.secondary_rays, st.shadow_rays, st.bounces_run)
    finally:
        R.destroy()


@pytest.fixture(scope="module")
def expected(ctx, wf_scenes):
    out = []
    for name, w, h, b, cap in FRAMES:
        s, cam, _ = wf_scenes[name]
        out.append(_stand_alone(ctx, s, cam, w, h, b, cap))
    return out


def _check(wf, fid, exp, what):
    acc, rgba, counts = exp
    st = wf.stats(fid)
    got = (st.primary_rays, st.secondary_rays, st.shadow_rays, st.bounces_run)
    assert got == counts, f"{what}: ray counts {got} != stand-alone {counts}"
    assert np.array_equal(wf.accum(fid), acc), f"{what}: accumulators differ from the stand-alone frame"
    assert np.array_equal(wf.rgba8(fid), rgba), f"{what}: RGBA8 differs from the stand-alone frame"


def test_frames_together_equal_stand_alone(ctx, wf_scenes, expected):
    rtx = _rtx()
    wf = rtx.Wavefront(ctx, arena_bytes=2 << 30, max_frames=32)
    try:
        ids = []
        for k, (name, w, h, b, cap) in enumerate(FRAMES):
            s, cam, _ = wf_scenes[name]
            ids.append(wf.add(s, cam, w, h, b, tag=100 + k, ray_capacity=cap))
        wf.run()
        wf.sync()
        done = wf.poll()
        assert sorted(t for _, t in done) == [100 + k for k in range(len(FRAMES))]
        for k, fid in enumerate(ids):
            _check(wf, fid, expected[k], f"frame {k} {FRAMES[k]}")
        assert any(e[2][1] > 0 for e in expected), "some frames must have secondary rays"
        assert expected[7][2][1] <= 3 * 700, "the capped frame keeps at most its capacity per bounce"
    finally:
        wf.destroy()


@pytest.mark.parametrize("variant,block,sched,stack_lds,shadow_order,share", [
    (0, 64, 0, 32, 0, 0), (1, 128, 0, 32, 0, 0), (1, 128, 1, 32, 0, 0), (2, 256, 0, 32, 0, 0), (2, 128, 1, 32, 0, 0),
    (3, 128, 0, 32, 0, 0), (4, 64, 1, 32, 0, 0), (1, 128, 1, 16, 0, 0), (1, 128, 0, 16, 0, 0), (2, 256, 1, 16, 0, 0),
    (0, 64, 1, 8, 0, 0), (1, 256, 0, 8, 0, 0), (1, 128, 1, 32, 1, 0), (0, 128, 0, 32, 1, 0), (1, 128, 1, 32, 0, 65),
    (1, 64, 1, 16, 0, 1)])
def test_every_traversal_form(ctx, wf_scenes, expected, variant, block, sched, stack_lds, shadow_order, share):
    # the sibling-pair form falls back to the single-node form when a scene
    # has 32 levels (the chain); the persistent and stackless forms map to it;
    # sched: closest hit in launch order (0) or by persistent waves (1);
    # stack_lds: traversal stack entries in LDS (the rest in private memory);
    # shadow_order: shadow rays in the reference's child order (1) or in index
    # order (0; the chain scene's 32 levels keep the reference's order);
    # share: percent of the resident blocks a persistent launch takes (1: one
    # block per CU)
    rtx = _rtx()
    old = (ctx.tuned("trace_variant"), ctx.tuned("trace_block"), ctx.tuned("wf_sched"), ctx.tuned("stack_lds"),
           ctx.tuned("shadow_order"), ctx.tuned("wf_share"))
    ctx.tune("trace_variant", variant)
    ctx.tune("trace_block", block)
    ctx.tune("wf_sched", sched)
    ctx.tune("stack_lds", stack_lds)
    ctx.tune("shadow_order", shadow_order)
    ctx.tune("wf_share", share)
    wf = rtx.Wavefront(ctx, arena_bytes=2 << 30, max_frames=32)
    try:
        ids = []
        for name, w, h, b, cap in FRAMES:
            s, cam, _ = wf_scenes[name]
            ids.append(wf.add(s, cam, w, h, b, ray_capacity=cap))
        wf.run()
        for k, fid in enumerate(ids):
            _check(wf, fid, expected[k], f"variant {variant} block {block} frame {k}")
    finally:
        wf.destroy()
        ctx.tune("trace_variant", old[0])
        ctx.tune("trace_block", old[1])
        ctx.tune("wf_sched", old[2])
        ctx.tune("stack_lds", old[3])
        ctx.tune("shadow_order", old[4])
        ctx.tune("wf_share", old[5])


def test_frames_joining_at_different_steps(ctx, wf_scenes, expected):
    # continuous admission: frames join while others are mid-way; released
    # frames' memory and slots are reused
    rtx = _rtx()
    wf = rtx.Wavefront(ctx, arena_bytes=2 << 30, max_frames=4)
    try:
        order = list(range(len(FRAMES)))
        live = {}
        checked = set()
        steps = 0
        while order or live:
            while order and len(live) < 3:
                k = order.pop(0)
                name, w, h, b, cap = FRAMES[k]
                s, cam, _ = wf_scenes[name]
                fid = wf.add(s, cam, w, h, b, tag=k, ray_capacity=cap)
                live[fid] = k
            wf.step()
            steps += 1
            for fid, tag in wf.poll():
                _check(wf, fid, expected[tag], f"frame {tag} (joined mid-way)")
                wf.release(fid)
                checked.add(tag)
                del live[fid]
            if not order and live and steps > 100:
                wf.sync()
        assert checked == set(range(len(FRAMES)))
        assert wf.memory()[0] == 0, "every frame's memory returned to the arena"
    finally:
        wf.destroy()


def test_arena_admission_and_release(ctx, wf_scenes):
    rtx = _rtx()
    s, cam, _ = wf_scenes["c1"]
    fb = rtx.Wavefront.frame_bytes(128, 128, 2)
    assert fb > 128 * 128 * 16
    wf = rtx.Wavefront(ctx, arena_bytes=2 * fb + fb // 2, max_frames=8)
    try:
        a = wf.try_add(s, cam, 128, 128, 2)
        b = wf.try_add(s, cam, 128, 128, 2)
        assert a is not None and b is not None
        assert wf.try_add(s, cam, 128, 128, 2) is None, "a full arena must refuse the frame (RTX_ENOMEM)"
        assert wf.memory()[0] == 2 * fb
        with pytest.raises(rtx.RtxError, match="larger than the whole arena|more memory than the whole arena"):
            wf.add(s, cam, 1024, 1024, 3)
        with pytest.raises(rtx.RtxError, match="not complete"):
            wf.release(a)
        wf.run()
        wf.sync()
        acc_a = wf.accum(a)
        wf.release(a)
        c = wf.try_add(s, cam, 128, 128, 2)
        assert c is not None, "released memory must admit a new frame"
        wf.run()
        assert np.array_equal(wf.accum(c), acc_a), "the same frame rendered again"
        wf.release(b)
        wf.release(c)
        assert wf.memory()[0] == 0
    finally:
        wf.destroy()


def test_wavefront_frame_matches_oracle(ctx, wf_scenes):
    rtx = _rtx()
    s, cam, sd = wf_scenes["c1"]
    w, h = 160, 128
    wf = rtx.Wavefront(ctx, arena_bytes=1 << 30, max_frames=4)
    try:
        other = wf.add(*wf_scenes["spheres"][:2], 96, 64, 2)
        fid = wf.add(s, cam, w, h, 1)
        wf.run()
        acc = wf.accum(fid)
        st = wf.stats(fid)
        assert wf.stats(other).primary_rays == 96 * 64
    finally:
        wf.destroy()
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    oacc, ost, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h, 1)
    assert st.primary_rays + st.secondary_rays == orays
    assert np.array_equal(acc, oacc), "wavefront frame differs from the oracle"
