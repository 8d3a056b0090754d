"""GPU traversal over CPU SAH trees (rtx_sah_build, the reference's
alternative accelerator, src/rt/bvh.cpp): the scene's tree, index buffer and
material map after the build equal the reference builder's (the committed
outputs of its own code), and the hits and shadow flags of every traversal
form, and whole frames, equal the oracle's traversal over that same
reference tree, bit for bit, for the stack forms (0-3) and the stackless
walk (4, trace-bvh-no-stack.cl).  Cases: a 5000-triangle soup, geometry below
zero, duplicated triangles, a flat axis, a regular grid, the 36-level tree
(beyond the stack: the MAX_LEVELS rule and the fallback from the
sibling-pair form) and a shaded terrain scene."""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from sah_cases import terrain_scene
from test_gpu_parity import load_scene

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden" / "sah"


def axis_rays(seed, n):
    """Rays along +x just above the axis of the 36-level tree: they pass
    through the boxes of every triangle, so their stacks fill up."""
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-2, 0.5, n), rng.uniform(0.0005, 0.02, n), rng.uniform(0.0005, 0.02, n)], 1)
    d = np.stack([np.ones(n), rng.normal(size=n) * 0.002, rng.normal(size=n) * 0.002], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ori4 = np.zeros((n, 4), np.float32)
    dir4 = np.zeros((n, 4), np.float32)
    ori4[:, :3] = o
    ori4[:, 3] = 0.0001
    dir4[:, :3] = d
    dir4[:, 3] = 1e36
    return ori4, dir4


def _rays(seed, n, pos):
    rng = np.random.default_rng(seed)
    lo, hi = pos.min(0), pos.max(0)
    c, ext = (lo + hi) / 2, (hi - lo) / 2 + 1.0
    tgt = pos[rng.integers(0, len(pos), n)] + rng.normal(size=(n, 3)).astype(np.float32) * 0.1 * ext
    o = c + rng.uniform(-1.5, 1.5, (n, 3)) * ext
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ori4 = np.zeros((n, 4), np.float32)
    dir4 = np.zeros((n, 4), np.float32)
    ori4[:, :3] = o
    ori4[:, 3] = 0.0001
    dir4[:, :3] = d
    dir4[:, 3] = 1e36
    return ori4, dir4


@pytest.mark.parametrize("case", ["soup_5000", "negative", "duplicates", "flat", "grid", "deep"])
def test_sah_tree_traversal(ctx, case):
    import rtx
    from rtx import make_lights, LIGHT_POINT, LIGHT_DIR
    z = np.load(GOLD / f"{case}.npz")
    pos, idx = z["pos"], z["idx"]
    T = len(idx)
    rng = np.random.default_rng(5)
    nrm = rng.normal(size=pos.shape).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    uv = rng.uniform(0, 1, (len(pos), 2)).astype(np.float32)
    mm = rng.integers(0, 3, T).astype(np.int32)
    c = (pos.min(0) + pos.max(0)) / 2
    lights = make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=tuple(c + (pos.max(0) - c) * 1.5)),
                                           dict(type=LIGHT_DIR, dir=(0.3, -1.0, 0.2))])
    s = load_scene(ctx, pos, idx, nrm, uv, mm, lights=lights)
    s.build_sah()
    tree = O.Tree(pos, idx, mm, z["box"], z["words"], z["order"])
    g = s.nodes()
    assert len(g) == tree.node_count
    assert np.array_equal(g["lo"][:, :3], tree.nodes["lo"][:, :3]) and \
        np.array_equal(g["hi"][:, :3], tree.nodes["hi"][:, :3]), "node boxes differ from the reference builder's"
    assert np.array_equal(g["a"], tree.nodes["a"]) and np.array_equal(g["b"], tree.nodes["b"])
    assert np.array_equal(g["parent"][1:], tree.nodes["parent"][1:])
    assert np.array_equal(s.indices().reshape(-1), tree.idx) and np.array_equal(s.matmap(), tree.matmap)
    n = 4096
    ori4, dir4 = axis_rays(7, n) if case == "deep" else _rays(31, n, pos)
    o = O.trace(tree, nrm, uv, ori4, dir4)
    if case == "deep":
        bid, _ = O.trace_brute(tree, ori4, dir4)
        assert ((o["id"] >= 0) != (bid >= 0)).sum() > 100, "the MAX_LEVELS rule should end some rays"
    o["flags"] = O.trace_shadow(tree, O.lights_from(lights), o["flags"], o["p4"])
    assert (o["id"] >= 0).sum() > n // 20
    r = rtx.Rays(ctx, n)
    r.write(ori4, dir4)
    hb = rtx.Hits(ctx, n)
    old = ctx.tuned("trace_variant")
    try:
        for variant in (0, 1, 2, 3):
            ctx.tune("trace_variant", variant)
            rtx.trace(s, r, hb, n)
            rtx.trace_shadow(s, r, hb, n)
            got = hb.read(n)
            for k in ("flags", "id", "t", "bary", "n4", "p4", "uv2"):
                assert np.array_equal(got[k], o[k]), f"{case}: {k} differs (variant {variant})"
    finally:
        ctx.tune("trace_variant", old)
    # the stackless walk (form 4) against the oracle's stackless walk: on the
    # 36-level tree it does not end rays at MAX_LEVELS
    with O.stackless():
        ons = O.trace(tree, nrm, uv, ori4, dir4)
        ons["flags"] = O.trace_shadow(tree, O.lights_from(lights), ons["flags"], ons["p4"])
    if case == "deep":
        assert ((ons["id"] >= 0) & (o["id"] < 0)).sum() > 100
    try:
        ctx.tune("trace_variant", 4)
        rtx.trace(s, r, hb, n)
        rtx.trace_shadow(s, r, hb, n)
        got = hb.read(n)
        for k in ("flags", "id", "t", "bary", "n4", "p4", "uv2"):
            assert np.array_equal(got[k], ons[k]), f"{case}: {k} differs (stackless)"
    finally:
        ctx.tune("trace_variant", old)
    with pytest.raises(rtx.RtxError, match="SAH"):
        s.refit()
    r.destroy()
    hb.destroy()
    s.destroy()


def test_sah_frame(ctx):
    import rtx
    from rtx import make_materials, make_lights, LIGHT_POINT
    pos, nrm, uv, idx, mm = terrain_scene()
    z = np.load(GOLD / "terrain.npz")
    assert np.array_equal(z["pos"], pos) and np.array_equal(z["idx"], idx)
    mats = make_materials([(0.6, 0.5, 0.4, 0.3, 0.0, 0.0), (1.0, 1.0, 1.0, 1.0, 0.8, 1.5)])
    lights = make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(3.0, 6.0, -4.0))])
    s = load_scene(ctx, pos, idx, nrm, uv, mm, mats=mats, lights=lights)
    s.build_sah()
    w, h = 160, 120
    cam = rtx.camera_set((0.0, 3.0, -6.0), (0.0, 0.8, 0.0), (0.0, 1.0, 0.0), np.pi / 4, w / h)
    R = rtx.Renderer(ctx, w, h, 3)
    st = R.frame(s, cam, rebuild=False)
    tree = O.Tree(pos, idx, mm, z["box"], z["words"], z["order"])
    oacc, ost, orays = O.render(tree, nrm, uv, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3)
    assert st.primary_rays + st.secondary_rays == orays and st.secondary_rays > 0
    assert np.array_equal(R.fb.accum(), oacc)
    # the stackless walk renders the oracle's stackless frame
    old = ctx.tuned("trace_variant")
    try:
        ctx.tune("trace_variant", 4)
        R.frame(s, cam, rebuild=False)
        with O.stackless():
            oacc_ns, _, _ = O.render(tree, nrm, uv, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3)
        assert np.array_equal(R.fb.accum(), oacc_ns)
    finally:
        ctx.tune("trace_variant", old)
    # a later LBVH build replaces the SAH tree
    s.build(30, 1)
    b = O.Build(pos, idx[z["order"]], mm[z["order"]], 30, 1)
    R.frame(s, cam, rebuild=False)
    oacc2, _, _ = O.render(b, nrm, uv, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3)
    assert np.array_equal(R.fb.accum(), oacc2)
    R.destroy()
    s.destroy()
