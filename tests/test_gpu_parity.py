"""GPU parity: every stage of the HIP path against the rt-seq oracle, through
the C ABI (librtx.so), on seeded inputs.

Bar (per stage):
  * Morton codes, sorted permutation, permuted index buffer / material map,
    node topology (canonical level order), node AABBs, global AABB: bit-exact.
  * primary rays, hits (id, t, barycentrics, normal, hit point, uv), shadow
    flags, secondary rays, framebuffer accumulators and RGBA8: bit-exact
    (both sides evaluate the same IEEE-754 single-precision sequence).  The
    looser contract (hit ids on >= 99.99% of rays, t within 1e-5 relative,
    pixels within 1/255 on >= 99.9%) is asserted as well, so a future
    reassociated kernel is judged by it.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ID_AGREE = 0.9999
T_REL = 1e-5
PIX_AGREE = 0.999


def _rtx():
    import rtx
    return rtx


def random_scene(seed, T, V=None, extent=10.0, flat_axis=None):
    rng = np.random.default_rng(seed)
    V = V or max(3, T)
    pos = rng.uniform(-extent, extent, (V, 3)).astype(np.float32) + np.float32(0.0)
    if flat_axis is not None:
        pos[:, flat_axis] = 1.25
    idx = rng.integers(0, V, (T, 3)).astype(np.uint32)
    nrm = rng.normal(size=(V, 3)).astype(np.float32)
    uv = rng.uniform(0, 1, (V, 2)).astype(np.float32)
    mm = rng.integers(0, 3, T).astype(np.int32)
    return pos, idx, nrm, uv, mm


def load_scene(ctx, pos, idx, nrm=None, uv=None, mm=None, mats=None, lights=None):
    rtx = _rtx()
    s = rtx.Scene(ctx)
    s.set_geometry(pos, idx, nrm, uv, mm)
    if mats is not None:
        s.set_materials(mats)
    if lights is not None:
        s.set_lights(lights)
    return s


def check_build(s, b, T):
    codes = s.codes()
    assert np.array_equal(codes.astype(np.uint64), b.codes), "sorted Morton codes differ"
    assert np.array_equal(s.perm(), b.perm), "permutation differs from the stable (code, index) order"
    assert np.array_equal(s.indices().reshape(-1), b.idx), "permuted index buffer differs"
    assert np.array_equal(s.matmap(), b.matmap), "permuted material map differs"
    n, lc = s.info()
    assert n == b.node_count, f"node count {n} != {b.node_count}"
    assert np.array_equal(lc[: b.depth], b.level_counts[: b.depth]), "per-level counts differ"
    g = s.nodes()
    o = b.nodes
    assert np.array_equal(g["a"], o["a"]) and np.array_equal(g["b"], o["b"]), "topology differs"
    assert np.array_equal(g["parent"][1:], o["parent"][1:]), "parents differ"
    assert np.array_equal(g["leaf"] != 0, o["leaf"] != 0), "leaf flags differ"
    assert np.array_equal(g["lo"][:, :3], o["lo"][:, :3]) and np.array_equal(g["hi"][:, :3], o["hi"][:, :3]), \
        "node AABBs differ"
    lo, hi = s.global_bbox()
    assert np.array_equal(lo, b.gbox[:3]) and np.array_equal(hi, b.gbox[3:]), "global AABB differs"


@pytest.mark.parametrize("T,seed", [(1, 1), (2, 2), (3, 3), (4, 4), (7, 5), (255, 6), (256, 7), (1000, 8),
                                    (4096, 9), (4097, 10), (50_000, 11)])
def test_lbvh_random(ctx, T, seed):
    pos, idx, nrm, uv, mm = random_scene(seed, T)
    s = load_scene(ctx, pos, idx, nrm, uv, mm)
    s.build(30, 1)
    b = O.Build(pos, idx, mm, 30, 1)
    check_build(s, b, T)
    s.destroy()


@pytest.mark.parametrize("depth,min_leaf", [(30, 1), (30, 4), (24, 1), (32, 2), (9, 1), (8, 1), (10, 2), (11, 1),
                                            (2, 1), (1, 1), (12, 8)])
def test_lbvh_params(ctx, depth, min_leaf):
    pos, idx, nrm, uv, mm = random_scene(100 + depth, 20_000)
    s = load_scene(ctx, pos, idx, nrm, uv, mm)
    s.build(depth, min_leaf)
    b = O.Build(pos, idx, mm, depth, min_leaf)
    check_build(s, b, idx.shape[0])
    s.destroy()


@pytest.mark.parametrize("bits,items", [(4, 16), (4, 32), (8, 32), (8, 16)])
@pytest.mark.parametrize("T,depth,seed", [(50_000, 30, 31), (20_000, 32, 32), (9000, 9, 33), (5, 30, 34),
                                          (70_000, 7, 35)])
def test_lbvh_sort_settings(ctx, bits, items, T, depth, seed):
    # every digit width and tile size of the radix sort gives the same stable
    # (code, index) order, hence the same tree; odd and even pass counts
    pos, idx, nrm, uv, mm = random_scene(seed, T)
    b = O.Build(pos, idx, mm, depth, 1)
    old = (ctx.tuned("sort_bits"), ctx.tuned("sort_items"))
    try:
        ctx.tune("sort_bits", bits)
        ctx.tune("sort_items", items)
        s = load_scene(ctx, pos, idx, nrm, uv, mm)
        s.build(depth, 1)
        check_build(s, b, T)
        pos2 = (pos * np.float32(0.9) + np.float32(0.5)).astype(np.float32)
        s.update_positions(pos2)  # a rebuild from the permuted buffers
        s.build(depth, 1)
        check_build(s, O.Build(pos2, b.idx.reshape(-1, 3), b.matmap, depth, 1), T)
        s.destroy()
    finally:
        ctx.tune("sort_bits", old[0])
        ctx.tune("sort_items", old[1])


@pytest.mark.parametrize("grid", [1, 2, 16])
def test_lbvh_build_grid(ctx, grid):
    # the hierarchy and refit levels' persistent blocks per CU change only
    # how many tiles are in flight, never the tree
    pos, idx, nrm, uv, mm = random_scene(50 + grid, 120_000)
    b = O.Build(pos, idx, mm, 30, 1)
    old = ctx.tuned("build_grid")
    try:
        ctx.tune("build_grid", grid)
        s = load_scene(ctx, pos, idx, nrm, uv, mm)
        s.build(30, 1)
        check_build(s, b, idx.shape[0])
        s.destroy()
    finally:
        ctx.tune("build_grid", old)


def test_lbvh_16m_triangles(ctx):
    # the sort-scaling size (2^24 triangles, a random soup over 2^23 vertices):
    # codes, permutation, topology and AABBs bit-exact against the oracle
    T = 1 << 24
    pos, idx, nrm, uv, mm = random_scene(24, T, V=T // 2)
    s = load_scene(ctx, pos, idx, nrm, uv, mm)
    s.build(30, 1)
    b = O.Build(pos, idx, mm, 30, 1)
    check_build(s, b, T)
    s.destroy()


def test_lbvh_64m_triangles_properties(ctx):
    # the largest sort-scaling size (2^26 triangles) through properties that do
    # not depend on the size: codes nondecreasing, the permutation a bijection
    # that keeps equal codes in index order (the stable order), the permuted
    # index buffer and material map equal to the inputs gathered through it,
    # and the root's box the global box of the vertices used
    T = 1 << 26
    rng = np.random.default_rng(26)
    V = T // 2
    pos = rng.uniform(-10.0, 10.0, (V, 3)).astype(np.float32)
    idx = rng.integers(0, V, (T, 3)).astype(np.uint32)
    mm = rng.integers(0, 3, T).astype(np.int32)
    s = load_scene(ctx, pos, idx, None, None, mm)
    s.build(30, 1)
    codes = s.codes().astype(np.int64)
    perm = s.perm().astype(np.int64)
    assert codes.size == T and perm.size == T
    assert np.all(codes[1:] >= codes[:-1]), "codes not sorted"
    seen = np.zeros(T, bool)
    seen[perm] = True
    assert seen.all(), "the permutation is not a bijection"
    tie = codes[1:] == codes[:-1]
    assert np.all(perm[1:][tie] > perm[:-1][tie]), "equal codes out of index order"
    assert np.array_equal(s.indices().reshape(-1, 3), idx[perm]), "permuted index buffer"
    assert np.array_equal(s.matmap(), mm[perm]), "permuted material map"
    used = np.zeros(V, bool)
    used[idx.reshape(-1)] = True
    lo, hi = s.global_bbox()
    assert np.array_equal(lo, pos[used].min(0)) and np.array_equal(hi, pos[used].max(0)), "global AABB"
    n, _ = s.info()
    assert 0 < n < 2 * T, f"node count {n}"
    s.destroy()


def test_lbvh_duplicates_and_flat_axis(ctx):
    # many identical triangles (equal codes never split) and a flat global axis
    pos, idx, nrm, uv, mm = random_scene(42, 3000, V=50, flat_axis=1)
    idx[::3] = idx[0]
    s = load_scene(ctx, pos, idx, nrm, uv, mm)
    s.build(30, 1)
    b = O.Build(pos, idx, mm, 30, 1)
    check_build(s, b, idx.shape[0])
    s.destroy()


@pytest.fixture(scope="module")
def c1(ctx):
    rtx = _rtx()
    from rtx import scenes
    sd = scenes.config_c1()
    s = scenes.load(rtx.Scene(ctx), sd)
    s.build(30, 1)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    cam = scenes.camera_of(sd)
    yield sd, s, b, cam
    s.destroy()


def test_lbvh_c1(c1):
    sd, s, b, cam = c1
    check_build(s, b, sd.T)


def test_rebuild_chains_permutation(ctx):
    # a second build starts from the permuted buffers (src/rt/bvh-builder.cpp:481-511)
    pos, idx, nrm, uv, mm = random_scene(77, 5000)
    s = load_scene(ctx, pos, idx, nrm, uv, mm)
    s.build(30, 1)
    b1 = O.Build(pos, idx, mm, 30, 1)
    pos2 = (pos * np.float32(0.9) + np.float32(0.5)).astype(np.float32)
    s.update_positions(pos2)
    s.build(30, 1)
    b2 = O.Build(pos2, b1.idx.reshape(-1, 3), b1.matmap, 30, 1)
    check_build(s, b2, 5000)
    assert np.array_equal(s.orig_ids(), b1.perm[b2.perm])
    s.destroy()


def test_refit(ctx):
    pos, idx, nrm, uv, mm = random_scene(78, 8000)
    s = load_scene(ctx, pos, idx, nrm, uv, mm)
    s.build(30, 1)
    b = O.Build(pos, idx, mm, 30, 1)
    pos2 = (pos + np.float32(0.25) * np.sin(pos)).astype(np.float32)
    s.update_positions(pos2)
    s.refit()
    b.refit(pos2)
    g = s.nodes()
    assert np.array_equal(g["a"], b.nodes["a"]) and np.array_equal(g["b"], b.nodes["b"])
    assert np.array_equal(g["lo"][:, :3], b.nodes["lo"][:, :3]) and np.array_equal(g["hi"][:, :3],
                                                                                      b.nodes["hi"][:, :3])
    s.destroy()


@pytest.mark.parametrize("w,h,quad,offset,count", [(512, 512, 32, 0, None), (100, 70, 32, 0, None),
                                                   (97, 131, 16, 1000, 5000), (33, 33, 32, 0, None)])
def test_raygen_primary(ctx, c1, w, h, quad, offset, count):
    rtx = _rtx()
    sd, s, b, cam = c1
    n = (w * h - offset) if count is None else count
    r = rtx.Rays(ctx, n)
    r.primary(cam, w, h, quad, 1, None, offset, n)
    go, gd, gp, gc = r.read(n)
    oo, od, op, oc = O.raygen_primary(O.camera_from(cam), w, h, quad, None, offset, n)
    assert r.count() == n
    assert np.array_equal(gp, op), "pixel mapping differs"
    assert np.array_equal(go, oo) and np.array_equal(gd, od), "ray origin/direction differ"
    assert np.array_equal(gc, oc)
    if offset == 0 and count is None:
        assert np.array_equal(np.sort(gp), np.arange(w * h)), "pixels not covered exactly once"
    r.destroy()


def hits_equal(g, o, n):
    flags_ok = np.array_equal(g["flags"] & 0xff, o["flags"] & 0xff)
    same_id = g["id"] == o["id"]
    agree = same_id.mean() if n else 1.0
    both = (g["id"] >= 0) & (o["id"] >= 0) & same_id
    rel = np.abs(g["t"][both] - o["t"][both]) / np.maximum(np.abs(o["t"][both]), 1e-30)
    return flags_ok, agree, (rel.max() if rel.size else 0.0)


def test_trace_c1_primary(ctx, c1):
    rtx = _rtx()
    sd, s, b, cam = c1
    w, h = sd.width, sd.height
    n = w * h
    r = rtx.Rays(ctx, n)
    hb = rtx.Hits(ctx, n)
    r.primary(cam, w, h)
    rtx.trace(s, r, hb, n)
    g = hb.read(n)
    oo, od, _, _ = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, sd.nrm, sd.uv, oo, od)
    flags_ok, agree, trel = hits_equal(g, o, n)
    assert agree >= ID_AGREE and trel <= T_REL
    assert flags_ok
    assert np.array_equal(g["id"], o["id"]), "hit ids differ"
    for k in ("t", "bary", "n4", "p4", "uv2"):
        assert np.array_equal(g[k], o[k]), f"hit field {k} differs"
    worst, _ = O.hit_t_f64(b, oo, od, g["id"], g["t"])
    assert worst < 1e-4, f"fp64 hit-distance error {worst}"
    r.destroy()
    hb.destroy()


def test_shadow_c1(ctx, c1):
    rtx = _rtx()
    sd, s, b, cam = c1
    w, h = sd.width, sd.height
    n = w * h
    r = rtx.Rays(ctx, n)
    hb = rtx.Hits(ctx, n)
    r.primary(cam, w, h)
    rtx.trace(s, r, hb, n)
    rtx.trace_shadow(s, r, hb, n)
    g = hb.read(n)
    oo, od, _, _ = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, sd.nrm, sd.uv, oo, od)
    of = O.trace_shadow(b, O.lights_from(sd.lights), o["flags"], o["p4"])
    assert np.array_equal(g["flags"], of), "shadow flags differ"
    r.destroy()
    hb.destroy()


def test_secondary_and_shade_c1(ctx, c1):
    rtx = _rtx()
    sd, s, b, cam = c1
    w, h = sd.width, sd.height
    n = w * h
    cap = 2 * n
    ra, rb = rtx.Rays(ctx, cap), rtx.Rays(ctx, cap)
    hb = rtx.Hits(ctx, cap)
    fb = rtx.FrameBuffer(ctx, w, h)
    ra.primary(cam, w, h)
    rtx.trace(s, ra, hb, n)
    rtx.trace_shadow(s, ra, hb, n)
    rtx.shade(s, ra, hb, fb, n, True)
    m = rtx.raygen_secondary(s, ra, hb, rb, n)
    lights = O.lights_from(sd.lights)
    oo, od, op, oc = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, sd.nrm, sd.uv, oo, od)
    o["flags"] = O.trace_shadow(b, lights, o["flags"], o["p4"])
    acc = np.zeros((n, 3), np.uint32)
    O.shade(sd.materials, b.matmap, lights, od, op, oc, o, acc)
    assert np.array_equal(fb.accum(), acc), "primary shading differs"
    so, sdir, sp, sc, total = O.raygen_secondary(sd.materials, b.matmap, oo, od, op, oc, o, cap)
    assert m == so.shape[0] == total
    go, gd, gp, gc = rb.read(m)
    assert np.array_equal(gp, sp) and np.array_equal(gc, sc), "secondary pixel/contribution differ"
    assert np.array_equal(go, so) and np.array_equal(gd, sdir), "secondary rays differ"
    fb.resolve()
    assert np.array_equal(fb.rgba8(), O.resolve(acc))
    for x in (ra, rb, hb, fb):
        x.destroy()


def test_secondary_capacity_drop(ctx, c1):
    # a parent whose children would reach the capacity emits none
    rtx = _rtx()
    sd, s, b, cam = c1
    w, h = 128, 128
    n = w * h
    cap = 1000
    ra, rb = rtx.Rays(ctx, n), rtx.Rays(ctx, cap)
    hb = rtx.Hits(ctx, n)
    ra.primary(cam, w, h)
    rtx.trace(s, ra, hb, n)
    m = rtx.raygen_secondary(s, ra, hb, rb, n)
    oo, od, op, oc = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, sd.nrm, sd.uv, oo, od)
    so, sdir, sp, sc, total = O.raygen_secondary(sd.materials, b.matmap, oo, od, op, oc, o, cap)
    assert m == so.shape[0] and m <= cap
    go, gd, gp, gc = rb.read(m)
    assert np.array_equal(go, so) and np.array_equal(gd, sdir) and np.array_equal(gp, sp)
    for x in (ra, rb, hb):
        x.destroy()


def test_frame_c1(ctx, c1):
    rtx = _rtx()
    sd, s, b, cam = c1
    R = rtx.Renderer(ctx, sd.width, sd.height, sd.max_bounces)
    st = R.frame(s, cam, rebuild=False)
    acc = R.fb.accum()
    rgba = R.fb.rgba8()
    oacc, ost, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam),
                                sd.width, sd.height, sd.max_bounces)
    assert st.primary_rays + st.secondary_rays == orays
    assert st.shadow_rays == ost["shadow"]
    assert np.array_equal(acc, oacc), "framebuffer accumulators differ"
    ref = O.resolve(oacc)
    close = (np.abs(rgba.astype(int) - ref.astype(int)) <= 1).all(axis=1).mean()
    assert close >= PIX_AGREE
    assert np.array_equal(rgba, ref)
    # read-back-free frame gives the same image
    R.frame_async(s, cam, rebuild=False)
    ctx.sync()
    st2 = R.last_stats()
    assert np.array_equal(R.fb.accum(), oacc)
    assert st2.secondary_rays == st.secondary_rays and st2.shadow_rays == st.shadow_rays
    R.destroy()


def test_frame_rebuild_c1(ctx, c1):
    # full frame including the LBVH rebuild from the permuted buffers
    rtx = _rtx()
    from rtx import scenes
    sd = scenes.config_c1()
    s = scenes.load(rtx.Scene(ctx), sd)
    cam = scenes.camera_of(sd)
    R = rtx.Renderer(ctx, 256, 256, 1)
    R.frame(s, cam, rebuild=True)
    R.frame(s, cam, rebuild=True)  # second build starts from the permuted order
    acc = R.fb.accum()
    b1 = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    b2 = O.Build(sd.pos, b1.idx.reshape(-1, 3), b1.matmap, 30, 1)
    oacc, _, _ = O.render(b2, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), 256, 256, 1)
    assert np.array_equal(acc, oacc)
    R.destroy()
    s.destroy()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("block", [64, 128, 256])
def test_traversal_variants_identical(ctx, c1, variant, block):
    # every tuned traversal form reproduces the oracle's hits and shadow flags
    rtx = _rtx()
    sd, s, b, cam = c1
    w, h = 200, 150
    n = w * h
    r = rtx.Rays(ctx, 2 * n)
    r2 = rtx.Rays(ctx, 2 * n)
    hb = rtx.Hits(ctx, 2 * n)
    old = (ctx.tuned("trace_variant"), ctx.tuned("trace_block"))
    ctx.tune("trace_variant", variant)
    ctx.tune("trace_block", block)
    try:
        r.primary(cam, w, h)
        rtx.trace(s, r, hb, n)
        rtx.trace_shadow(s, r, hb, n)
        g = hb.read(n)
        m = rtx.raygen_secondary(s, r, hb, r2, n)
        rtx.trace(s, r2, hb, m)
        rtx.trace_shadow(s, r2, hb, m)
        g2 = hb.read(m)
    finally:
        ctx.tune("trace_variant", old[0])
        ctx.tune("trace_block", old[1])
    lights = O.lights_from(sd.lights)
    oo, od, op, oc = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, sd.nrm, sd.uv, oo, od)
    o["flags"] = O.trace_shadow(b, lights, o["flags"], o["p4"])
    for k in ("flags", "id", "t", "bary", "n4", "p4"):
        assert np.array_equal(g[k], o[k]), f"primary {k} differs (variant {variant}, block {block})"
    so, sdir, sp, sc, _ = O.raygen_secondary(sd.materials, b.matmap, oo, od, op, oc, o, 2 * n)
    o2 = O.trace(b, sd.nrm, sd.uv, so, sdir)
    o2["flags"] = O.trace_shadow(b, lights, o2["flags"], o2["p4"])
    for k in ("flags", "id", "t", "bary", "n4", "p4"):
        assert np.array_equal(g2[k], o2[k]), f"secondary {k} differs (variant {variant}, block {block})"
    for x in (r, r2, hb):
        x.destroy()


def test_graph_replay_matches_oracle(ctx, c1):
    # the captured frame (LBVH rebuild + render) replayed several times, with
    # eager work and synchronizations in between; every replay must match
    rtx = _rtx()
    from rtx import scenes
    sd = scenes.config_c1()
    s = scenes.load(rtx.Scene(ctx), sd)
    cam = scenes.camera_of(sd)
    w, h = 192, 144
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], w / h)
    R = rtx.Renderer(ctx, w, h, 2)
    R.frame_async(s, cam, rebuild=True)   # eager frame: allocates everything
    ctx.sync()
    R.capture(s, cam, rebuild=True)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    for k in range(3):
        R.replay()
        ctx.sync()
        b = O.Build(sd.pos, b.idx.reshape(-1, 3), b.matmap, 30, 1)  # each replay rebuilds from the permuted order
        oacc, st, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam),
                                   w, h, 2)
        assert np.array_equal(R.fb.accum(), oacc), f"replay {k} differs"
        gst = R.last_stats()
        assert gst.primary_rays + gst.secondary_rays == orays and gst.shadow_rays == st["shadow"]
        R.frame(s, cam, rebuild=False)      # eager work between replays
        b2 = b
        assert np.array_equal(R.fb.accum(), O.render(b2, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights),
                                                     O.camera_from(cam), w, h, 2)[0])
    R.destroy()
    s.destroy()


@pytest.mark.parametrize("variant,refill,side", [(3, 1, 1), (3, 16, 1), (3, 64, 1), (1, 16, 1), (1, 16, 0)])
def test_frame_and_graph_per_variant(ctx, variant, refill, side):
    # whole frames (async and captured) under each traversal form, including
    # the persistent ray-queue form at its extreme refill thresholds, and
    # with every stage on one stream (side = 0: linear graphs)
    rtx = _rtx()
    from rtx import scenes
    sd = scenes.config_c1()
    s = scenes.load(rtx.Scene(ctx), sd)
    w, h = 160, 120
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], w / h)
    old = (ctx.tuned("trace_variant"), ctx.tuned("refill"), ctx.tuned("side_stream"))
    ctx.tune("trace_variant", variant)
    ctx.tune("refill", refill)
    ctx.tune("side_stream", side)
    try:
        R = rtx.Renderer(ctx, w, h, 2)
        R.frame_async(s, cam, rebuild=True)
        ctx.sync()
        b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
        oacc, st, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam),
                                   w, h, 2)
        assert np.array_equal(R.fb.accum(), oacc), "async frame differs"
        gst = R.last_stats()
        assert gst.primary_rays + gst.secondary_rays == orays and gst.shadow_rays == st["shadow"]
        R.capture(s, cam, rebuild=True)
        for k in range(2):
            R.replay()
            ctx.sync()
            b = O.Build(sd.pos, b.idx.reshape(-1, 3), b.matmap, 30, 1)
            oacc = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h,
                            2)[0]
            assert np.array_equal(R.fb.accum(), oacc), f"replay {k} differs"
        R.destroy()
    finally:
        ctx.tune("trace_variant", old[0])
        ctx.tune("refill", old[1])
        ctx.tune("side_stream", old[2])
        s.destroy()
