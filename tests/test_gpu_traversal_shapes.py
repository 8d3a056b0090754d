"""GPU parity of every traversal form on trees of every shape the build makes.

The C1 tests trace one scene; here the trees are one leaf (T <= 3), trees
of 1, 2, 9, 12, 30 and 32 levels with leaves at odd and even depths, larger
leaves (min_leaf 4 and 8), duplicated triangles (runs of equal codes: fat
leaves, tested by the whole wave) and a flat axis.  Random rays from inside
and around the scene give long, incoherent traversals.  Hits and shadow
flags of each form must equal the oracle's bit for bit, with the whole
traversal stack in LDS and with ring stacks of 16 and 8 LDS entries (the
deeper levels in private memory), with shadow rays in the reference's child
order and in index order.
"""
import numpy as np
import pytest

import oracle as O
from chain import chain_scene
from test_gpu_parity import load_scene, random_scene

pytestmark = pytest.mark.gpu


def _rays(seed, n, extent=12.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-extent, extent, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ori4 = np.zeros((n, 4), np.float32)
    dir4 = np.zeros((n, 4), np.float32)
    ori4[:, :3] = o
    ori4[:, 3] = 0.0001
    dir4[:, :3] = d.astype(np.float32)
    dir4[:, 3] = 1e36
    return ori4, dir4


def _lights():
    from rtx import make_lights, LIGHT_POINT, LIGHT_DIR
    return make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(3.0, 25.0, -4.0), color=(1.0, 1.0, 1.0)),
                                         dict(type=LIGHT_DIR, dir=(0.3, -1.0, 0.2), color=(0.5, 0.5, 0.5))])


CASES = [  # T, depth, min_leaf, seed, kind
    (1, 30, 1, 1, "random"), (2, 30, 1, 2, "random"), (3, 30, 1, 3, "random"), (7, 30, 1, 4, "random"),
    (300, 30, 1, 5, "random"), (5000, 30, 1, 6, "random"), (5000, 32, 1, 7, "random"), (5000, 9, 1, 8, "random"),
    (5000, 2, 1, 9, "random"), (5000, 1, 1, 10, "random"), (20000, 30, 4, 11, "random"),
    (20000, 12, 8, 12, "random"), (6000, 30, 1, 13, "duplicates"), (6000, 30, 1, 14, "flat"),
]


@pytest.mark.parametrize("T,depth,min_leaf,seed,kind", CASES)
def test_traversal_forms_match_oracle(ctx, T, depth, min_leaf, seed, kind):
    import rtx
    if kind == "flat":
        pos, idx, nrm, uv, mm = random_scene(seed, T, flat_axis=1)
    else:
        pos, idx, nrm, uv, mm = random_scene(seed, T)
        if kind == "duplicates":  # every triangle three times: runs of equal codes, fat leaves
            idx = np.repeat(idx[: T // 3], 3, axis=0)
            mm = np.repeat(mm[: T // 3], 3)
    lights = _lights()
    s = load_scene(ctx, pos, idx, nrm, uv, mm, lights=lights)
    s.build(depth, min_leaf)
    b = O.Build(pos, idx, mm, depth, min_leaf)
    n = 4096
    ori4, dir4 = _rays(100 + seed, n)
    r = rtx.Rays(ctx, n)
    r.write(ori4, dir4)
    hb = rtx.Hits(ctx, n)
    o = O.trace(b, nrm, uv, ori4, dir4)
    o["flags"] = O.trace_shadow(b, O.lights_from(lights), o["flags"], o["p4"])
    if T >= 300:
        assert (o["id"] >= 0).sum() > n // 8, "the rays should hit the scene"
    old = (ctx.tuned("trace_variant"), ctx.tuned("stack_lds"), ctx.tuned("shadow_order"))
    try:
        for variant, sl, so in ((0, 32, 0), (0, 16, 0), (0, 8, 0), (1, 32, 0), (1, 16, 0), (1, 8, 0), (2, 32, 0),
                                (2, 16, 0), (2, 8, 0), (0, 32, 1), (1, 32, 1), (1, 16, 1)):
            # so: shadow rays in the reference's child order (1) or in index order (0)
            ctx.tune("trace_variant", variant)
            ctx.tune("stack_lds", sl)
            ctx.tune("shadow_order", so)
            rtx.trace(s, r, hb, n)
            rtx.trace_shadow(s, r, hb, n)
            g = hb.read(n)
            for k in ("flags", "id", "t", "bary", "n4", "p4"):
                assert np.array_equal(g[k], o[k]), \
                    f"{k} differs (variant {variant}, stack_lds {sl}, shadow_order {so}, T {T}, depth {depth})"
        # the stackless walk (parent pointers of the LBVH) against the oracle's
        with O.stackless():
            ons = O.trace(b, nrm, uv, ori4, dir4)
            ons["flags"] = O.trace_shadow(b, O.lights_from(lights), ons["flags"], ons["p4"])
        ctx.tune("trace_variant", 4)
        ctx.tune("stack_lds", 32)
        rtx.trace(s, r, hb, n)
        rtx.trace_shadow(s, r, hb, n)
        g = hb.read(n)
        for k in ("flags", "id", "t", "bary", "n4", "p4"):
            assert np.array_equal(g[k], ons[k]), f"{k} differs (stackless, T {T}, depth {depth})"
    finally:
        ctx.tune("trace_variant", old[0])
        ctx.tune("stack_lds", old[1])
        ctx.tune("shadow_order", old[2])
    r.destroy()
    hb.destroy()
    s.destroy()


@pytest.mark.parametrize("variant,stack_lds", [(0, 32), (1, 32), (2, 32), (3, 32), (4, 32), (0, 16), (1, 16),
                                               (1, 8), (2, 8)])
def test_max_levels_rule(ctx, variant, stack_lds):
    # A ray whose near children all lie on the chain has 32 entries on its
    # stack at the bottom leaf; the reference then ends the ray instead of
    # popping (src/kernel/trace-bvh.cl:361-367, shadow-trace-bvh.cl alike).
    # Rays aimed at the slivers with positive directions take that path, so
    # for some of them the reference (and the oracle) report no hit where a
    # brute-force search finds one.  Every traversal form must agree, with
    # every stack size (the ring stacks hold levels 16..31 or 8..31 of these
    # paths in private memory).
    import rtx
    from rtx import make_lights, LIGHT_POINT
    pos, idx = chain_scene()
    T = len(idx)
    b = O.Build(pos, idx, None, 32, 1)
    assert (b.level_counts[:32] > 0).all(), "the chain should have 32 inner levels"
    rng = np.random.default_rng(5)
    cent = pos[idx].mean(axis=1)
    n = 4096
    tgt = cent[4 + rng.integers(0, T - 4, n)]
    d = np.abs(rng.normal(size=(n, 3))) + 0.2
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ori4 = np.zeros((n, 4), np.float32)
    dir4 = np.zeros((n, 4), np.float32)
    ori4[:, :3] = tgt - 2.0 * d
    ori4[:, 3] = 0.0001
    dir4[:, :3] = d
    dir4[:, 3] = 1e36
    nrm = np.tile(np.float32([0.0, 0.0, 1.0]), (len(pos), 1))
    uv = np.zeros((len(pos), 2), np.float32)
    lights = make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(3.0, 3.0, 3.0), color=(1.0, 1.0, 1.0))])
    o = O.trace(b, nrm, uv, ori4, dir4)
    bid, _ = O.trace_brute(b, ori4, dir4)
    assert ((o["id"] >= 0) != (bid >= 0)).sum() > 100, "the rule should decide some rays"
    o["flags"] = O.trace_shadow(b, O.lights_from(lights), o["flags"], o["p4"])
    if variant == 4:  # the stackless walk has no MAX_LEVELS: it finds what the brute force finds
        with O.stackless():
            o = O.trace(b, nrm, uv, ori4, dir4)
            o["flags"] = O.trace_shadow(b, O.lights_from(lights), o["flags"], o["p4"])
        assert np.array_equal(o["id"] >= 0, bid >= 0)
    s = load_scene(ctx, pos, idx, nrm, uv, np.zeros(T, np.int32), lights=lights)
    s.build(32, 1)
    r = rtx.Rays(ctx, n)
    r.write(ori4, dir4)
    hb = rtx.Hits(ctx, n)
    old = (ctx.tuned("trace_variant"), ctx.tuned("stack_lds"))
    try:
        ctx.tune("trace_variant", variant)
        ctx.tune("stack_lds", stack_lds)
        rtx.trace(s, r, hb, n)
        rtx.trace_shadow(s, r, hb, n)
        g = hb.read(n)
    finally:
        ctx.tune("trace_variant", old[0])
        ctx.tune("stack_lds", old[1])
    for k in ("flags", "id", "t", "bary", "n4", "p4"):
        assert np.array_equal(g[k], o[k]), f"{k} differs (variant {variant}, stack_lds {stack_lds})"
    r.destroy()
    hb.destroy()
    s.destroy()
