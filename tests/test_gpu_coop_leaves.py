"""GPU parity of the wave-cooperative leaf tests.

Leaves holding long runs of equal Morton codes are tested by the whole wave
together once they reach the `coop_leaf` size.  The result must be the
sequential in-order loop's: the smallest t, and on an exact tie the later
triangle.  The scene has
  * a tiny, finely tessellated sphere (4096 triangles in a few Morton
    cells), whose leaves hold hundreds of triangles;
  * a sphere whose every triangle appears twice (same vertex indices), so
    every hit on it is an exact tie between two triangles of one leaf (equal
    centroids give equal codes, and equal codes are never split apart);
  * a refractive sphere and a floor, for secondary and shadow rays.
Every threshold (0 = never, 1 = every leaf) and the single-node traversal
forms (plain and persistent) must give the oracle's hits, shadow flags and
secondary rays bit for bit.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rtx():
    import rtx
    return rtx


def pole_scene():
    from rtx import scenes
    from rtx import make_lights, make_materials, LIGHT_POINT, LIGHT_DIR
    spheres = [  # slices, stacks, centre, radius, duplicate every triangle
        (96, 48, (0.0, 0.0, 0.0), 1.0, True),
        (128, 64, (2.5, 0.0, 0.0), 0.75, False),
        (64, 32, (0.0, 1.3, 0.0), 0.05, False),
    ]
    P, N, U, I, MM = [], [], [], [], []
    base = 0
    for k, (slices, stacks, c, r, dup) in enumerate(spheres):
        p, n, uv, _ = scenes.uv_sphere(slices, stacks, c, r)
        idx = scenes.grid_indices(stacks + 1, slices + 1, base)
        if dup:
            idx = np.repeat(idx, 2, axis=0)
        P.append(p), N.append(n), U.append(uv), I.append(idx)
        MM.append(np.full(idx.shape[0], k, np.int32))
        base += p.shape[0]
    # floor (two triangles) below the spheres
    fl = np.array([[-6, -1.5, -6], [6, -1.5, -6], [6, -1.5, 6], [-6, -1.5, 6]], np.float64)
    P.append(fl), N.append(np.tile([0.0, 1.0, 0.0], (4, 1))), U.append(np.zeros((4, 2)))
    I.append(np.array([[base, base + 2, base + 1], [base, base + 3, base + 2]], np.uint32))
    MM.append(np.full(2, 3, np.int32))
    f32 = lambda a: (np.asarray(a, np.float64) + 0.0).astype(np.float32)
    pos, nrm, uv = f32(np.concatenate(P)), f32(np.concatenate(N)), f32(np.concatenate(U))
    idx, mm = np.concatenate(I).astype(np.uint32), np.concatenate(MM)
    mats = make_materials([(0.9, 0.9, 0.9, 1.0, 0.6, 0.0), (0.8, 0.9, 1.0, 1.0, 0.9, 1.5),
                           (1.0, 0.3, 0.3, 1.0, 0.5, 0.0), (0.5, 0.5, 0.5, 0.5, 0.0, 0.0)])
    lights = make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(0.0, 10.0, 0.0), color=(1.0, 1.0, 1.0)),
                                           dict(type=LIGHT_DIR, dir=(0.3, -1.0, 0.2), color=(0.5, 0.5, 0.5))])
    return pos, nrm, uv, idx, mm, mats, lights


@pytest.fixture(scope="module")
def poles(ctx):
    rtx = _rtx()
    pos, nrm, uv, idx, mm, mats, lights = pole_scene()
    s = rtx.Scene(ctx)
    s.set_geometry(pos, idx, nrm, uv, mm)
    s.set_materials(mats)
    s.set_lights(lights)
    s.build(30, 1)
    b = O.Build(pos, idx, mm, 30, 1)
    yield dict(s=s, b=b, nrm=nrm, uv=uv, mats=mats, lights=lights)
    s.destroy()


def test_scene_has_fat_leaves_and_ties(poles):
    import rtx
    b = poles["b"]
    n = b.nodes
    leaf = n["leaf"] != 0
    cnt = (n["b"] - n["a"])[leaf]
    assert cnt.max() >= 64, "the tiny sphere should form leaves of many triangles"
    # the oracle resolves every tie on the duplicated sphere to the later
    # triangle of the pair: input 2j + 1 (the pair is adjacent after the
    # stable sort, so the later input is the later sorted position)
    p, at, up, w, h = CAMS["top"]
    cam = rtx.camera_set(p, at, up, np.pi / 4, w / h)
    oo, od, _, _ = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, poles["nrm"], poles["uv"], oo, od)
    ids = o["id"][o["id"] >= 0]
    dup = ids[b.matmap[ids] == 0]
    assert len(dup) > 1000
    assert np.all(b.perm[dup] % 2 == 1), "a tie went to the earlier triangle of a duplicated pair"


CAMS = {
    "top": ((0.0, 5.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 201, 201),
    "tiny": ((0.25, 1.55, -0.3), (0.0, 1.3, 0.0), (0.0, 1.0, 0.0), 160, 120),
    "oblique": ((3.0, 3.5, -4.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 160, 120),
}


@pytest.mark.parametrize("cam_name", list(CAMS))
@pytest.mark.parametrize("coop", [0, 1, 2, 8, 64])
@pytest.mark.parametrize("variant", [0, 1, 3])
def test_coop_leaf_thresholds_identical(ctx, poles, cam_name, coop, variant):
    rtx = _rtx()
    s, b = poles["s"], poles["b"]
    p, at, up, w, h = CAMS[cam_name]
    cam = rtx.camera_set(p, at, up, np.pi / 4, w / h)
    n = w * h
    r, r2 = rtx.Rays(ctx, 2 * n + 2), rtx.Rays(ctx, 2 * n + 2)
    hb = rtx.Hits(ctx, 2 * n + 2)
    old = (ctx.tuned("coop_leaf"), ctx.tuned("trace_variant"))
    ctx.tune("coop_leaf", coop)
    ctx.tune("trace_variant", variant)
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
        ctx.tune("coop_leaf", old[0])
        ctx.tune("trace_variant", old[1])
    lights = O.lights_from(poles["lights"])
    oo, od, op, oc = O.raygen_primary(O.camera_from(cam), w, h)
    o = O.trace(b, poles["nrm"], poles["uv"], oo, od)
    o["flags"] = O.trace_shadow(b, lights, o["flags"], o["p4"])
    assert (o["id"] >= 0).sum() > n // 4, "the view should hit geometry"
    for k in ("flags", "id", "t", "bary", "n4", "p4"):
        assert np.array_equal(g[k], o[k]), f"primary {k} differs (coop {coop}, variant {variant}, {cam_name})"
    so, sdir, sp, sc, _ = O.raygen_secondary(poles["mats"], b.matmap, oo, od, op, oc, o, 2 * n + 2)
    assert m == len(so)
    o2 = O.trace(b, poles["nrm"], poles["uv"], so, sdir)
    o2["flags"] = O.trace_shadow(b, lights, o2["flags"], o2["p4"])
    for k in ("flags", "id", "t", "bary", "n4", "p4"):
        assert np.array_equal(g2[k], o2[k]), f"secondary {k} differs (coop {coop}, variant {variant}, {cam_name})"
    for x in (r, r2, hb):
        x.destroy()
