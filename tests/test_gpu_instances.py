"""GPU parity of the multi-root (instanced) traversal: trace_multi and
shadow_trace_multi (src/kernel/trace-bvh.cl:419-463,
src/kernel/shadow-trace-bvh.cl:317-373) against the oracle, bit for bit, for
every traversal form.

Cases: rigid instances with scales sharing one mesh; two scenes of the same
triangles under the same transform (every hit an exact tie, which the first
root wins); interpenetrating instances; a projective transform (w != 1, the
division branch of multiply_point); new transforms without a rebuild; and a
scene rebuilt under its instances.  Lights: a directional light, a point
light, and spot lights on both sides of that kernel's cone test.
"""
import numpy as np
import pytest

import oracle as O
from inst_scene import rot, xform, two_meshes, sphere_mesh, rays_at
from test_gpu_parity import load_scene

pytestmark = pytest.mark.gpu

FIELDS = ("flags", "t", "id", "bary", "n4", "p4", "uv2")


def _lights():
    from rtx import make_lights, LIGHT_DIR, LIGHT_SPOT, LIGHT_POINT
    return make_lights((0.1, 0.1, 0.1), [
        dict(type=LIGHT_DIR, dir=(0.2, -1.0, 0.1), color=(1.0, 1.0, 1.0)),
        dict(type=LIGHT_POINT, pos=(1.0, 6.0, -5.0), color=(1.0, 1.0, 1.0)),
        # toward this light from the scene is about (0, +1, 0): below the cut-off, lit untraced
        dict(type=LIGHT_SPOT, pos=(0.0, 12.0, 0.0), dir=(0.0, -1.0, 0.0), angle=0.8, color=(1.0, 1.0, 1.0)),
        # toward this light is about (0, -1, 0): above the cut-off, traced
        dict(type=LIGHT_SPOT, pos=(0.5, -12.0, 0.0), dir=(0.0, -1.0, 0.0), angle=0.8, color=(1.0, 1.0, 1.0)),
    ])


def _projective():
    # the inverse has the bottom row (0.01, 0, 0.03, 1): the rays (+z) keep
    # w > 0 in the object frame, both the points and the directions are
    # divided by it, and the instance is still hit
    R = rot((1, 1, 0), 0.5)
    t = np.array((0.5, -1.0, 0.5))
    Mi = np.eye(4)
    Mi[:3, :3] = R.T
    Mi[:3, 3] = -R.T @ t
    Mi[3, :3] = (0.01, 0.0, 0.03)
    return np.linalg.inv(Mi).astype(np.float32), Mi.astype(np.float32)


CASES = {
    "rigid": (two_meshes, [(0, *xform(rot((0, 1, 0), 0.7), 1.0, (-2.0, 0.3, 0.0))),
                           (1, *xform(rot((1, 0, 1), 1.9), 2.0, (2.0, -0.5, 1.0))),
                           (0, *xform(rot((0, 0, 1), -0.4), 0.5, (0.0, 2.0, -1.0)))]),
    "tie": (lambda: [sphere_mesh(), sphere_mesh()], [(1, *xform(rot((0, 1, 0), 0.9), 1.5)),
                                                     (0, *xform(rot((0, 1, 0), 0.9), 1.5))]),
    "overlap": (two_meshes, [(0, *xform(rot((0, 1, 0), 0.3), 1.2, (0.4, 0.0, 0.0))),
                             (1, *xform(rot((1, 0, 0), 0.8), 1.6, (-0.5, 0.2, 0.3))),
                             (0, *xform(rot((0, 0, 1), 1.1), 0.9, (0.0, -0.3, -0.6))),
                             (1, *xform(np.eye(3), 1.0, (0.0, 2.6, 0.0)))]),
    "projective": (two_meshes, [(0, *_projective()), (1, *xform(np.eye(3), 1.0, (0.0, 2.5, 0.0)))]),
}


def _load(ctx, meshes, lights):
    out = []
    for p, n, uv, idx in meshes:
        s = load_scene(ctx, p, idx, n, uv, np.zeros(len(idx), np.int32), lights=lights)
        s.build(30, 1)
        out.append(s)
    return out


def _oracle(meshes, builds, insts, lights, ori4, dir4):
    ms = O.MultiScene(builds, [m[1] for m in meshes], [m[2] for m in meshes], insts)
    o = O.trace_multi(ms, ori4, dir4)
    o["flags"] = O.trace_shadow_multi(ms, O.lights_from(lights), o["flags"], o["p4"])
    return o, ms


def _check_all_forms(ctx, inst, r, hb, n, o, what):
    """Every traversal form; the single-node forms (0, 1 and 3, which maps to
    1) under the three root schedules: each lane on its own roots in one loop,
    over the roots whose world box its ray meets (2) or over all of them (1),
    and the wave-uniform root loop (0); shadow rays of the single-node forms
    in index order (0) and in the reference's child order (1)."""
    old, old_lr, old_coop = ctx.tuned("trace_variant"), ctx.tuned("inst_lane_roots"), ctx.tuned("coop_leaf")
    old_so = ctx.tuned("shadow_order")
    try:
        for variant in (0, 1, 2, 3):
            for lane_roots in ((2, 1, 0) if variant != 2 else (2,)):
                # coop_leaf 1: every leaf tested by the whole wave (with the
                # owner lane's triangle pointer in the lane-roots kernels)
                for coop in ((old_coop, 1) if variant == 1 else (old_coop,)):
                    for so in ((0, 1) if variant in (0, 1) else (0,)):
                        ctx.tune("trace_variant", variant)
                        ctx.tune("inst_lane_roots", lane_roots)
                        ctx.tune("coop_leaf", coop)
                        ctx.tune("shadow_order", so)
                        inst.trace(r, hb, n)
                        inst.trace_shadow(r, hb, n)
                        g = hb.read(n)
                        for k in FIELDS:
                            assert np.array_equal(g[k], o[k]), \
                                f"{what}: {k} differs (variant {variant}, lane roots {lane_roots}, coop leaf " \
                                f"{coop}, shadow order {so})"
    finally:
        ctx.tune("trace_variant", old)
        ctx.tune("inst_lane_roots", old_lr)
        ctx.tune("coop_leaf", old_coop)
        ctx.tune("shadow_order", old_so)


@pytest.mark.parametrize("case", list(CASES))
def test_instances_match_oracle(ctx, case):
    import rtx
    make, insts = CASES[case]
    meshes = make()
    lights = _lights()
    scs = _load(ctx, meshes, lights)
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    n = 6000
    ori4, dir4 = rays_at(11, n)
    o, ms = _oracle(meshes, builds, insts, lights, ori4, dir4)
    hit = (o["flags"] & 1) != 0
    assert hit.sum() > 300, "the rays should hit the instances"
    for i, (m, _, _) in enumerate(insts):
        mine = (o["id"][hit] >= ms.tri_base[m]) & (o["id"][hit] < ms.tri_base[m] + builds[m].T)
        if case == "tie":  # every hit is a tie, won by the first root (mesh 1, laid out first)
            assert mine.all() if i == 0 else not mine.any()
        else:
            assert mine.sum() > 50, f"instance {i}'s mesh should be hit"
    for l in (0, 1, 3):
        assert ((o["flags"] >> (8 + l)) & 1)[hit].any(), f"light {l} should shadow some hits"
    assert not ((o["flags"] >> 10) & 1).any(), "the first spot light lights every point untraced"
    inst = rtx.Instances(ctx, [scs[m] for m, _, _ in insts], [t for _, t, _ in insts], [ti for _, _, ti in insts])
    inst.set_lights(lights)
    base = inst.tri_base()
    assert [int(b) for b in base] == [ms.tri_base[m] for m, _, _ in insts]
    r = rtx.Rays(ctx, n)
    r.write(ori4, dir4)
    hb = rtx.Hits(ctx, n)
    _check_all_forms(ctx, inst, r, hb, n, o, case)
    # the ray count kept on the device
    inst.trace(r, hb)
    inst.trace_shadow(r, hb)
    g = hb.read(n)
    assert all(np.array_equal(g[k], o[k]) for k in FIELDS), f"{case}: device-count trace differs"
    inst.destroy()
    r.destroy()
    hb.destroy()
    for s in scs:
        s.destroy()


def test_new_transforms_and_rebuilt_scene(ctx):
    """Rigid animation ships new roots without a rebuild; a scene rebuilt
    under its instances (moved vertices) is traced with its new tree."""
    import rtx
    meshes = two_meshes()
    lights = _lights()
    scs = _load(ctx, meshes, lights)
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    n = 4096
    ori4, dir4 = rays_at(12, n)
    r = rtx.Rays(ctx, n)
    r.write(ori4, dir4)
    hb = rtx.Hits(ctx, n)
    insts = [(0, *xform(rot((0, 1, 0), 0.2), 1.0, (-1.5, 0.0, 0.0))), (1, *xform(np.eye(3), 1.3, (1.5, 0.0, 0.0))),
             (0, *xform(np.eye(3), 0.6, (0.0, 1.8, 0.0)))]
    inst = rtx.Instances(ctx, [scs[m] for m, _, _ in insts], [t for _, t, _ in insts], [ti for _, _, ti in insts])
    inst.set_lights(lights)
    for step in range(4):
        ang = 0.35 * step
        cur = [(0, *xform(rot((0, 1, 0), 0.2 + ang), 1.0, (-1.5, 0.2 * step, 0.0))),
               (1, *xform(rot((1, 0, 0), ang), 1.3, (1.5, 0.0, 0.1 * step))),
               (0, *xform(rot((0, 0, 1), -ang), 0.6, (0.0, 1.8 - 0.1 * step, 0.0)))]
        inst.set_transforms([t for _, t, _ in cur], [ti for _, _, ti in cur])
        o, _ = _oracle(meshes, builds, cur, lights, ori4, dir4)
        inst.trace(r, hb, n)
        inst.trace_shadow(r, hb, n)
        g = hb.read(n)
        for k in FIELDS:
            assert np.array_equal(g[k], o[k]), f"step {step}: {k} differs"
    # move the vertices of mesh 0 and rebuild its scene: the instances follow
    p0, n0, uv0, idx0 = meshes[0]
    p_new = (p0 * np.float32(1.25) + np.float32([0.0, 0.1, 0.0])).astype(np.float32)
    scs[0].update_positions(p_new)
    scs[0].build(30, 1)
    ctx.sync()
    b0 = O.Build(p_new, builds[0].idx.reshape(-1, 3), None, 30, 1)  # input in the last build's order
    meshes2 = [(p_new, n0, uv0, idx0), meshes[1]]
    o, _ = _oracle(meshes2, [b0, builds[1]], cur, lights, ori4, dir4)
    inst.trace(r, hb, n)
    inst.trace_shadow(r, hb, n)
    g = hb.read(n)
    for k in FIELDS:
        assert np.array_equal(g[k], o[k]), f"rebuilt scene: {k} differs"
    inst.destroy()
    r.destroy()
    hb.destroy()
    for s in scs:
        s.destroy()


def test_instance_errors(ctx):
    import rtx
    p, nrm, uv, idx = sphere_mesh(8, 6)
    s = load_scene(ctx, p, idx, nrm, uv, np.zeros(len(idx), np.int32))
    I4 = np.eye(4, dtype=np.float32)
    with pytest.raises(rtx.RtxError):
        rtx.Instances(ctx, [], np.zeros((0, 16), np.float32), np.zeros((0, 16), np.float32))
    inst = rtx.Instances(ctx, [s], I4, I4)
    r = rtx.Rays(ctx, 16)
    hb = rtx.Hits(ctx, 16)
    with pytest.raises(rtx.RtxError, match="no BVH"):
        inst.trace(r, hb, 16)
    s.build(30, 1)
    inst.trace(r, hb, 0)
    inst.destroy()
    r.destroy()
    hb.destroy()
    s.destroy()


@pytest.mark.parametrize("case", ["rigid", "overlap"])
def test_frame_over_instances(ctx, case):
    """Renderer::render_to_framebuffer over a scene of objects: primary rays,
    the multi-root trace and shadow rays, secondary rays and shading through
    the shared material map (a mesh takes its first instance's material:
    reflective and refractive ones make bounces), bit for bit."""
    import rtx
    from rtx import make_materials
    make, insts = CASES[case]
    meshes = make()
    lights = _lights()
    scs = _load(ctx, meshes, lights)
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    inst = rtx.Instances(ctx, [scs[m] for m, _, _ in insts], [t for _, t, _ in insts], [ti for _, _, ti in insts])
    inst.set_lights(lights)
    rows = [(0.9, 0.2, 0.2, 0.5, 0.6, 1.5), (0.2, 0.8, 0.3, 1.0, 0.7, 0.0), (0.3, 0.3, 0.9, 0.2, 0.0, 0.0),
            (0.8, 0.8, 0.8, 1.0, 0.9, 1.33)]
    mats = make_materials(rows[: len(insts)])
    inst.set_materials(mats)
    w, h = 160, 120
    cam = rtx.camera_set((0.0, 0.8, -5.0), (0.0, 0.3, 0.0), (0.0, 1.0, 0.0), np.pi / 3, w / h)
    R = rtx.Renderer(ctx, w, h, 3)
    st = R.frame_instances(inst, cam)
    ms = O.MultiScene(builds, [m[1] for m in meshes], [m[2] for m in meshes], insts)
    oacc, ost, orays = O.render_multi(ms, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3)
    assert (oacc.sum(1) > 0).mean() > 0.15 and st.secondary_rays > 0
    assert st.primary_rays + st.secondary_rays == orays
    assert np.array_equal(R.fb.accum(), oacc)
    # the other root schedules give the same frame (bounce rays are incoherent)
    for lane_roots in (1, 0):
        ctx.tune("inst_lane_roots", lane_roots)
        try:
            R.frame_instances(inst, cam)
        finally:
            ctx.tune("inst_lane_roots", 2)
        assert np.array_equal(R.fb.accum(), oacc), f"root schedule {lane_roots}"
    # new transforms: the next frame follows them, no rebuild
    moved = [(m, *xform(rot((0, 1, 0), 0.4 * (i + 1)), 1.0, (1.2 * i - 1.2, 0.1 * i, 0.0))) for i, (m, _, _) in
             enumerate(insts)]
    inst.set_transforms([t for _, t, _ in moved], [ti for _, _, ti in moved])
    R.frame_instances(inst, cam)
    ms2 = O.MultiScene(builds, [m[1] for m in meshes], [m[2] for m in meshes], moved)
    oacc2, _, _ = O.render_multi(ms2, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3)
    assert np.array_equal(R.fb.accum(), oacc2)
    R.destroy()
    inst.destroy()
    for s in scs:
        s.destroy()


def test_textured_frame_over_instances(ctx):
    """Textures and the cubemap on the instances' shading scene: a mesh's
    texture index comes from its first instance's material, misses take the
    sky, bit for bit; with the cubemap off again misses add nothing.  Images
    before materials is an error."""
    import rtx
    from rtx import make_materials
    from test_gpu_textures import _images
    make, insts = CASES["rigid"]
    meshes = make()
    lights = _lights()
    scs = _load(ctx, meshes, lights)
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    inst = rtx.Instances(ctx, [scs[m] for m, _, _ in insts], [t for _, t, _ in insts], [ti for _, _, ti in insts])
    inst.set_lights(lights)
    textures, faces = _images(4)
    with pytest.raises(rtx.RtxError):
        inst.set_textures(textures)
    rows = [(0.9, 0.2, 0.2, 0.5, 0.6, 0.0), (0.2, 0.8, 0.3, 1.0, 0.0, 0.0), (0.3, 0.3, 0.9, 0.2, 0.0, 0.0),
            (0.8, 0.8, 0.8, 1.0, 0.9, 1.33)]
    mats = make_materials(rows[: len(insts)])
    mi = mats.view(np.int32).reshape(-1, 8)
    mi[:, 7] = [k % 3 - 1 for k in range(1, len(insts) + 1)]  # 0, 1, -1, 0, ...: image, image, none
    inst.set_materials(mats)
    inst.set_textures(textures)
    inst.set_cubemap(faces)
    w, h = 160, 120
    cam = rtx.camera_set((0.0, 0.8, -5.0), (0.0, 0.3, 0.0), (0.0, 1.0, 0.0), np.pi / 3, w / h)
    R = rtx.Renderer(ctx, w, h, 3)
    R.frame_instances(inst, cam)
    ms = O.MultiScene(builds, [m[1] for m in meshes], [m[2] for m in meshes], insts)
    oacc, _, _ = O.render_multi(ms, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3,
                                tex=O.TexSet(textures, faces))
    plain, _, _ = O.render_multi(ms, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3)
    assert (oacc != plain).any(1).mean() > 0.5, "the sky and the textures should change much of the frame"
    assert np.array_equal(R.fb.accum(), oacc)
    inst.set_cubemap(None)
    R.frame_instances(inst, cam)
    oacc2, _, _ = O.render_multi(ms, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3,
                                 tex=O.TexSet(textures))
    assert np.array_equal(R.fb.accum(), oacc2)
    R.destroy()
    inst.destroy()
    for s in scs:
        s.destroy()


def test_frame_after_scene_geometry_changes(ctx):
    """A frame over instances, then one scene's geometry replaced by a mesh of
    another triangle count and rebuilt: the next frame renumbers the shared
    ids and rebuilds the material map (each mesh its first instance's
    material), bit for bit with the oracle's frame over the new meshes."""
    import rtx
    from rtx import make_materials
    meshes = two_meshes()
    lights = _lights()
    scs = _load(ctx, meshes, lights)
    insts = [(0, *xform(rot((0, 1, 0), 0.5), 1.0, (-1.5, 0.2, 0.0))),
             (1, *xform(rot((1, 0, 0), 0.9), 1.4, (1.4, -0.2, 0.5))),
             (0, *xform(np.eye(3), 0.6, (0.0, 1.7, -0.5)))]
    inst = rtx.Instances(ctx, [scs[m] for m, _, _ in insts], [t for _, t, _ in insts], [ti for _, _, ti in insts])
    inst.set_lights(lights)
    mats = make_materials([(0.9, 0.3, 0.2, 0.5, 0.6, 0.0), (0.2, 0.8, 0.3, 1.0, 0.3, 1.5),
                           (0.3, 0.3, 0.9, 0.2, 0.0, 0.0)])
    inst.set_materials(mats)
    w, h = 160, 120
    cam = rtx.camera_set((0.0, 0.8, -5.0), (0.0, 0.3, 0.0), (0.0, 1.0, 0.0), np.pi / 3, w / h)
    R = rtx.Renderer(ctx, w, h, 3)
    for step, mesh0 in enumerate((meshes[0], sphere_mesh(40, 28, 0.9))):
        if step:
            p, nn, uv, idx = mesh0
            scs[0].set_geometry(p, idx, nn, uv, np.zeros(len(idx), np.int32))
            scs[0].build(30, 1)
        cur = [mesh0, meshes[1]]
        builds = [O.Build(m[0], m[3], None, 30, 1) for m in cur]
        st = R.frame_instances(inst, cam)
        ms = O.MultiScene(builds, [m[1] for m in cur], [m[2] for m in cur], insts)
        oacc, _, orays = O.render_multi(ms, mats, O.lights_from(lights), O.camera_from(cam), w, h, 3)
        assert st.primary_rays + st.secondary_rays == orays
        assert np.array_equal(R.fb.accum(), oacc), f"step {step}: accumulators differ"
        assert np.array_equal(inst.tri_base(), np.asarray([0, len(mesh0[3]), 0], np.uint32))
    R.destroy()
    inst.destroy()
    for s in scs:
        s.destroy()


def _world_boxes(meshes, insts):
    out = []
    for m, tr, _ in insts:
        M = tr.astype(np.float64)
        w = meshes[m][0].astype(np.float64) @ M[:3, :3].T + M[:3, 3]
        out.append((w.min(0), w.max(0)))
    return out


def test_culling_boundaries(ctx):
    """Rays at the edges of root culling (inst_lane_roots 2) against the
    oracle, which walks every root: axis-parallel rays along the faces, edges
    and middles of the instances' world boxes (zero direction components),
    rays from 1e4 away, rays whose tMax ends just short of, at or just past
    their hit, and instances with large and tiny scales; every form and
    schedule, shadow rays included."""
    import rtx
    meshes = two_meshes()
    lights = _lights()
    scs = _load(ctx, meshes, lights)
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    insts = [(0, *xform(rot((0, 1, 0), 0.3), 1.0, (-2.0, 0.0, 0.0))),
             (1, *xform(np.eye(3), 1.0, (2.0, 0.5, 0.0))),
             (0, *xform(rot((1, 1, 0), 1.1), 250.0, (0.0, 0.0, 900.0))),
             (1, *xform(np.eye(3), 0.004, (0.0, 3.0, 0.0)))]
    rng = np.random.default_rng(77)
    O4, D4 = [], []

    def ray(o, d, tmax=1e30):
        O4.append([*o, 0.0])
        D4.append([*d, tmax])
    for lo, hi in _world_boxes(meshes, insts):
        mid = (lo + hi) / 2
        ext = float((hi - lo).max())
        for a in range(3):
            b, c = [k for k in range(3) if k != a]
            for vb in (lo[b], mid[b], hi[b]):
                for vc in (lo[c], mid[c], hi[c]):
                    for sgn in (1.0, -1.0):
                        o = np.zeros(3)
                        o[a] = (lo[a] - ext) if sgn > 0 else (hi[a] + ext)
                        o[b], o[c] = vb, vc
                        d = np.zeros(3)
                        d[a] = sgn
                        ray(o, d)
        for _ in range(60):  # from far away toward the box
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            o = mid + 1e4 * u
            tgt = mid + rng.uniform(-0.5, 0.5, 3) * (hi - lo)
            d = (tgt - o) / np.linalg.norm(tgt - o)
            ray(o, d)
    O4 = np.asarray(O4, np.float32)
    D4 = np.asarray(D4, np.float32)
    # tMax at, just short of and just past the hits of a first pass
    o0, _ = _oracle(meshes, builds, insts, lights, O4, D4)
    hit = np.nonzero((o0["flags"] & 1) != 0)[0]
    assert hit.size > 100
    t = o0["t"][hit]
    extra_o, extra_d = [], []
    for tm in (t, np.nextafter(t, np.float32(0)), np.nextafter(t, np.float32(np.inf)),
               (t * np.float32(1 - 1e-6)).astype(np.float32)):
        extra_o.append(O4[hit])
        dd = D4[hit].copy()
        dd[:, 3] = tm
        extra_d.append(dd)
    O4 = np.concatenate([O4, *extra_o])
    D4 = np.concatenate([D4, *extra_d])
    n = len(O4)
    o, _ = _oracle(meshes, builds, insts, lights, O4, D4)
    r = rtx.Rays(ctx, n)
    r.write(O4, D4)
    hb = rtx.Hits(ctx, n)
    inst = rtx.Instances(ctx, [scs[m] for m, _, _ in insts], [t for _, t, _ in insts], [ti for _, _, ti in insts])
    inst.set_lights(lights)
    _check_all_forms(ctx, inst, r, hb, n, o, "culling boundaries")
    inst.destroy()
    r.destroy()
    hb.destroy()
    for s in scs:
        s.destroy()


def test_many_instances_culled_in_chunks(ctx):
    """100 instances (a 10 x 10 grid of the two meshes under rotations and
    scales): root culling runs over two chunks of roots (64 + 36); closest
    hits and shadow rays equal the oracle's in every form and schedule."""
    import rtx
    meshes = two_meshes()
    lights = _lights()
    scs = _load(ctx, meshes, lights)
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    insts = []
    for k in range(100):
        gx, gy = k % 10, k // 10
        insts.append((k % 2, *xform(rot((gx + 1.0, 1.0, gy + 1.0), 0.37 * k), 0.3 + 0.02 * (k % 7),
                                    (1.1 * (gx - 4.5), 1.1 * (gy - 4.5), 0.2 * (k % 3)))))
    n = 6000
    ori4, dir4 = rays_at(31, n, spread=8.0)
    rng = np.random.default_rng(5)
    tgt = rng.uniform(-5.5, 5.5, (n, 3))
    tgt[:, 2] = rng.uniform(-0.5, 0.7, n)
    d = tgt - ori4[:, :3]
    dir4[:, :3] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o, _ = _oracle(meshes, builds, insts, lights, ori4, dir4)
    tris_hit = np.unique(o["id"][(o["flags"] & 1) != 0])
    assert ((o["flags"] & 1) != 0).sum() > 1000 and tris_hit.size > 50
    r = rtx.Rays(ctx, n)
    r.write(ori4, dir4)
    hb = rtx.Hits(ctx, n)
    inst = rtx.Instances(ctx, [scs[m] for m, _, _ in insts], [t for _, t, _ in insts], [ti for _, _, ti in insts])
    inst.set_lights(lights)
    _check_all_forms(ctx, inst, r, hb, n, o, "100 instances")
    inst.destroy()
    r.destroy()
    hb.destroy()
    for s in scs:
        s.destroy()


def _random_set(rng):
    """Instances of three meshes: mostly rigid with scales up to 100x either
    way, some projective; 2 to 80 of them (around the 64-root chunk)."""
    insts = []
    for _ in range(int(rng.choice([2, 16, 63, 64, 65, 80]))):
        m = int(rng.integers(0, 3))
        if rng.uniform() < 0.1:
            M = np.eye(4)
            M[:3, :3] = rot(rng.normal(size=3), rng.uniform(0, 3)) * rng.uniform(0.5, 2)
            M[:3, 3] = rng.uniform(-3, 3, 3)
            Mi = np.linalg.inv(M)
            Mi[3, :3] = rng.uniform(-0.02, 0.02, 3)
            insts.append((m, np.linalg.inv(Mi).astype(np.float32), Mi.astype(np.float32)))
        else:
            sc = float(np.exp(rng.uniform(np.log(0.01), np.log(100.0)))) if rng.uniform() < 0.2 else rng.uniform(0.3, 1.5)
            insts.append((m, *xform(rot(rng.normal(size=3), rng.uniform(0, 6)), sc, rng.uniform(-6, 6, 3))))
    return insts


def _random_rays(rng, n):
    """Rays at the scene, a fifth from 1e3 times farther, some axis-parallel,
    a fifth with a finite tMax."""
    o = rng.uniform(-8, 8, (n, 3))
    far = rng.uniform(size=n) < 0.2
    o[far] *= 1e3
    d = rng.uniform(-6, 6, (n, 3)) - o
    ax = rng.uniform(size=n) < 0.15
    keep = rng.integers(0, 3, n)
    d[ax] = 0.0
    d[ax, keep[ax]] = np.sign(rng.normal(size=ax.sum())) + 0.0
    o[ax] = rng.uniform(-6, 6, (ax.sum(), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ori4 = np.zeros((n, 4), np.float32)
    dir4 = np.zeros((n, 4), np.float32)
    ori4[:, :3] = o
    dir4[:, :3] = d
    dir4[:, 3] = np.where(rng.uniform(size=n) < 0.2, rng.uniform(0.5, 15, n), 1e30)
    return ori4, dir4


@pytest.mark.parametrize("seed", range(8))
def test_random_instance_sets(ctx, seed):
    """Random instance sets and rays (see _random_set, _random_rays): every
    form and root schedule equals the oracle bit for bit."""
    import rtx
    rng = np.random.default_rng(1000 + seed)
    meshes = [sphere_mesh(12, 8, 1.0), sphere_mesh(20, 14, 0.6), sphere_mesh(9, 6, 1.3)]
    lights = _lights()
    scs = _load(ctx, meshes, lights)
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    insts = _random_set(rng)
    n = 3000
    ori4, dir4 = _random_rays(rng, n)
    o, _ = _oracle(meshes, builds, insts, lights, ori4, dir4)
    assert ((o["flags"] & 1) != 0).sum() > 50
    r = rtx.Rays(ctx, n)
    r.write(ori4, dir4)
    hb = rtx.Hits(ctx, n)
    inst = rtx.Instances(ctx, [scs[m] for m, _, _ in insts], [t for _, t, _ in insts], [ti for _, _, ti in insts])
    inst.set_lights(lights)
    _check_all_forms(ctx, inst, r, hb, n, o, f"random set {seed} ({len(insts)} instances)")
    inst.destroy()
    r.destroy()
    hb.destroy()
    for s in scs:
        s.destroy()
