"""Pins of the oracle's multi-root traversal (trace_multi / shadow_trace_multi,
src/kernel/trace-bvh.cl:419-463 and src/kernel/shadow-trace-bvh.cl:317-373).

The reference has no tests or fixtures for this path, so the oracle is held
against what the mathematics fixes: an identity root is the single-root
trace; instances under rigid transforms and uniform scales hit what a brute
force over the same triangles moved into the world hits, at the same world
distance; the first root wins an exact tie; normals go through `tr`; a
shadow ray is blocked exactly when a brute-force any-hit over the moved
triangles finds a blocker; and a spot light is judged by that kernel's own
cone test.
"""
import numpy as np

import oracle as O

import rtx
from rtx import scenes
from inst_scene import rot as _rot, xform as _tr, sphere_mesh as _sphere_mesh, rays_at as _rays


def _world(ms, instances, builds, pos_list):
    """The instances' triangles moved into the world, indexed as the shared
    index buffer (global triangle id = mesh offset + sorted position)."""
    P, I = [], []
    base = 0
    for m, tr, _ in instances:
        b = builds[m]
        p = np.c_[b.pos4[:, :3].astype(np.float64), np.ones(len(b.pos4))] @ np.asarray(tr, np.float64).T
        P.append(p[:, :3])
        I.append(b.idx.reshape(-1, 3) + base)
        base += len(p)
    return np.concatenate(P).astype(np.float32), np.concatenate(I).astype(np.uint32)


def _setup(instances_spec):
    pos, nrm, uv, idx = _sphere_mesh()
    pos2, nrm2, uv2, idx2 = _sphere_mesh(16, 12, 0.7)
    meshes = [(pos, nrm, uv, idx), (pos2, nrm2, uv2, idx2)]
    builds = [O.Build(m[0], m[3], None, 30, 1) for m in meshes]
    ms = O.MultiScene(builds, [m[1] for m in meshes], [m[2] for m in meshes], instances_spec)
    return meshes, builds, ms


def test_identity_root_is_single_trace():
    sd = scenes.config_c1()
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    I4 = np.eye(4, dtype=np.float32)
    ms = O.MultiScene([b], [sd.nrm], [sd.uv], [(0, I4, I4)])
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], 1.0)
    oo, od, _, _ = O.raygen_primary(O.camera_from(cam), 96, 96)
    h1 = O.trace(b, sd.nrm, sd.uv, oo, od)
    hm = O.trace_multi(ms, oo, od)
    assert (h1["id"] >= 0).sum() > 1000
    assert np.array_equal(h1["id"], hm["id"])
    hit = h1["id"] >= 0
    # world t is the distance to the transformed hit point: the same within rounding
    assert np.allclose(hm["t"][hit], h1["t"][hit], rtol=2e-6)


def test_transformed_instances_match_brute_force():
    insts = [(0, *_tr(_rot((0, 1, 0), 0.7), 1.0, (-2.0, 0.3, 0.0))),
             (1, *_tr(_rot((1, 0, 1), 1.9), 2.0, (2.0, -0.5, 1.0))),
             (0, *_tr(_rot((0, 0, 1), -0.4), 0.5, (0.0, 2.0, -1.0)))]
    meshes, builds, ms = _setup(insts)
    ori4, dir4 = _rays(3, 3000)
    h = O.trace_multi(ms, ori4, dir4)
    wpos, widx = _world(ms, insts, builds, None)
    wb = O.Build(wpos, widx, None, 30, 1)
    bid, bt = O.trace_brute(wb, ori4, dir4)
    # brute force ids are of the world mesh (its own sort): map through positions
    hit = h["id"] >= 0
    assert hit.sum() > 500
    assert np.array_equal(hit, bid >= 0)
    assert np.allclose(h["t"][hit], bt[hit], rtol=2e-5)
    # hit points lie on the moved triangles: the world point at t along the ray
    p = ori4[hit, :3] + dir4[hit, :3] * h["t"][hit, None]
    assert np.allclose(h["p4"][hit, :3], p, atol=1e-5)


def test_first_root_wins_a_tie_and_normals_go_through_tr():
    # meshes 0 and 1 are the same triangles under identical transforms: every
    # hit is an exact tie between the roots, and the first root's id (mesh 0's
    # range) must be reported
    pos, nrm, uv, idx = _sphere_mesh()
    b0 = O.Build(pos, idx, None, 30, 1)
    b1 = O.Build(pos, idx, None, 30, 1)
    tr, tri = _tr(_rot((0, 1, 0), 0.9), 1.5, (0.0, 0.0, 0.0))
    ms = O.MultiScene([b0, b1], [nrm, nrm], [uv, uv], [(1, tr, tri), (0, tr, tri)])
    ori4, dir4 = _rays(4, 2000)
    h = O.trace_multi(ms, ori4, dir4)
    hit = h["id"] >= 0
    assert hit.sum() > 300
    # instance 0 uses mesh 1, laid out first: ids below its triangle count
    assert np.all(h["id"][hit] < b1.T)
    # the normal of the sphere at the hit is radial; tr rotates and scales it
    # by 1.5 (the reference does not renormalize), and it faces the ray
    p = h["p4"][hit, :3]
    n = h["n4"][hit, :3]
    radial = p / np.linalg.norm(p, axis=1, keepdims=True)
    cosang = np.abs(np.sum(n * radial, axis=1)) / np.linalg.norm(n, axis=1)
    assert np.median(cosang) > 0.99
    assert np.allclose(np.median(np.linalg.norm(n, axis=1)), 1.5, rtol=0.02)
    assert np.all(np.sum(n * dir4[hit, :3], axis=1) <= 0)


def test_shadow_multi_matches_brute_force_and_spot_rule():
    from rtx import make_lights, LIGHT_DIR, LIGHT_SPOT
    insts = [(0, *_tr(np.eye(3), 1.0, (0.0, 0.0, 0.0))), (1, *_tr(np.eye(3), 1.0, (0.0, 2.5, 0.0)))]
    meshes, builds, ms = _setup(insts)
    ori4, dir4 = _rays(5, 2000)
    h = O.trace_multi(ms, ori4, dir4)
    lights = O.lights_from(make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_DIR, dir=(0.0, -1.0, 0.0),
                                                             color=(1.0, 1.0, 1.0))]))
    f = O.trace_shadow_multi(ms, lights, h["flags"], h["p4"])
    wpos, widx = _world(ms, insts, builds, None)
    wb = O.Build(wpos, widx, None, 30, 1)
    fb = O.trace_shadow_brute(wb, lights, h["flags"], h["p4"])
    hit = (h["flags"] & 1) != 0
    assert ((f >> 8) & 1)[hit].sum() > 50, "the upper sphere should shadow part of the lower one"
    assert np.mean(f[hit] == fb[hit]) > 0.999
    # a spot light above the scene pointing down at it: the multi-root kernel
    # compares dot(toward the light, spot dir) (about -1 here) with the angle
    # and lights every point untraced, while the single-root kernel's test
    # (dot(from the light, spot dir) against the angle) traces them and finds
    # the upper sphere's shadow
    spot = O.lights_from(make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_SPOT, pos=(0.0, 10.0, 0.0),
                                                           dir=(0.0, -1.0, 0.0), angle=0.9,
                                                           color=(1.0, 1.0, 1.0))]))
    fs = O.trace_shadow_multi(ms, spot, h["flags"], h["p4"])
    assert not ((fs >> 8) & 1).any()
    f1 = O.trace_shadow_brute(wb, spot, h["flags"], h["p4"])
    assert ((f1 >> 8) & 1)[hit].sum() > 50
