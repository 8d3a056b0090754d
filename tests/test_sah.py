"""The library's CPU SAH builder (rtx_sah_build_host / rtx_sah_build) against
the reference's own BVH::construct (src/rt/bvh.cpp), bit for bit: node boxes,
children or triangle ranges, parents, leaf flags and the triangle order.

Expected outputs in tests/golden/sah/*.npz were written by the reference
builder itself (oracle/_ref/sahdump, built from the reference tree by
oracle/Makefile.ref; tests/golden/make_sah_golden.py).  Cases: 1, 2, 3 and 17
triangles, a 5000-triangle soup, geometry entirely below zero (the empty
bucket box's hi of FLT_MIN), duplicated triangles, a flat axis and a regular
grid (equal centroids).  When the reference tree and that build are present,
the OBJ models are also compared live.  Host code only."""
import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
GOLD = HERE / "golden" / "sah"
CASES = sorted(p.stem for p in GOLD.glob("*.npz"))
DUMP = HERE.parent / "oracle" / "_ref" / "sahdump"


def _check(nodes, order, box, words, ref_order, what):
    assert len(nodes) == len(box), f"{what}: node count {len(nodes)} != {len(box)}"
    got_box = np.concatenate([nodes["lo"][:, :3], nodes["hi"][:, :3]], 1).astype(np.float32)
    assert np.array_equal(got_box.view(np.uint32), box.view(np.uint32)), f"{what}: node boxes differ"
    assert np.array_equal(nodes["a"], words[:, 0]) and np.array_equal(nodes["b"], words[:, 1]), \
        f"{what}: children / ranges differ"
    assert np.array_equal(nodes["parent"], words[:, 2]), f"{what}: parents differ"
    assert np.array_equal(nodes["leaf"] != 0, words[:, 3] != 0), f"{what}: leaf flags differ"
    assert np.array_equal(order, ref_order), f"{what}: triangle order differs"


@pytest.mark.parametrize("case", CASES)
def test_sah_matches_reference_fixture(case):
    import rtx
    z = np.load(GOLD / f"{case}.npz")
    nodes, order = rtx.sah_build_host(z["pos"], z["idx"])
    _check(nodes, order, z["box"], z["words"], z["order"], case)


def test_sah_tree_invariants():
    import rtx
    z = np.load(GOLD / "soup_5000.npz")
    nodes, order = rtx.sah_build_host(z["pos"], z["idx"])
    T = len(z["idx"])
    assert sorted(order.tolist()) == list(range(T))
    leaf = nodes["leaf"] != 0
    # leaves cover [0, T) exactly once; internal nodes have adjacent children
    cover = np.zeros(T, int)
    for n in np.nonzero(leaf)[0]:
        cover[nodes["a"][n]:nodes["b"][n]] += 1
    assert (cover == 1).all()
    assert (nodes["b"][~leaf] == nodes["a"][~leaf] + 1).all()
    # a node's box is the union of its triangles' boxes
    p = z["pos"][z["idx"][order]]
    for n in np.nonzero(leaf)[0][:200]:
        tri = p[nodes["a"][n]:nodes["b"][n]].reshape(-1, 3)
        assert np.array_equal(nodes["lo"][n, :3], tri.min(0)) and np.array_equal(nodes["hi"][n, :3], tri.max(0))


@pytest.mark.skipif(not (DUMP.exists() and Path("/root/reference/models/obj").exists()),
                    reason="reference tree or oracle/_ref not present")
@pytest.mark.parametrize("name", ["teapot-low_res.obj", "frame_boat1.obj", "frame_water1.obj", "teapot.obj"])
def test_sah_reference_models_live(name):
    import rtx
    sys.path.insert(0, str(HERE / "golden"))
    from make_sah_golden import run_reference
    m = rtx.ObjModel(Path("/root/reference/models/obj") / name)
    box, words, ref_order = run_reference(m.pos, m.idx)
    nodes, order = rtx.sah_build_host(m.pos, m.idx)
    _check(nodes, order, box, words, ref_order, name)


def test_oracle_on_a_tree_deeper_than_the_stack():
    """The 36-level SAH tree: the oracle's traversal keeps to the reference's
    MAX_LEVELS rule (a ray ends at a pop with 32 entries) without storing
    entries beyond the stack; every hit it reports is a real hit no nearer
    than the brute force's, and the rule ends some rays early."""
    import oracle as O
    from test_gpu_sah import axis_rays
    z = np.load(GOLD / "deep.npz")
    tree = O.Tree(z["pos"], z["idx"], None, z["box"], z["words"], z["order"])
    ori4, dir4 = axis_rays(7, 4096)
    nrm = np.tile(np.float32([0, 0, 1]), (len(z["pos"]), 1))
    h = O.trace(tree, nrm, np.zeros((len(z["pos"]), 2), np.float32), ori4, dir4)
    bid, bt = O.trace_brute(tree, ori4, dir4)
    both = (h["id"] >= 0) & (bid >= 0)
    assert not ((h["id"] >= 0) & (bid < 0)).any()
    assert (h["t"][both] >= bt[both]).all()
    assert ((h["id"] >= 0) != (bid >= 0)).sum() > 100


def test_stackless_walk_pins():
    """The oracle's stackless walk (trace-bvh-no-stack.cl): with no stack it
    has no MAX_LEVELS cut, so on the 36-level tree it finds what a brute
    force finds (the same triangle and t: the rays have no exact ties); on a
    shallow tree it equals the stack walk; shadow flags equal brute force."""
    import oracle as O
    from rtx import make_lights, LIGHT_POINT
    from test_gpu_sah import axis_rays, _rays
    z = np.load(GOLD / "deep.npz")
    tree = O.Tree(z["pos"], z["idx"], None, z["box"], z["words"], z["order"])
    ori4, dir4 = axis_rays(7, 4096)
    nrm = np.tile(np.float32([0, 0, 1]), (len(z["pos"]), 1))
    uv = np.zeros((len(z["pos"]), 2), np.float32)
    with O.stackless():
        h = O.trace(tree, nrm, uv, ori4, dir4)
    bid, bt = O.trace_brute(tree, ori4, dir4)
    assert np.array_equal(h["id"], bid) and np.array_equal(h["t"][bid >= 0], bt[bid >= 0])
    z = np.load(GOLD / "soup_5000.npz")
    tree = O.Tree(z["pos"], z["idx"], None, z["box"], z["words"], z["order"])
    nrm = np.tile(np.float32([0, 0, 1]), (len(z["pos"]), 1))
    uv = np.zeros((len(z["pos"]), 2), np.float32)
    ori4, dir4 = _rays(31, 4096, z["pos"])
    stack = O.trace(tree, nrm, uv, ori4, dir4)
    with O.stackless():
        ns = O.trace(tree, nrm, uv, ori4, dir4)
    for k in ("id", "t", "bary", "n4", "p4"):
        assert np.array_equal(stack[k], ns[k]), k
    lights = O.lights_from(make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(0.0, 80.0, 0.0))]))
    with O.stackless():
        f = O.trace_shadow(tree, lights, stack["flags"], stack["p4"])
    fb = O.trace_shadow_brute(tree, lights, stack["flags"], stack["p4"])
    assert ((f >> 8) & 1).sum() > 100 and np.array_equal(f, fb)
