"""The library's OBJ/MTL loader (rtx_obj_*) against the reference's own
loader (ModelOBJ, src/rt/obj-loader.cpp), bit for bit.

Expected outputs in tests/golden/obj/*.refN.npz were written by the
reference loader itself (oracle/_ref/objdump, built from the reference tree
by oracle/Makefile.ref; tests/golden/make_obj_golden.py), with
rebuildNormals = N.  The fixtures are four of the reference's models and
small files for the loader's rules: every face form, polygons, negative
indices, generated normals, MTL keywords, unknown materials and the alpha
order of material groups.  When the reference tree and that build are
present, every model the reference ships is also compared live.  Host code
only: no GPU.
"""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))
from objdump_format import from_npz, parse  # noqa: E402

OBJ_DIR = HERE / "golden" / "obj"
FIXTURES = sorted(OBJ_DIR.glob("*.obj"))
DUMP = HERE.parent / "oracle" / "_ref" / "objdump"
REF_MODELS = Path("/root/reference/models/obj")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _compare(m, ref, what):
    import rtx  # noqa: F401
    for k in ("pos", "nrm", "uv"):
        assert getattr(m, k).shape == ref[k].shape, f"{what}: {k} shape"
        assert np.array_equal(_bits(getattr(m, k)), _bits(ref[k])), f"{what}: {k} differs"
    assert np.array_equal(m.idx, ref["idx"]), f"{what}: triangles differ"
    assert len(m.groups) == len(ref["groups"]), f"{what}: group count"
    for g, (a, b) in enumerate(zip(m.groups, ref["groups"])):
        assert (a["first"], a["count"]) == (b["first"], b["count"]), f"{what}: group {g} range"
        assert np.float32(a["alpha"]) == np.float32(b["alpha"]) and \
            np.float32(a["shininess"]) == np.float32(b["shininess"]), f"{what}: group {g} material"
        assert np.array_equal(np.float32(a["diffuse"][:3]), np.float32(b["diffuse"])), f"{what}: group {g} diffuse"
        assert a["color_map"] == b["color_map"], f"{what}: group {g} color map"
        # get_meshes: the reference's mesh holds `count` zero triangles, then the real ones
        c = b["count"]
        assert b["idx"].shape[0] == 2 * c and not b["idx"][:c].any()
        assert np.array_equal(a["idx"], b["idx"][c:]), f"{what}: group {g} mesh triangles"
        for k in ("pos", "nrm", "uv"):
            assert np.array_equal(_bits(a[k]), _bits(b[k])), f"{what}: group {g} mesh {k}"


@pytest.mark.parametrize("rebuild", [0, 1])
@pytest.mark.parametrize("obj", FIXTURES, ids=[p.stem for p in FIXTURES])
def test_loader_matches_reference_fixture(obj, rebuild):
    import rtx
    m = rtx.ObjModel(obj, rebuild_normals=bool(rebuild))
    _compare(m, from_npz(obj.with_name(f"{obj.stem}.ref{rebuild}.npz")), f"{obj.name} rebuild={rebuild}")


def test_fixture_rules_are_exercised():
    import rtx
    m = rtx.ObjModel(OBJ_DIR / "forms.obj")
    assert len(m.idx) == 13 and len(m.groups) == 3
    names = {g["name"] for g in m.groups}
    assert names == {"red", "glass"}
    glass = next(g for g in m.groups if g["name"] == "glass")
    # Ni follows Ns and also sets the shininess (the reference dispatches on 'N')
    assert np.float32(glass["shininess"]) == np.float32(1.5) / np.float32(1000.0)
    assert np.float32(glass["alpha"]) == np.float32(1.0) - np.float32(0.25)
    assert glass["specular"][:3] == [0.0, 0.0, 0.0] and glass["bump_map"] == "glass_bump.png"
    a = rtx.ObjModel(OBJ_DIR / "alpha.obj")
    assert [g["name"] for g in a.groups] == ["solid", "solid2", "half", "half", "clear"]
    n = rtx.ObjModel(OBJ_DIR / "nonormals.obj")
    assert np.allclose(np.linalg.norm(n.nrm, axis=1), 1.0, atol=1e-6)


def test_loader_errors(tmp_path):
    import rtx
    with pytest.raises(rtx.RtxError, match="cannot open"):
        rtx.ObjModel(tmp_path / "missing.obj")
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n")
    with pytest.raises(rtx.RtxError, match="out of range"):
        rtx.ObjModel(bad)
    short = tmp_path / "short.obj"
    short.write_text("v 0 0 0\nv 1 0 0\nf 1 2\n")
    with pytest.raises(rtx.RtxError, match="fewer than 3"):
        rtx.ObjModel(short)


@pytest.mark.skipif(not (DUMP.exists() and REF_MODELS.exists()), reason="reference tree or oracle/_ref not present")
@pytest.mark.parametrize("name", sorted(p.name for p in REF_MODELS.glob("*.obj")) if REF_MODELS.exists() else [])
def test_every_reference_model_live(name):
    import rtx
    r = subprocess.run([str(DUMP), str(REF_MODELS / name), "0"], capture_output=True, text=True, check=True,
                       timeout=300)
    _compare(rtx.ObjModel(REF_MODELS / name), parse(r.stdout), name)


def test_geom_matrix_closed_forms():
    from rtx import scenes
    M = scenes.geom_matrix((1.0, -8.0, 3.0), (0.0, 0.0, 0.0), 2.0)
    assert np.array_equal(M, np.float32([[2, 0, 0, 1], [0, 2, 0, -8], [0, 0, 2, 3], [0, 0, 0, 1]]))
    # rotations: orthonormal, det 1, and a quarter turn about each axis maps the axes
    for rpy, (src, dst) in (((np.pi / 2, 0, 0), ((1, 0, 0), (0, 1, 0))),    # roll about z
                            ((0, np.pi / 2, 0), ((0, 1, 0), (0, 0, 1))),    # pitch about x
                            ((0, 0, np.pi / 2), ((0, 0, 1), (1, 0, 0)))):   # yaw about y
        R = scenes.geom_matrix((0, 0, 0), rpy, 1.0)[:3, :3].astype(np.float64)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-6) and abs(np.linalg.det(R) - 1) < 1e-6
        assert np.allclose(R @ np.array(src, float), dst, atol=1e-6), rpy
    # the transform of points and normals (w = 1 for both after homogenize)
    p, n = scenes.transform_vertices(M, np.float32([[1, 2, 3], [-0.5, 0, 0]]), np.float32([[0, 1, 0], [1, 0, 0]]))
    assert np.array_equal(p, np.float32([[3, -4, 9], [0, -8, 3]]))
    # normals: M (n, 0), no translation, w then set to 1: scaled by 2, not renormalized
    assert np.array_equal(n, np.float32([[0, 2, 0], [2, 0, 0]]))


def test_boat_preset_aggregate():
    import rtx
    from rtx import scenes
    sd = scenes.preset_boat(OBJ_DIR, res=(64, 36))
    water, boat = rtx.ObjModel(OBJ_DIR / "frame_water1.obj"), rtx.ObjModel(OBJ_DIR / "frame_boat1.obj")
    assert sd.T == len(water.idx) + len(boat.idx)
    assert np.array_equal(sd.idx[len(water.idx):], boat.idx + len(water.pos))
    assert np.array_equal(sd.pos[: len(water.pos)], water.pos * np.float32(2) + np.float32([0, -8, 0]))
    assert (sd.matmap[: len(water.idx)] == 0).all() and (sd.matmap[len(water.idx):] == 1).all()
