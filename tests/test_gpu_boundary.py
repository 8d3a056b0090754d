"""GPU tests of the reference-interface entry points beside the frame path:

  * the reference-layout vertex ingest (rtx_scene_set_vertices_ref, the
    80-byte Vertex of include/rt/mesh.hpp:23-30) builds the same tree and
    renders the same frame as rtx_scene_set_vertices, and both equal the
    oracle;
  * traces against an accelerator copy (rtx_bvh_copy + rtx_trace_with /
    rtx_trace_shadow_with, Tracer::trace(Scene&, DeviceMemory& bvh_mem, ...) of
    include/rt/tracer.hpp:21,26): after the scene moves and is rebuilt, the
    copy still gives the first tree's hits and shadow flags (the oracle's on
    the unmoved geometry), the scene itself the moved geometry's;
  * the same through the C++ stage classes (tests/cpp/stage_tracer.cpp), with
    BVHBuilder::set_log writing the build's stage times;
  * the autotuner's keys (rtx_tune_set / rtx_tune_get): documented ranges
    accepted and read back, values outside them refused.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tesina-rt_amd" / "bin" / "stage_tracer"
HIT_KEYS = ("flags", "id", "t", "bary", "n4", "p4")


def _vertex80(pos, nrm, uv, rng):
    """The reference's Vertex records: position, normal, tangent, bitangent
    (cl_float3/float4, 16 bytes each) and texCoord (cl_float2), aligned to 80
    bytes; tangent, bitangent and padding hold junk the ingest must ignore."""
    V = len(pos)
    v = rng.normal(size=(V, 20)).astype(np.float32) * 1e6
    v[:, 0:3] = pos
    v[:, 4:7] = nrm
    v[:, 16:18] = uv
    return v


def test_set_vertices_ref_matches_set_vertices(ctx):
    import rtx
    from rtx import scenes
    sd = scenes.config_c1()
    a = scenes.load(rtx.Scene(ctx), sd)
    b = rtx.Scene(ctx)
    b.set_vertices_ref(_vertex80(sd.pos, sd.nrm, sd.uv, np.random.default_rng(7)))
    b.set_triangles(sd.idx, sd.matmap)
    b.set_materials(sd.materials)
    b.set_lights(sd.lights)
    w, h = 200, 150
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], w / h)
    R = rtx.Renderer(ctx, w, h, 2)
    try:
        R.frame(a, cam, rebuild=True)
        acc_a = R.fb.accum()
        R.frame(b, cam, rebuild=True)
        acc_b = R.fb.accum()
        assert np.array_equal(a.nodes()["a"], b.nodes()["a"]) and np.array_equal(a.nodes()["lo"], b.nodes()["lo"])
        assert np.array_equal(acc_a, acc_b), "the 80-byte Vertex ingest renders a different frame"
        ob = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
        oacc, _, _ = O.render(ob, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h, 2)
        assert np.array_equal(acc_b, oacc)
    finally:
        R.destroy()
        a.destroy()
        b.destroy()


def _primary(rtx, ctx, cam, w, h):
    r = rtx.Rays(ctx, w * h)
    r.primary(cam, w, h)
    oo, od, _, _ = O.raygen_primary(O.camera_from(cam), w, h)
    return r, oo, od


def test_trace_with_accelerator_copy(ctx):
    import rtx
    from rtx import scenes
    sd = scenes.config_c1()
    s = scenes.load(rtx.Scene(ctx), sd)
    s.build(30, 1)
    other = scenes.load(rtx.Scene(ctx), sd)
    w, h = 160, 120
    n = w * h
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], w / h)
    r, oo, od = _primary(rtx, ctx, cam, w, h)
    hb = rtx.Hits(ctx, n)
    bvh = rtx.BVHCopy(ctx)
    try:
        with pytest.raises(rtx.RtxError, match="no LBVH"):
            bvh.copy(other)  # not built yet
        bvh.copy(s)
        ob = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
        lights = O.lights_from(sd.lights)
        want = O.trace(ob, sd.nrm, sd.uv, oo, od)
        want["flags"] = O.trace_shadow(ob, lights, want["flags"], want["p4"])
        # the scene moves and is rebuilt (from its permuted buffers)
        moved = sd.pos + np.float32([1.0, 0.0, 0.0])
        s.update_positions(moved)
        s.build(30, 1)
        rtx.trace(s, r, hb, n, bvh=bvh)
        rtx.trace_shadow(s, r, hb, n, bvh=bvh)
        g = hb.read(n)
        for k in HIT_KEYS:
            assert np.array_equal(g[k], want[k]), f"{k}: the copy must trace the first tree"
        ob2 = O.Build(moved, ob.idx.reshape(-1, 3), ob.matmap, 30, 1)
        want2 = O.trace(ob2, sd.nrm, sd.uv, oo, od)
        want2["flags"] = O.trace_shadow(ob2, lights, want2["flags"], want2["p4"])
        rtx.trace(s, r, hb, n)
        rtx.trace_shadow(s, r, hb, n)
        g2 = hb.read(n)
        for k in HIT_KEYS:
            assert np.array_equal(g2[k], want2[k]), f"{k}: the scene must trace its new tree"
        assert not np.array_equal(g2["t"], g["t"]), "the move must change the hits"
        other.build(30, 1)
        with pytest.raises(rtx.RtxError, match="not taken from this scene"):
            rtx.trace(other, r, hb, n, bvh=bvh)
    finally:
        bvh.destroy()
        for x in (r, hb, s, other):
            x.destroy()


def test_stage_tracer_overloads_and_builder_log(tmp_path):
    import rtx
    from rtx import scenes
    from test_gpu_stage_api import write_scene
    assert BIN.exists(), "stage_tracer not built (run __graft_entry__.build())"
    sd = scenes.config_c1()
    w, h = 128, 96
    write_scene(tmp_path, sd, w, h, 1, 30, 1, 1)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(BIN), str(tmp_path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    log = (out / "bvh.log").read_text()
    assert log.count("-----BVH construction times:") == 2, log
    for line in ("BBox builder time:", "Morton encoder time:", "Morton sorter:", "Total:"):
        assert line in log, f"{line} missing from the builder's log:\n{log}"
    rd = {k: np.fromfile(out / f"{k}.bin", dt) for k, dt in
          (("first_t", np.float32), ("first_id", np.int32), ("copy_t", np.float32), ("copy_id", np.int32),
           ("copy_flags", np.uint32), ("moved_t", np.float32), ("moved_id", np.int32), ("moved_flags", np.uint32))}
    assert np.array_equal(rd["copy_t"], rd["first_t"]) and np.array_equal(rd["copy_id"], rd["first_id"])
    assert not np.array_equal(rd["moved_t"], rd["first_t"])
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], w / h)
    oo, od, _, _ = O.raygen_primary(O.camera_from(cam), w, h)
    ob = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    lights = O.lights_from(sd.lights)
    want = O.trace(ob, sd.nrm, sd.uv, oo, od)
    want["flags"] = O.trace_shadow(ob, lights, want["flags"], want["p4"])
    assert np.array_equal(rd["copy_t"], want["t"]) and np.array_equal(rd["copy_id"], want["id"])
    assert np.array_equal(rd["copy_flags"], want["flags"])
    moved = sd.pos + np.float32([1.0, 0.0, 0.0])
    ob2 = O.Build(moved, ob.idx.reshape(-1, 3), ob.matmap, 30, 1)
    want2 = O.trace(ob2, sd.nrm, sd.uv, oo, od)
    want2["flags"] = O.trace_shadow(ob2, lights, want2["flags"], want2["p4"])
    assert np.array_equal(rd["moved_id"], want2["id"]) and np.array_equal(rd["moved_flags"], want2["flags"])


TUNE_KEYS = {  # key: (values accepted, values refused)
    "trace_variant": ([0, 4], [-1, 5]),
    "trace_block": ([64, 128, 256], [32, 512]),
    "xcds": ([1, 64], [0, 65]),
    "refill": ([1, 64], [0, 65]),
    "coop_leaf": ([0, 1 << 20], [-1, (1 << 20) + 1]),
    "side_stream": ([0, 1], [2]),
    "inst_lane_roots": ([0, 2], [3]),
    "wf_sched": ([0, 1], [2]),
    "wf_grain": ([1, 64], [0, 65]),
    "wf_share": ([0, 1, 100], [-1, 101]),
    "sort_bits": ([4, 8], [6, 16]),
    "sort_items": ([16, 32], [8, 64]),
    "build_grid": ([1, 16], [0, 17]),
    "shadow_order": ([0, 1], [2]),
    "stack_lds": ([8, 16, 32], [4, 24, 64]),
    "srv_step_pixels": ([0, 2_500_000], [-1]),
}


def test_tune_keys_round_trip_and_refuse_bad_values(ctx):
    # every autotuner key takes the values of its documented range, reads them
    # back, and refuses values outside it without changing the setting
    import rtx
    for key, (good, bad) in TUNE_KEYS.items():
        old = ctx.tuned(key)
        assert old >= 0, f"{key}: unknown to rtx_tune_get"
        try:
            for v in good:
                ctx.tune(key, v)
                assert ctx.tuned(key) == v, key
            for v in bad:
                with pytest.raises(rtx.RtxError, match="rtx_tune_set"):
                    ctx.tune(key, v)
                assert ctx.tuned(key) == good[-1], f"{key}: a refused value changed the setting"
        finally:
            ctx.tune(key, old)
    with pytest.raises(rtx.RtxError, match="unknown key"):
        ctx.tune("no_such_key", 1)
    assert ctx.tuned("no_such_key") == -1
