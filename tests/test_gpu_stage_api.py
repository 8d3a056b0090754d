"""The C++ stage classes (include/rt/stages.hpp) drive a frame the way the
reference's apps drive Renderer; the framebuffer must equal the oracle's.
Runs tests/cpp/stage_render.cpp (built by build()) against librtx.so."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tesina-rt_amd" / "bin" / "stage_render"


def write_scene(d: Path, sd, w, h, bounces, depth, min_leaf, frames):
    sd.pos.astype(np.float32).tofile(d / "pos.f32")
    sd.nrm.astype(np.float32).tofile(d / "nrm.f32")
    sd.uv.astype(np.float32).tofile(d / "uv.f32")
    sd.idx.astype(np.uint32).tofile(d / "idx.u32")
    sd.matmap.astype(np.int32).tofile(d / "matmap.i32")
    sd.materials.tofile(d / "mats.bin")
    (d / "lights.bin").write_bytes(bytes(sd.lights))
    c = sd.camera
    (d / "view.txt").write_text(" ".join(repr(float(x)) for x in (*c["pos"], *c["look_at"], *c["up"],
                                                                     np.float32(c["fov"]))))
    (d / "params.txt").write_text(f"{w} {h} {bounces} {depth} {min_leaf} {frames}")


@pytest.mark.parametrize("frames", [1, 2])
def test_cpp_renderer_matches_oracle(tmp_path, frames):
    import rtx
    from rtx import scenes
    assert BIN.exists(), "stage_render not built (run __graft_entry__.build())"
    sd = scenes.config_c1()
    w, h, bounces, depth = 160, 120, 2, 30
    write_scene(tmp_path, sd, w, h, bounces, depth, 1, frames)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(BIN), str(tmp_path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    acc = np.fromfile(out / "accum.u32", np.uint32).reshape(-1, 3)
    rays, sec, nodes, _ = (float(x) for x in (out / "stats.txt").read_text().split())
    cam = rtx.camera_set(sd.camera["pos"], sd.camera["look_at"], sd.camera["up"], sd.camera["fov"], w / h)
    b = O.Build(sd.pos, sd.idx, sd.matmap, depth, 1)
    for _ in range(frames - 1):  # every frame rebuilds from the permuted buffers
        b = O.Build(sd.pos, b.idx.reshape(-1, 3), b.matmap, depth, 1)
    # the reference Renderer's bundles hold 3 tiles (src/rt/renderer.cpp:376)
    oacc, st, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h,
                               bounces, cap=3 * w * h)
    assert int(rays) == orays and int(sec) == st["secondary"] and int(nodes) == b.node_count
    assert np.array_equal(acc, oacc)
    rgba = np.fromfile(out / "rgba.u8", np.uint8).reshape(-1, 4)
    assert np.array_equal(rgba, O.resolve(oacc))
