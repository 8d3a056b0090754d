"""The reference's boat scene (include/rt/test-params.hpp:10-50), built from
its OBJ models by the library's loader and the scene registry
(Scene::create_aggregate_mesh), rendered on the GPU from its three stats
views and compared with the oracle's frame bit for bit.  The water is
reflective and refractive (IOR 1.333), so every view runs secondary rays."""
from pathlib import Path

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
MODELS = Path(__file__).resolve().parent / "golden" / "obj"


@pytest.mark.parametrize("view", [0, 1, 2])
def test_boat_preset_frame(ctx, view):
    import rtx
    from rtx import scenes
    sd = scenes.preset_boat(MODELS, res=(192, 108), max_bounces=3, view=view)
    s = scenes.load(rtx.Scene(ctx), sd)
    cam = scenes.camera_of(sd)
    R = rtx.Renderer(ctx, sd.width, sd.height, sd.max_bounces)
    st = R.frame(s, cam, rebuild=True)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    oacc, ost, orays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam),
                                sd.width, sd.height, sd.max_bounces)
    assert (oacc.sum(1) > 0).mean() > 0.2, "the view should see the scene"
    assert st.primary_rays + st.secondary_rays == orays and st.secondary_rays > 0
    assert np.array_equal(R.fb.accum(), oacc)
    R.destroy()
    s.destroy()


@pytest.mark.parametrize("view", [0, 2])
def test_boat_through_cpp_stage_classes(tmp_path, view):
    """The same scene built in C++ as boat_set_scene does (load_obj_file_as_aggregate,
    add_object, create_aggregate_mesh) and rendered by the C++ Renderer."""
    import subprocess
    import rtx
    from rtx import scenes
    exe = Path(__file__).resolve().parent.parent / "tesina-rt_amd" / "bin" / "stage_boat"
    assert exe.exists(), "stage_boat not built (run __graft_entry__.build())"
    w, h = 192, 108
    r = subprocess.run([str(exe), str(MODELS), str(tmp_path), str(w), str(h), str(view)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    acc = np.fromfile(tmp_path / "accum.u32", np.uint32).reshape(-1, 3)
    sd = scenes.preset_boat(MODELS, res=(w, h), max_bounces=3, view=view)
    cam = scenes.camera_of(sd)
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    # the C++ Renderer's bundles hold 3 tiles (src/rt/renderer.cpp:376)
    oacc, _, _ = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam), w, h, 3,
                          cap=3 * w * h)
    assert (oacc.sum(1) > 0).mean() > 0.2
    assert np.array_equal(acc, oacc)
