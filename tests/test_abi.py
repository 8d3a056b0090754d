"""The C ABI library loads and exports every symbol include/rtx.h declares
(no GPU work is launched here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "rtx.h"
LIB = ROOT / "tesina-rt_amd" / "librtx.so"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rtx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "rtx_lbvh_build" in syms and "rtx_trace" in syms and len(syms) > 40


@pytest.mark.skipif(not LIB.exists(), reason="librtx.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"symbols declared in rtx.h but not exported: {missing}"


@pytest.mark.skipif(not LIB.exists(), reason="librtx.so not built")
def test_python_binding_covers_header():
    import rtx
    assert sorted(rtx.SYMBOLS) == declared_symbols()


@pytest.mark.skipif(not LIB.exists(), reason="librtx.so not built")
def test_host_only_entry_points():
    import numpy as np
    import rtx
    lib = rtx.lib()
    assert lib.rtx_abi_version() == 1
    # Camera::set is host code: usable without a GPU
    cam = rtx.camera_set((0, 0, -5), (0, 0, 0), (0, 1, 0), np.pi / 2, 2.0)
    assert abs(cam.dir[2] - 1.0) < 1e-7
    assert abs(cam.up[1] - 1.0) < 1e-6  # tan(pi/4) = 1
    assert abs(abs(cam.right[0]) - 2.0) < 1e-6  # aspect 2
    # error convention: negative status, message, no exit
    assert lib.rtx_scene_create(None, None) < 0
    assert b"null" in lib.rtx_last_error()


@pytest.mark.skipif(not LIB.exists(), reason="librtx.so not built")
def test_instance_entry_points_reject_null_handles():
    """The instance set's entry points check their arguments on the host:
    a negative status and a message naming the call, with no device work."""
    import rtx
    lib = rtx.lib()
    calls = [
        ("rtx_instances_create", lambda: lib.rtx_instances_create(None, None, None, None, 0, None)),
        ("rtx_instances_set_materials", lambda: lib.rtx_instances_set_materials(None, None)),
        ("rtx_instances_shading_scene", lambda: lib.rtx_instances_shading_scene(None, None)),
        ("rtx_instances_set_transforms", lambda: lib.rtx_instances_set_transforms(None, None, None)),
        ("rtx_trace_instances", lambda: lib.rtx_trace_instances(None, None, None, 0, None)),
        ("rtx_trace_shadow_instances", lambda: lib.rtx_trace_shadow_instances(None, None, None, 0, None)),
        ("rtx_render_instances", lambda: lib.rtx_render_instances(None, None, None, None, None)),
    ]
    for name, call in calls:
        assert call() < 0, name
        assert name.encode() in lib.rtx_last_error(), (name, lib.rtx_last_error())
