"""Build the gfx950 library librtx.so (HIP kernels + C ABI) in-tree.

    python tesina-rt_amd/build.py            # build librtx.so
    python tesina-rt_amd/build.py --clean

Every HIP source is compiled with hipcc for gfx950 only, with no FMA
contraction, correctly rounded f32 divide/sqrt, IEEE denormals and packed
FP32 instructions disabled; the objects are linked into one shared library
next to this file.  The C++ stage classes (rt/) are compiled into the same
library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
RTSRC = PKG / "rt"
OBJ = PKG / "_build"
LIB = PKG / "librtx.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{PKG}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources() -> list[Path]:
    hip = sorted(CSRC.glob("*.hip"))
    cpp = sorted(RTSRC.glob("*.cpp")) if RTSRC.exists() else []
    return hip + cpp


def _compile(src: Path) -> Path:
    OBJ.mkdir(exist_ok=True)
    obj = OBJ / (src.stem + ("_rt" if src.suffix == ".cpp" else "") + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_tools()
    if verbose:
        print(f"built {LIB}")
    return LIB


TOOLS = {"stage_render": ROOT / "tests" / "cpp" / "stage_render.cpp",
         "stage_objects": ROOT / "tests" / "cpp" / "stage_objects.cpp",
         "stage_boat": ROOT / "tests" / "cpp" / "stage_boat.cpp",
         "stage_tracer": ROOT / "tests" / "cpp" / "stage_tracer.cpp"}
BIN = PKG / "bin"


def build_tools() -> None:
    """C++ programs that drive the library through include/rt/stages.hpp
    (host code only: plain g++ against librtx.so)."""
    BIN.mkdir(exist_ok=True)
    for name, src in TOOLS.items():
        if not src.exists():
            continue
        cmd = ["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(src), "-o", str(BIN / name),
               f"-L{PKG}", "-lrtx", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"tool build failed: {name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def clean() -> None:
    if OBJ.exists():
        shutil.rmtree(OBJ)
    if BIN.exists():
        shutil.rmtree(BIN)
    if LIB.exists():
        LIB.unlink()


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    else:
        build(verbose=True)
