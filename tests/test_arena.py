"""The wavefront renderer's device arena and admission queue, on the host:
tests/cpp/arena_test.cpp compiled with g++ against the header the library
uses (tesina-rt_amd/csrc/arena.hpp), run with a tiny forced budget."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_arena_and_admission(tmp_path):
    exe = tmp_path / "arena_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'tesina-rt_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "arena_test.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
