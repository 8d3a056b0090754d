"""The traversal stacks on the host: tests/cpp/stack_test.cpp compiled with
g++ against the header the kernels use (tesina-rt_amd/csrc/stack.hpp), under
AddressSanitizer and UBSan.  The ring stacks (8 and 16 LDS entries, the rest
in private memory) must return the whole-LDS stack's entry on every pop of
every push/pop sequence of up to 20 operations and of long random walks past
MAX_LEVELS, with no access outside buffers of the kernels' exact sizes."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_ring_stack_matches_whole_stack_in_bounds(tmp_path):
    exe = tmp_path / "stack_test"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wall", "-Werror", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", f"-I{ROOT / 'tesina-rt_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "stack_test.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
