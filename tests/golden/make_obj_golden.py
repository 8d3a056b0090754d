"""Writes the expected outputs of the OBJ loader fixtures (tests/golden/obj)
by running the reference's own loader: oracle/_ref/objdump, built from the
reference tree by oracle/Makefile.ref.  Run from the repository root:

    make -C oracle -f Makefile.ref && python tests/golden/make_obj_golden.py
"""
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from objdump_format import parse, to_npz  # noqa: E402

DUMP = HERE.parent.parent / "oracle" / "_ref" / "objdump"


def main():
    for obj in sorted((HERE / "obj").glob("*.obj")):
        for rebuild in (0, 1):
            r = subprocess.run([str(DUMP), str(obj), str(rebuild)], capture_output=True, text=True, check=True)
            out = obj.with_name(f"{obj.stem}.ref{rebuild}.npz")
            to_npz(parse(r.stdout), out)
            print(out.name, out.stat().st_size)


if __name__ == "__main__":
    main()
