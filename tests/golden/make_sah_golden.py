"""Writes the expected outputs of the SAH builder fixtures by running the
reference's own CPU SAH BVH (oracle/_ref/sahdump, built from the reference
tree by oracle/Makefile.ref).  Run from the repository root:

    make -C oracle -f Makefile.ref && python tests/golden/make_sah_golden.py
"""
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent / "tesina-rt_amd"))
from sah_cases import CASES  # noqa: E402

DUMP = HERE.parent.parent / "oracle" / "_ref" / "sahdump"


def run_reference(pos, idx):
    """(nodes: lo[3] hi[3] a b parent leaf, order) from the reference builder."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
    with tempfile.TemporaryDirectory() as d:
        fi, fo = Path(d) / "in.bin", Path(d) / "out.bin"
        with open(fi, "wb") as f:
            np.array([pos.shape[0], idx.shape[0]], np.uint32).tofile(f)
            pos.tofile(f)
            idx.tofile(f)
        subprocess.run([str(DUMP), str(fi), str(fo)], check=True, timeout=600)
        raw = np.fromfile(fo, np.uint8)
    N = int(raw[:4].view(np.uint32)[0])
    rec = raw[4:4 + 40 * N].view(np.uint32).reshape(N, 10)
    box = rec[:, :6].copy().view(np.float32)
    words = rec[:, 6:].copy()
    order = raw[4 + 40 * N:].view(np.uint32).copy()
    return box, words, order


def main():
    out = HERE / "sah"
    out.mkdir(exist_ok=True)
    for name, make in CASES.items():
        pos, idx = make()
        box, words, order = run_reference(pos, idx)
        np.savez_compressed(out / f"{name}.npz", pos=np.float32(pos), idx=np.uint32(idx), box=box, words=words,
                            order=order)
        print(name, len(idx), box.shape[0])


if __name__ == "__main__":
    main()
