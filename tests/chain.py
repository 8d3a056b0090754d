"""Scene helper shared by the oracle and GPU tests of the traversal's MAX_LEVELS rule."""
import numpy as np


def chain_scene():
    """Triangles whose Morton codes (depth 32) are 0, 1, 2, 4, ..., 2^31, four
    of each: a chain of 32 inner levels.  Two tiny anchors fix the global box
    to [0, 1]^3.  Code 2^j is one bit of one axis (bit i of a code is bit
    (i + 1) / 3 of axis (2 + i) % 3 at depth 32); each of its triangles is a
    sliver along that axis, centred in the bit's cell."""
    P, I = [], []

    def tri(v):
        base = len(P)
        P.extend([tuple(map(float, x)) for x in v])
        I.append([base, base + 1, base + 2])

    e = 1e-4
    for _ in range(2):
        tri([(0, 0, 0), (e, 0, 0), (0, e, 0)])
        tri([(1, 1, 1), (1 - e, 1, 1), (1, 1 - e, 1)])
    cell = 1.0 / 2048
    for j in range(32):
        a, b = (2 + j) % 3, (j + 1) // 3
        c = np.full(3, 0.5 * cell)
        c[a] = (2 ** b + 0.5) * cell
        L, h = 0.9 * min(c[a], 1 - c[a]), 0.45 * cell
        ea, e2 = np.eye(3)[a], np.eye(3)[(a + 1) % 3]
        for _ in range(4):
            tri([c - L * ea - h * e2, c + L * ea - h * e2, c + h * e2])
    return np.array(P, np.float32), np.array(I, np.uint32)
