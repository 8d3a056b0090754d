"""Triangle sets for the SAH builder's parity tests."""
import numpy as np


def _soup(seed, T, lo=-50.0, hi=50.0, size=2.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, (T, 3))
    pos = (c[:, None, :] + rng.normal(size=(T, 3, 3)) * size).reshape(-1, 3).astype(np.float32)
    return pos, np.arange(3 * T, dtype=np.uint32).reshape(T, 3)


def _negative():
    # every coordinate below zero: the bucket boxes' empty hi (FLT_MIN) stays
    # above all the geometry
    return _soup(3, 3000, -120.0, -20.0)


def _duplicates():
    pos, idx = _soup(4, 500)
    return pos, np.concatenate([idx, idx, idx[:100]])


def _flat():
    pos, idx = _soup(5, 2000)
    pos[:, 1] = 0.0
    return pos, idx


def _grid():
    n = 40
    xs = np.arange(n + 1, dtype=np.float32)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    pos = np.stack([X.ravel(), np.zeros(X.size, np.float32), Z.ravel()], 1)
    i = np.arange(n)
    a = (i[:, None] * (n + 1) + i[None, :]).ravel()
    idx = np.concatenate([np.stack([a, a + 1, a + n + 1], 1), np.stack([a + 1, a + n + 2, a + n + 1], 1)])
    return pos, idx.astype(np.uint32)


def deep():
    """Triangles at x = 3^k: the SAH splits off the far one at each level, a
    tree of 36 levels (beyond the traversal stack's 32)."""
    T = 70
    x = np.float32(3.0) ** np.arange(T, dtype=np.float32)
    c = np.stack([x, np.zeros(T), np.zeros(T)], 1).astype(np.float32)
    s = (x * 0.01).astype(np.float32)[:, None, None]
    pos = (c[:, None, :] + np.float32([[0, 0, 0], [0.3, 1, 0], [0, 0.2, 1]])[None] * s).reshape(-1, 3)
    return pos.astype(np.float32), np.arange(3 * T, dtype=np.uint32).reshape(T, 3)


def terrain_scene():
    """A small shaded scene: a 40 x 40 heightfield (mirror-free) and a
    reflective sphere; returns (pos, nrm, uv, idx, matmap)."""
    from rtx import scenes
    n = 41
    rng = np.random.default_rng(11)
    hts = rng.uniform(0, 1, (n, n))
    p0, n0, u0, i0 = scenes.heightfield(n, 4.0, hts * 0.3)
    sp, sn, su, si = scenes.uv_sphere(30, 30, (0.0, 1.2, 0.0), 1.0, base=p0.shape[0])
    pos = np.concatenate([p0, sp]).astype(np.float32)
    nrm = np.concatenate([n0, sn]).astype(np.float32)
    uv = np.concatenate([u0, su]).astype(np.float32)
    idx = np.concatenate([i0, si]).astype(np.uint32)
    mm = np.concatenate([np.zeros(len(i0), np.int32), np.ones(len(si), np.int32)])
    return pos, nrm, uv, idx, mm


CASES = {
    "deep": deep,
    "terrain": lambda: terrain_scene()[0:4:3],
    "soup_1": lambda: _soup(1, 1),
    "soup_2": lambda: _soup(2, 2),
    "soup_3": lambda: _soup(6, 3),
    "soup_17": lambda: _soup(7, 17),
    "soup_5000": lambda: _soup(8, 5000),
    "negative": _negative,
    "duplicates": _duplicates,
    "flat": _flat,
    "grid": _grid,
}
