"""Instanced scenes for the multi-root traversal tests: sphere meshes, rigid
transforms with scales, and rays aimed into the scene."""
import numpy as np

from rtx import scenes


def rot(axis, ang):
    c, s = np.cos(ang), np.sin(ang)
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + s * K + (1 - c) * K @ K


def xform(R=np.eye(3), scale=1.0, t=(0.0, 0.0, 0.0)):
    """Object-to-world matrix (row-major 4x4, float32) and its inverse."""
    M = np.eye(4)
    M[:3, :3] = scale * R
    M[:3, 3] = t
    return M.astype(np.float32), np.linalg.inv(M).astype(np.float32)


def sphere_mesh(slices=24, stacks=16, r=1.0):
    p, n, uv, _ = scenes.uv_sphere(slices, stacks, (0.0, 0.0, 0.0), r)
    idx = scenes.grid_indices(stacks + 1, slices + 1, 0).astype(np.uint32)
    return p.astype(np.float32), n.astype(np.float32), uv.astype(np.float32), idx


def two_meshes():
    return [sphere_mesh(), sphere_mesh(16, 12, 0.7)]


def rays_at(seed, n, centre=(0.0, 0.0, 0.0), spread=6.0):
    """Rays from behind z = -spread toward random points of a 6-unit cube."""
    rng = np.random.default_rng(seed)
    o = np.asarray(centre) + rng.normal(size=(n, 3)) * 0.3 + np.array([0.0, 0.0, -spread])
    tgt = np.asarray(centre) + rng.uniform(-3, 3, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ori4 = np.zeros((n, 4), np.float32)
    dir4 = np.zeros((n, 4), np.float32)
    ori4[:, :3] = o
    ori4[:, 3] = 0.0001
    dir4[:, :3] = d
    dir4[:, 3] = 1e36
    return ori4, dir4
