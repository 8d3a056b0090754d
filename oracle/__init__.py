"""Python wrapper of the rt-seq oracle (liboracle.so, built from rtseq.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

NODE_DTYPE = np.dtype([("hi", "<f4", 4), ("lo", "<f4", 4), ("a", "<u4"), ("b", "<u4"), ("parent", "<u4"),
                       ("split_axis", "i1"), ("leaf", "i1"), ("pad", "u1", 2)])
LIGHT_DTYPE = np.dtype([("type", "<i4"), ("pos", "<f4", 3), ("dir", "<f4", 3), ("color", "<f4", 3),
                        ("angle", "<f4"), ("radius", "<f4")])
LIGHTS_DTYPE = np.dtype([("ambient", "<f4", 3), ("count", "<i4"), ("l", LIGHT_DTYPE, 8)])
CAMERA_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("right", "<f4", 3), ("up", "<f4", 3)])
STATS_DTYPE = np.dtype([("primary", "<u8"), ("secondary", "<u8"), ("shadow", "<u8"), ("bounces_run", "<u8"),
                        ("per_bounce", "<u8", 16)])

SPOT, DIR, POINT = 0, 1, 2

_lib = None


REF_TREE = Path("/root/reference")


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    # the reference's own OBJ loader, when its tree is here (oracle/_ref)
    if REF_TREE.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "-f", "Makefile.ref", f"REF={REF_TREE}"], check=True,
                       capture_output=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.orc_morton_one.restype = C.c_uint64
        _lib.orc_raygen_secondary.restype = C.c_uint64
        _lib.orc_render.restype = C.c_uint64
        _lib.orc_render_tex.restype = C.c_uint64
        _lib.orc_render_multi.restype = C.c_uint64
        _lib.orc_hit_t_f64.restype = C.c_double
    return _lib


def P(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def pos4(pos3: np.ndarray) -> np.ndarray:
    p = np.zeros((pos3.shape[0], 4), np.float32)
    p[:, :3] = pos3
    return p


def lights_from(rtx_lights) -> np.ndarray:
    """Convert an rtx.Lights ctypes struct (same layout) to the oracle's array."""
    a = np.frombuffer(bytes(rtx_lights), dtype=LIGHTS_DTYPE).copy()
    return a


def camera_from(rtx_cam) -> np.ndarray:
    return np.frombuffer(bytes(rtx_cam), dtype=CAMERA_DTYPE).copy()


def global_bbox(pos3, idx):
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    p4 = pos4(pos3)
    i = np.ascontiguousarray(idx, np.uint32).reshape(-1)
    lib().orc_global_bbox(P(p4), P(i), C.c_uint32(i.size // 3), P(lo), P(hi))
    return lo, hi


def morton_one(lo, hi, glo, ghi, depth: int) -> int:
    f = [np.ascontiguousarray(x, np.float32) for x in (lo, hi, glo, ghi)]
    return int(lib().orc_morton_one(P(f[0]), P(f[1]), P(f[2]), P(f[3]), C.c_int(depth)))


def morton(pos3, idx, depth: int) -> np.ndarray:
    p4 = pos4(pos3)
    i = np.ascontiguousarray(idx, np.uint32).reshape(-1)
    T = i.size // 3
    out = np.zeros(T, np.uint64)
    lib().orc_morton(P(p4), P(i), C.c_uint32(T), C.c_int(depth), P(out))
    return out


def sort_perm(codes: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(codes, np.uint64)
    out = np.zeros(c.size, np.uint32)
    lib().orc_sort(P(c), C.c_uint32(c.size), P(out))
    return out


def bitonic_perm(codes: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(codes, np.uint64)
    out = np.zeros(c.size, np.uint32)
    lib().orc_bitonic_emulate(P(c), C.c_uint32(c.size), P(out))
    return out


class Build:
    """Result of orc_lbvh_build; idx/matmap are the permuted copies."""

    def __init__(self, pos3, idx, matmap=None, depth=30, min_leaf=1):
        self.pos4 = pos4(np.asarray(pos3, np.float32))
        self.idx = np.ascontiguousarray(np.asarray(idx, np.uint32).reshape(-1)).copy()
        T = self.idx.size // 3
        self.T = T
        self.matmap = (np.zeros(T, np.int32) if matmap is None else np.ascontiguousarray(matmap, np.int32).copy())
        self.codes = np.zeros(T, np.uint64)
        self.perm = np.zeros(T, np.uint32)
        self.nodes = np.zeros(2 * T + 1, NODE_DTYPE)
        n = C.c_uint32()
        self.level_counts = np.zeros(64, np.uint32)
        self.gbox = np.zeros(6, np.float32)
        rc = lib().orc_lbvh_build(P(self.pos4), P(self.idx), P(self.matmap), C.c_uint32(T), C.c_int(depth),
                                  C.c_int(min_leaf), P(self.codes), P(self.perm), P(self.nodes),
                                  C.c_uint32(self.nodes.size), C.byref(n), P(self.level_counts), P(self.gbox))
        if rc != 0:
            raise RuntimeError(f"orc_lbvh_build failed ({rc})")
        self.nodes = self.nodes[: n.value].copy()
        self.node_count = int(n.value)
        self.depth, self.min_leaf = depth, min_leaf

    def refit(self, pos3):
        self.pos4 = pos4(np.asarray(pos3, np.float32))
        rc = lib().orc_lbvh_refit(P(self.pos4), P(self.idx), C.c_uint32(self.T), P(self.nodes),
                                  C.c_uint32(self.node_count), C.c_int(self.depth), P(self.level_counts))
        if rc != 0:
            raise RuntimeError("orc_lbvh_refit failed")


class stackless:
    """Context in which every trace of the oracle walks the tree without a
    stack (src/kernel/trace-bvh-no-stack.cl, shadow-trace-bvh-no-stack.cl)."""

    def __enter__(self):
        lib().orc_set_traversal(1)
        return self

    def __exit__(self, *a):
        lib().orc_set_traversal(0)


class Tree:
    """A tree given from outside, e.g. the reference's CPU SAH BVH
    (src/rt/bvh.cpp): node boxes (lo[3], hi[3]), words (a, b, parent, leaf)
    and the triangle order; the index buffer and material map are permuted
    into that order (BVH::construct_and_map), as Build's are.  The oracle's
    traversal functions take it where they take a Build."""

    def __init__(self, pos3, idx, matmap, box, words, order):
        self.pos4 = pos4(np.asarray(pos3, np.float32))
        order = np.asarray(order, np.int64)
        self.idx = np.ascontiguousarray(np.asarray(idx, np.uint32).reshape(-1, 3)[order].reshape(-1))
        self.T = len(order)
        mm = np.zeros(self.T, np.int32) if matmap is None else np.asarray(matmap, np.int32)
        self.matmap = np.ascontiguousarray(mm[order])
        N = len(box)
        self.nodes = np.zeros(N, NODE_DTYPE)
        self.nodes["lo"][:, :3] = box[:, :3]
        self.nodes["hi"][:, :3] = box[:, 3:6]
        self.nodes["a"], self.nodes["b"] = words[:, 0], words[:, 1]
        self.nodes["parent"] = words[:, 2]
        self.nodes["leaf"] = (words[:, 3] != 0).astype(np.int8)
        self.node_count = N


def raygen_primary(cam: np.ndarray, w: int, h: int, quad: int = 32, psamples=None, offset: int = 0, count=None):
    ps = np.ascontiguousarray(np.array([0.0, 0.0, 1.0], np.float32) if psamples is None else psamples, np.float32)
    count = w * h - offset if count is None else count
    o = np.zeros((count, 4), np.float32)
    d = np.zeros((count, 4), np.float32)
    px = np.zeros(count, np.int32)
    c = np.zeros(count, np.float32)
    lib().orc_raygen_primary(P(cam), C.c_int(w), C.c_int(h), C.c_int(quad), P(ps), C.c_uint64(offset),
                             C.c_uint64(count), P(o), P(d), P(px), P(c))
    return o, d, px, c


def empty_hits(n):
    return dict(flags=np.zeros(n, np.uint32), t=np.zeros(n, np.float32), id=np.zeros(n, np.int32),
                bary=np.zeros((n, 2), np.float32), n4=np.zeros((n, 4), np.float32), p4=np.zeros((n, 4), np.float32),
                uv2=np.zeros((n, 2), np.float32))


def trace(b: Build, nrm3, uv2, ori4, dir4, count_visits=False):
    n = ori4.shape[0]
    h = empty_hits(n)
    nrm4 = pos4(np.asarray(nrm3, np.float32))
    uv = np.ascontiguousarray(uv2, np.float32)
    visits = np.zeros(2, np.uint64)
    lib().orc_trace(P(b.pos4), P(nrm4), P(uv), P(b.idx), P(b.nodes), P(np.ascontiguousarray(ori4, np.float32)),
                    P(np.ascontiguousarray(dir4, np.float32)), C.c_uint64(n), P(h["flags"]), P(h["t"]), P(h["id"]),
                    P(h["bary"]), P(h["n4"]), P(h["p4"]), P(h["uv2"]), P(visits) if count_visits else None)
    if count_visits:
        h["visits"] = int(visits[0])
        h["tri_tests"] = int(visits[1])
    return h


def trace_brute(b: Build, ori4, dir4):
    n = ori4.shape[0]
    ids = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    lib().orc_trace_brute(P(b.pos4), P(b.idx), C.c_uint32(b.T), P(np.ascontiguousarray(ori4, np.float32)),
                          P(np.ascontiguousarray(dir4, np.float32)), C.c_uint64(n), P(ids), P(t))
    return ids, t


def trace_shadow(b: Build, lights: np.ndarray, flags, p4, visits=None):
    """visits: optional uint64[2] array accumulating (nodes, triangle tests)."""
    f = np.ascontiguousarray(flags, np.uint32).copy()
    lib().orc_trace_shadow(P(b.pos4), P(b.idx), P(b.nodes), P(lights), C.c_uint64(f.size), P(f),
                           P(np.ascontiguousarray(p4, np.float32)), P(visits))
    return f


def trace_shadow_brute(b: Build, lights: np.ndarray, flags, p4):
    f = np.ascontiguousarray(flags, np.uint32).copy()
    lib().orc_trace_shadow_brute(P(b.pos4), P(b.idx), C.c_uint32(b.T), P(lights), C.c_uint64(f.size), P(f),
                                 P(np.ascontiguousarray(p4, np.float32)))
    return f


def raygen_secondary(mats: np.ndarray, matmap, ori4, dir4, pixel, contrib, hits: dict, cap: int):
    n = ori4.shape[0]
    o = np.zeros((max(cap, 1), 4), np.float32)
    d = np.zeros((max(cap, 1), 4), np.float32)
    px = np.zeros(max(cap, 1), np.int32)
    c = np.zeros(max(cap, 1), np.float32)
    total = np.zeros(1, np.uint64)
    m = lib().orc_raygen_secondary(P(mats), P(np.ascontiguousarray(matmap, np.int32)),
                                   P(np.ascontiguousarray(ori4, np.float32)), P(np.ascontiguousarray(dir4, np.float32)),
                                   P(np.ascontiguousarray(pixel, np.int32)), P(np.ascontiguousarray(contrib, np.float32)),
                                   P(hits["flags"]), P(hits["id"]), P(hits["n4"]), P(hits["p4"]), C.c_uint64(n),
                                   C.c_uint64(cap), P(o), P(d), P(px), P(c), P(total))
    m = int(m)
    return o[:m], d[:m], px[:m], c[:m], int(total[0])


class _TexSetC(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("tab", C.c_void_p), ("count", C.c_int32), ("cube", C.c_int32)]


class TexSet:
    """Textures of the shading stage: material textures (h x w x 4 uint8
    arrays, texel (x, y) at [y, x]) and optionally the six cubemap faces
    (+x, -x, +y, -y, +z, -z)."""

    def __init__(self, textures=(), cubemap=None):
        imgs = [np.ascontiguousarray(t, np.uint8) for t in textures]
        self.count = len(imgs)
        self.cube = 0
        if cubemap is not None:
            faces = [np.ascontiguousarray(f, np.uint8) for f in cubemap]
            assert len(faces) == 6
            imgs += faces
            self.cube = 1
        tab, off = [], 0
        for im in imgs:
            h, w = im.shape[:2]
            tab += [off, w, h]
            off += w * h
        self.rgba = np.ascontiguousarray(np.concatenate([im.reshape(-1, 4) for im in imgs]) if imgs
                                         else np.zeros((1, 4), np.uint8))
        self.tab = np.asarray(tab or [0, 1, 1], np.int64)
        self.c = _TexSetC(self.rgba.ctypes.data, self.tab.ctypes.data, self.count, self.cube)


def sample_texture(ts: TexSet, entry: int, st, repeat: bool = True) -> np.ndarray:
    st = np.ascontiguousarray(st, np.float32).reshape(-1, 2)
    out = np.zeros((st.shape[0], 3), np.float32)
    lib().orc_sample_texture(C.byref(ts.c), C.c_int32(entry), P(st), C.c_uint64(st.shape[0]), C.c_int32(int(repeat)),
                             P(out))
    return out


def sample_cubemap(ts: TexSet, dir4) -> np.ndarray:
    d = np.ascontiguousarray(dir4, np.float32)
    out = np.zeros((d.shape[0], 3), np.float32)
    lib().orc_sample_cubemap(C.byref(ts.c), P(d), C.c_uint64(d.shape[0]), P(out))
    return out


def shade(mats, matmap, lights, dir4, pixel, contrib, hits, accum, tex: TexSet | None = None):
    n = dir4.shape[0]
    uv2 = hits.get("uv2")
    lib().orc_shade_tex(P(mats), P(np.ascontiguousarray(matmap, np.int32)), P(lights),
                        P(np.ascontiguousarray(dir4, np.float32)), P(np.ascontiguousarray(pixel, np.int32)),
                        P(np.ascontiguousarray(contrib, np.float32)), P(hits["flags"]), P(hits["id"]), P(hits["n4"]),
                        P(hits["p4"]), P(None if uv2 is None else np.ascontiguousarray(uv2, np.float32)),
                        C.c_uint64(n), C.byref(tex.c) if tex is not None else None, P(accum))
    return accum


def resolve(accum: np.ndarray) -> np.ndarray:
    out = np.zeros((accum.size // 3, 4), np.uint8)
    lib().orc_resolve(P(np.ascontiguousarray(accum, np.uint32)), C.c_uint64(accum.size // 3), P(out))
    return out


def render(b: Build, nrm3, uv2, mats, lights, cam, w, h, max_bounces, quad=32, cap=0, tex: TexSet | None = None):
    accum = np.zeros((w * h, 3), np.uint32)
    st = np.zeros(1, STATS_DTYPE)
    nrm4 = pos4(np.asarray(nrm3, np.float32))
    uv = np.ascontiguousarray(uv2, np.float32)
    cap = cap or (w * h) << max_bounces
    rays = lib().orc_render_tex(P(b.pos4), P(nrm4), P(uv), P(b.idx), P(b.matmap), P(b.nodes), P(mats), P(lights),
                                P(cam), C.c_int(w), C.c_int(h), C.c_int(quad), C.c_int(max_bounces), C.c_uint64(cap),
                                C.byref(tex.c) if tex is not None else None, P(accum), P(st))
    return accum, st[0], int(rays)


def hit_t_f64(b: Build, ori4, dir4, ids, t32):
    n = ori4.shape[0]
    t64 = np.zeros(n, np.float64)
    worst = lib().orc_hit_t_f64(P(b.pos4), P(b.idx), P(np.ascontiguousarray(ori4, np.float32)),
                                P(np.ascontiguousarray(dir4, np.float32)), P(np.ascontiguousarray(ids, np.int32)),
                                P(np.ascontiguousarray(t32, np.float32)), C.c_uint64(n), P(t64))
    return float(worst), t64


# ---------------------------------------------------------------- multi-root
ROOT_DTYPE = np.dtype([("node", "<i4"), ("tr", "<f4", 16), ("tr_inv", "<f4", 16)])
assert ROOT_DTYPE.itemsize == 132


class MultiScene:
    """Meshes in one shared node array, index buffer and vertex buffer, as
    Scene::create_bvhs lays them out (src/rt/scene.cpp:644-700): every new
    mesh, in the order the instances first use it, gets the next node, triangle
    and vertex offsets; roots give a mesh's first node and the instance's
    transforms (object to world `tr`, row-major 4x4, and its inverse)."""

    def __init__(self, builds, nrms, uvs, instances):
        # builds: per mesh O.Build; instances: [(mesh, tr 4x4, tr_inv 4x4)]
        order = []
        for m, _, _ in instances:
            if m not in order:
                order.append(m)
        node_off, tri_off, vtx_off = {}, {}, {}
        nodes, idx, pos, nrm, uv = [], [], [], [], []
        no = to = vo = 0
        for m in order:
            b = builds[m]
            node_off[m], tri_off[m], vtx_off[m] = no, to, vo
            nd = b.nodes.copy()
            leaf = nd["leaf"] != 0
            nd["a"] = np.where(leaf, nd["a"] + to, nd["a"] + no)
            nd["b"] = np.where(leaf, nd["b"] + to, nd["b"] + no)
            nd["parent"] = nd["parent"] + no
            nodes.append(nd)
            idx.append(b.idx + vo)
            pos.append(b.pos4)
            nrm.append(pos4(np.asarray(nrms[m], np.float32)))
            uv.append(np.ascontiguousarray(uvs[m], np.float32))
            no += len(nd)
            to += b.T
            vo += b.pos4.shape[0]
        self.nodes = np.ascontiguousarray(np.concatenate(nodes))
        self.idx = np.ascontiguousarray(np.concatenate(idx), np.uint32)
        self.pos4 = np.ascontiguousarray(np.concatenate(pos), np.float32)
        self.nrm4 = np.ascontiguousarray(np.concatenate(nrm), np.float32)
        self.uv = np.ascontiguousarray(np.concatenate(uv), np.float32)
        self.tri_base = tri_off
        self.inst_mesh = [m for m, _, _ in instances]
        self.roots = np.zeros(len(instances), ROOT_DTYPE)
        for i, (m, tr, tri) in enumerate(instances):
            self.roots[i]["node"] = node_off[m]
            self.roots[i]["tr"] = np.asarray(tr, np.float32).reshape(16)
            self.roots[i]["tr_inv"] = np.asarray(tri, np.float32).reshape(16)


def trace_multi(ms: MultiScene, ori4, dir4):
    n = ori4.shape[0]
    h = empty_hits(n)
    lib().orc_trace_multi(P(ms.pos4), P(ms.nrm4), P(ms.uv), P(ms.idx), P(ms.nodes), P(ms.roots),
                          C.c_int(len(ms.roots)), P(np.ascontiguousarray(ori4, np.float32)),
                          P(np.ascontiguousarray(dir4, np.float32)), C.c_uint64(n), P(h["flags"]), P(h["t"]),
                          P(h["id"]), P(h["bary"]), P(h["n4"]), P(h["p4"]), P(h["uv2"]))
    return h


def render_multi(ms: MultiScene, inst_materials, lights, cam, w, h, max_bounces, quad=32, cap=0,
                 tex: TexSet | None = None):
    """A frame over the instances (orc_render_multi): inst_materials = one
    32-B material per instance; each mesh takes its first instance's
    (Scene::create_bvhs).  Returns (accum, stats, rays)."""
    raw = np.asarray(inst_materials, np.uint8).reshape(-1, 32)
    order = sorted(ms.tri_base, key=ms.tri_base.get)  # the meshes in order of first use
    starts = [ms.tri_base[m] for m in order] + [ms.idx.size // 3]
    mats = np.ascontiguousarray(np.concatenate([raw[ms.inst_mesh.index(m)] for m in order]))
    matmap = np.ascontiguousarray(np.concatenate([np.full(starts[k + 1] - starts[k], k, np.int32)
                                                  for k in range(len(order))]))
    accum = np.zeros((w * h, 3), np.uint32)
    st = np.zeros(1, STATS_DTYPE)
    cap = cap or (w * h) << max_bounces
    rays = lib().orc_render_multi(P(ms.pos4), P(ms.nrm4), P(ms.uv), P(ms.idx), P(ms.nodes), P(ms.roots),
                                  C.c_int(len(ms.roots)), P(matmap), P(mats), P(lights), P(cam), C.c_int(w),
                                  C.c_int(h), C.c_int(quad), C.c_int(max_bounces), C.c_uint64(cap),
                                  C.byref(tex.c) if tex is not None else None, P(accum), P(st))
    return accum, st[0], int(rays)


def trace_shadow_multi(ms: MultiScene, lights: np.ndarray, flags, p4):
    f = np.ascontiguousarray(flags, np.uint32).copy()
    lib().orc_trace_shadow_multi(P(ms.pos4), P(ms.idx), P(ms.nodes), P(ms.roots), C.c_int(len(ms.roots)),
                                 P(lights), C.c_uint64(f.size), P(f), P(np.ascontiguousarray(p4, np.float32)))
    return f
