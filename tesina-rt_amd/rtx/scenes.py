"""Procedural scenes of the benchmark configurations, generated on the host.

Geometry is computed in float64 and rounded to float32 once, so the GPU path
and the CPU oracle see identical inputs.  Randomness: 1024 independent
xoshiro256** streams, each seeded by splitmix64 from (seed + stream), stepped
together; value k of a draw comes from stream k % 1024, step k // 1024;
floats are (u64 >> 40) * 2^-24.

Workloads (seed = 0x5EED0000 + config id):
  C1  181x181 heightfield over [-5,5]^2 (U[0,1] heights, 5x5 box filter,
      64,800 tris) + 100x100 UV sphere (20,000 tris), 512x512, 1 point light,
      1 mirror bounce.
  C2  200 UV spheres of 50x50 (5,000 tris each, 1,000,000 tris), radii
      logU[0.05,1], centres U[-10,10]^3, vertices jittered by U[-1%,1%] of the
      radius; 30% refractive (ior 1.5), 30% mirror, 40% diffuse; 1920x1080,
      2 point lights, Whitted depth 3.
  C3  724x724 heightfield (1,045,458 tris) displaced by 8 travelling sine
      waves per frame; 1280x720, 1 light, depth 2.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import LIGHT_DIR, LIGHT_POINT, make_lights, make_materials

SEED_BASE = 0x5EED0000
_M64 = (1 << 64) - 1
_STREAMS = 1024


def _splitmix64(x: np.ndarray):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
    z = x.copy()
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x, z ^ (z >> np.uint64(31))


def _rotl(x, k):
    return (x << np.uint64(k)) | (x >> np.uint64(64 - k))


class Rng:
    """1024 xoshiro256** streams stepped in lockstep."""

    def __init__(self, seed: int):
        with np.errstate(over="ignore"):
            x = (np.uint64(seed) + np.arange(_STREAMS, dtype=np.uint64)) & np.uint64(_M64)
            st = []
            for _ in range(4):
                x, z = _splitmix64(x)
                st.append(z)
        self.s = st

    def _next(self) -> np.ndarray:
        s0, s1, s2, s3 = self.s
        with np.errstate(over="ignore"):
            res = _rotl(s1 * np.uint64(5), 7) * np.uint64(9)
            t = s1 << np.uint64(17)
            s2 = s2 ^ s0
            s3 = s3 ^ s1
            s1 = s1 ^ s2
            s0 = s0 ^ s3
            s2 = s2 ^ t
            s3 = _rotl(s3, 45)
        self.s = [s0, s1, s2, s3]
        return res

    def u64(self, n: int) -> np.ndarray:
        steps = (n + _STREAMS - 1) // _STREAMS
        out = np.empty((steps, _STREAMS), np.uint64)
        for i in range(steps):
            out[i] = self._next()
        return out.reshape(-1)[:n]

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.u64(n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
        return lo + (hi - lo) * u


@dataclass
class SceneData:
    pos: np.ndarray  # V x 3 float32
    nrm: np.ndarray  # V x 3 float32
    uv: np.ndarray  # V x 2 float32
    idx: np.ndarray  # T x 3 uint32
    matmap: np.ndarray  # T int32
    materials: np.ndarray  # M x 32 bytes
    lights: object  # rtx.Lights
    camera: dict  # pos, look_at, up, fov
    width: int
    height: int
    max_bounces: int
    waves: np.ndarray | None = None  # C3: n x 6 (a, kappa, dx, dz, omega, phi)
    meta: dict = field(default_factory=dict)

    @property
    def T(self):
        return self.idx.shape[0]


def _f32(a):
    # +0.0 canonicalizes -0.0 so no signed zero reaches the min/max unions
    return (np.asarray(a, np.float64) + 0.0).astype(np.float32)


def grid_indices(nx: int, nz: int, base: int = 0) -> np.ndarray:
    """Two triangles per grid cell of an (nx x nz)-vertex grid, row-major."""
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(nz - 1), indexing="ij")
    v00 = (i * nz + j).ravel()
    v01 = v00 + 1
    v10 = v00 + nz
    v11 = v10 + 1
    t = np.empty((v00.size * 2, 3), np.int64)
    t[0::2] = np.stack([v00, v10, v11], 1)
    t[1::2] = np.stack([v00, v11, v01], 1)
    return (t + base).astype(np.uint32)


def heightfield(n: int, extent: float, heights: np.ndarray):
    """n x n vertices over [-extent, extent]^2 with the given heights (n x n, float64)."""
    x = np.linspace(-extent, extent, n)
    X, Z = np.meshgrid(x, x, indexing="ij")
    pos = np.stack([X, heights, Z], -1).reshape(-1, 3)
    h = x[1] - x[0]
    gx = np.gradient(heights, h, axis=0)
    gz = np.gradient(heights, h, axis=1)
    nrm = np.stack([-gx, np.ones_like(gx), -gz], -1).reshape(-1, 3)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    U, W = np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n), indexing="ij")
    uv = np.stack([U, W], -1).reshape(-1, 2)
    return _f32(pos), _f32(nrm), _f32(uv), grid_indices(n, n)


def box_filter(a: np.ndarray, k: int = 5) -> np.ndarray:
    r = k // 2
    p = np.pad(a, r, mode="edge")
    out = np.zeros_like(a)
    for di in range(k):
        for dj in range(k):
            out += p[di:di + a.shape[0], dj:dj + a.shape[1]]
    return out / (k * k)


def uv_sphere(slices: int, stacks: int, center, radius: float, jitter: np.ndarray | None = None, base: int = 0):
    """(stacks+1) x (slices+1) vertices; 2 triangles per quad, so the pole
    quads carry one degenerate (zero-area) triangle each."""
    th = np.linspace(0.0, np.pi, stacks + 1)
    ph = np.linspace(0.0, 2 * np.pi, slices + 1)
    T_, P_ = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T_) * np.cos(P_), np.cos(T_), np.sin(T_) * np.sin(P_)], -1).reshape(-1, 3)
    pos = np.asarray(center, np.float64) + radius * n
    if jitter is not None:
        pos = pos + jitter.reshape(-1, 3)
    U, W = np.meshgrid(np.linspace(0, 1, stacks + 1), np.linspace(0, 1, slices + 1), indexing="ij")
    uv = np.stack([W, U], -1).reshape(-1, 2)
    idx = grid_indices(stacks + 1, slices + 1, base)
    return pos, n, uv, idx


def config_c1(seed: int = SEED_BASE + 1) -> SceneData:
    rng = Rng(seed)
    n = 181
    hts = box_filter(rng.uniform(n * n).reshape(n, n))
    p0, n0, u0, i0 = heightfield(n, 5.0, hts)
    sp, sn, su, si = uv_sphere(100, 100, (0.0, 1.5, 0.0), 1.0, base=p0.shape[0])
    pos = np.concatenate([p0, _f32(sp)])
    nrm = np.concatenate([n0, _f32(sn)])
    uv = np.concatenate([u0, _f32(su)])
    idx = np.concatenate([i0, si])
    matmap = np.concatenate([np.zeros(i0.shape[0], np.int32), np.ones(si.shape[0], np.int32)])
    mats = make_materials([(0.6, 0.5, 0.4, 0.3, 0.0, 0.0), (1.0, 1.0, 1.0, 1.0, 0.8, 0.0)])
    lights = make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(4.0, 8.0, -4.0), color=(1.0, 1.0, 1.0))])
    cam = dict(pos=(0.0, 4.0, -8.0), look_at=(0.0, 0.5, 0.0), up=(0.0, 1.0, 0.0), fov=np.pi / 4)
    return SceneData(pos, nrm, uv, idx, matmap, mats, lights, cam, 512, 512, 1, meta=dict(name="C1", seed=seed))


def config_c2(seed: int = SEED_BASE + 2, n_spheres: int = 200, res=(1920, 1080), slices: int = 50,
              stacks: int = 50) -> SceneData:
    rng = Rng(seed)
    radii = np.exp(rng.uniform(n_spheres, np.log(0.05), np.log(1.0)))
    centres = rng.uniform(3 * n_spheres, -10.0, 10.0).reshape(n_spheres, 3)
    kind = rng.uniform(n_spheres)
    colours = rng.uniform(3 * n_spheres, 0.2, 1.0).reshape(n_spheres, 3)
    vps = (stacks + 1) * (slices + 1)
    jit = rng.uniform(3 * vps * n_spheres, -0.01, 0.01).reshape(n_spheres, vps, 3)
    P, N, U, I, MM, rows = [], [], [], [], [], []
    base = 0
    for s in range(n_spheres):
        p, nn, uv, idx = uv_sphere(slices, stacks, centres[s], radii[s], jit[s] * radii[s], base)
        P.append(p), N.append(nn), U.append(uv), I.append(idx)
        MM.append(np.full(idx.shape[0], s, np.int32))
        r, g, b = colours[s]
        if kind[s] < 0.3:
            rows.append((r, g, b, 1.0, 0.9, 1.5))  # refractive
        elif kind[s] < 0.6:
            rows.append((r, g, b, 1.0, 0.6, 0.0))  # mirror
        else:
            rows.append((r, g, b, 0.5, 0.0, 0.0))  # diffuse
        base += p.shape[0]
    mats = make_materials(rows)
    lights = make_lights((0.1, 0.1, 0.1), [
        dict(type=LIGHT_POINT, pos=(15.0, 20.0, -25.0), color=(0.7, 0.7, 0.7)),
        dict(type=LIGHT_POINT, pos=(-20.0, 10.0, -15.0), color=(0.5, 0.5, 0.6)),
    ])
    cam = dict(pos=(0.0, 0.0, -35.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=np.pi / 4)
    return SceneData(_f32(np.concatenate(P)), _f32(np.concatenate(N)), _f32(np.concatenate(U)),
                     np.concatenate(I), np.concatenate(MM), mats, lights, cam, res[0], res[1], 3,
                     meta=dict(name="C2", seed=seed, spheres=n_spheres))


def waves_c3(seed: int = SEED_BASE + 3, n: int = 8) -> np.ndarray:
    rng = Rng(seed ^ 0xA5A5)
    a = rng.uniform(n, 0.01, 0.1)
    lam = np.exp(rng.uniform(n, np.log(0.05), np.log(1.0)))
    kappa = 2 * np.pi / lam
    ang = rng.uniform(n, 0.0, 2 * np.pi)
    phi = rng.uniform(n, 0.0, 2 * np.pi)
    omega = np.sqrt(9.81 * kappa)
    return np.stack([a, kappa, np.cos(ang), np.sin(ang), omega, phi], 1).astype(np.float32)


def config_c3(seed: int = SEED_BASE + 3, n: int = 724, res=(1280, 720)) -> SceneData:
    pos, nrm, uv, idx = heightfield(n, 5.0, np.zeros((n, n)))
    matmap = np.zeros(idx.shape[0], np.int32)
    mats = make_materials([(0.1, 0.3, 0.5, 1.0, 0.95, 1.333)])
    lights = make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(3.0, 10.0, -6.0), color=(1.0, 1.0, 1.0))])
    cam = dict(pos=(0.0, 3.0, -7.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=np.pi / 4)
    return SceneData(pos, nrm, uv, idx, matmap, mats, lights, cam, res[0], res[1], 2, waves=waves_c3(seed),
                     meta=dict(name="C3", seed=seed))


def _mixed_scene(seed: int, i: int, tris: float) -> SceneData:
    """Scene i of a C4/C5 set: sphere sets (5000-triangle UV spheres) at even
    i, heightfields displaced by a static wave field at odd i."""
    if i % 2 == 0:
        return config_c2(seed=seed + 16 * i, n_spheres=max(2, int(round(tris / 5000))), res=(256, 256))
    sd = config_c3(seed=seed + 16 * i, n=max(8, int(round(np.sqrt(tris / 2.0))) + 1), res=(256, 256))
    sd.pos[:, 1] = displace_host(sd, 0.5)[:, 1]
    return sd


def bounding_sphere(sd: SceneData):
    """Centre and radius of the scene's axis-aligned box."""
    lo, hi = sd.pos.min(0), sd.pos.max(0)
    return 0.5 * (lo + hi), 0.5 * float(np.linalg.norm(hi - lo))


def orbit_camera(sd: SceneData, direction, width: int, height: int, sphere=None):
    """Camera on a sphere of 2.5x the scene's bounding radius, looking at its
    centre (C4 frames, C5 requests); `sphere` = bounding_sphere(sd) when
    already known."""
    from . import camera_set
    c, r = sphere if sphere is not None else bounding_sphere(sd)
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.99 else (1.0, 0.0, 0.0)
    return camera_set(c + 2.5 * r * d, c, up, np.pi / 4, width / height)


@dataclass
class BatchFrame:
    sd: SceneData
    width: int
    height: int
    bounces: int
    direction: np.ndarray
    ray_capacity: int


def config_c4(seed: int = SEED_BASE + 4, n_scenes: int = 64) -> list:
    """C4 batch: n_scenes scenes with T ~ logU[10k, 2M] (sphere sets and
    heightfields alternating), one frame each at a resolution uniform over
    256^2, 512^2, 1280x720 and 1920x1080, depth uniform over 1..4, a camera on
    a random direction; secondary bundles of 2 x pixels + 2 rays."""
    rng = Rng(seed)
    tris = np.exp(rng.uniform(n_scenes, np.log(1e4), np.log(2e6)))
    pick = np.random.default_rng(4)
    res_set = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    out = []
    for i, ti in enumerate(tris):
        sd = _mixed_scene(seed, i, ti)
        w, h = res_set[int(pick.integers(0, 4))]
        depth = int(pick.integers(1, 5))
        out.append(BatchFrame(sd, w, h, depth, pick.normal(size=3), 2 * w * h + 2))
    return out


def config_c5(seed: int = SEED_BASE + 5, n_scenes: int = 12, total_tris: int = 20_000_000) -> list:
    """C5 resident scenes: T ~ logU[10k, 4M] per scene, rescaled so that the
    sum is about 20M triangles (sphere sets and heightfields alternating)."""
    rng = Rng(seed)
    t = np.exp(rng.uniform(n_scenes, np.log(1e4), np.log(4e6)))
    t = t * (total_tris / t.sum())
    return [_mixed_scene(seed, i, ti) for i, ti in enumerate(t)]


def c5_requests(sds, rate: float, n: int, seed: int, start_ns: int = 0):
    """A C5 request stream: Poisson arrivals at `rate` per second from
    start_ns (add the submission time to arrival_ns before submitting),
    scenes Zipf(1.1) over the resident set, resolution uniform over 256^2,
    512^2 and 1280x720, depth uniform over 1..3, a camera on a random
    direction.  Returns request dicts for rtx.Server.submit (scene = index)."""
    spheres = [bounding_sphere(sd) for sd in sds]
    rng = np.random.default_rng(seed)
    classes = [(256, 256), (512, 512), (1280, 720)]
    zipf = 1.0 / np.arange(1, len(sds) + 1) ** 1.1
    zipf /= zipf.sum()
    arrivals = start_ns + (np.cumsum(rng.exponential(1.0 / rate, n)) * 1e9).astype(np.int64)
    reqs = []
    for i in range(n):
        k = int(rng.choice(len(sds), p=zipf))
        w, h = classes[int(rng.integers(0, 3))]
        cam = orbit_camera(sds[k], rng.normal(size=3), w, h, spheres[k])
        reqs.append(dict(scene=k, width=w, height=h, bounces=int(rng.integers(1, 4)), camera=cam,
                         arrival_ns=int(arrivals[i]), tag=i))
    return reqs


def displace_host(sd: SceneData, t: float) -> np.ndarray:
    """Host copy of the wave displacement (float32, same formula as the
    device kernel) — used only to stage positions for update_positions."""
    p = sd.pos.astype(np.float32).copy()
    y = np.zeros(p.shape[0], np.float32)
    for a, k, dx, dz, om, ph in sd.waves:
        y += a * np.sin(k * (dx * p[:, 0] + dz * p[:, 2]) - om * np.float32(t) + ph)
    p[:, 1] = y
    return p


def load(scene, sd: SceneData):
    """Upload a SceneData into an rtx.Scene."""
    scene.set_geometry(sd.pos, sd.idx, sd.nrm, sd.uv, sd.matmap)
    scene.set_materials(sd.materials)
    scene.set_lights(sd.lights)
    return scene


def camera_of(sd: SceneData):
    from . import camera_set
    c = sd.camera
    if "dir" in c:
        return camera_set(c["pos"], None, c["up"], c["fov"], sd.width / sd.height, direction=c["dir"])
    return camera_set(c["pos"], c["look_at"], c["up"], c["fov"], sd.width / sd.height)


# ------------------------------------------------------------------ objects --
# The reference's scenes of objects: meshes from OBJ files, each object a
# mesh under GeometricProperties (pos, rpy, scale) with a material, merged by
# Scene::create_aggregate_mesh into one mesh whose vertices are transformed
# on the host.  Float32 arithmetic in the reference's order of evaluation.

def _mat4_mul(a, b):
    """sqmat<4>::operator* (include/rt/matrix.hpp:55-66): each entry summed
    over k = 0..3 from 0 in float32."""
    r = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            acc = np.float32(0.0)
            for k in range(4):
                acc = np.float32(acc + np.float32(a[i, k] * b[k, j]))
            r[i, j] = acc
    return r


def geom_matrix(pos=(0.0, 0.0, 0.0), rpy=(0.0, 0.0, 0.0), scale=1.0) -> np.ndarray:
    """GeometricProperties::updateTransform (src/rt/geom.cpp:35-42): M =
    translation * rotation * scale, the rotation MY * MP * MR of
    rotationMatrix4x4 (src/rt/math.cpp:93-127), cos/sin in double then float."""
    import math
    f = np.float32
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = np.asarray(pos, np.float32)
    cr, sr = f(math.cos(f(rpy[0]))), f(math.sin(f(rpy[0])))
    cp, sp = f(math.cos(f(rpy[1]))), f(math.sin(f(rpy[1])))
    cy, sy = f(math.cos(f(rpy[2]))), f(math.sin(f(rpy[2])))
    MR = np.zeros((4, 4), np.float32)
    MR[2, 2], MR[0, 0], MR[0, 1], MR[1, 0], MR[1, 1] = 1.0, cr, -sr, sr, cr
    MP = np.zeros((4, 4), np.float32)
    MP[0, 0], MP[1, 1], MP[1, 2], MP[2, 1], MP[2, 2] = 1.0, cp, -sp, sp, cp
    MY = np.zeros((4, 4), np.float32)
    MY[1, 1], MY[0, 0], MY[0, 2], MY[2, 0], MY[2, 2] = 1.0, cy, sy, -sy, cy
    R = _mat4_mul(_mat4_mul(MY, MP), MR)
    R[3, 3] = 1.0
    sc = np.asarray(scale, np.float32)
    S = np.zeros((4, 4), np.float32)
    S[0, 0], S[1, 1], S[2, 2] = (sc, sc, sc) if sc.ndim == 0 else sc
    S[3, 3] = 1.0
    return _mat4_mul(_mat4_mul(T, R), S)


def transform_vertices(M: np.ndarray, pos: np.ndarray, nrm: np.ndarray):
    """GeometricProperties::transform (src/rt/geom.cpp:45-64): M times (p, 1)
    and (n, 0), each row summed in index order, then homogenize with
    1/w (the normal's w set to 1 first)."""
    M = np.asarray(M, np.float32)

    def mv(v, w):
        out = np.zeros((v.shape[0], 4), np.float32)
        for i in range(4):
            acc = np.zeros(v.shape[0], np.float32)
            for j in range(3):
                acc = acc + M[i, j] * v[:, j]
            out[:, i] = acc + M[i, 3] * np.float32(w)
        return out
    p4 = mv(np.asarray(pos, np.float32), 1.0)
    n4 = mv(np.asarray(nrm, np.float32), 0.0)
    n4[:, 3] = 1.0
    ip = (np.float32(1.0) / p4[:, 3]).astype(np.float32)
    inn = (np.float32(1.0) / n4[:, 3]).astype(np.float32)
    return (p4[:, :3] * ip[:, None]).astype(np.float32), (n4[:, :3] * inn[:, None]).astype(np.float32)


def aggregate(objects, lights, camera, width, height, max_bounces, name="objects") -> SceneData:
    """Scene::create_aggregate_mesh (src/rt/scene.cpp:495-539): objects =
    [(model with pos/nrm/uv/idx, M, material row (r, g, b, shininess,
    reflectiveness, ior))]; material i belongs to object i."""
    P, N, U, I, MM, rows = [], [], [], [], [], []
    base = 0
    for k, (m, M, row) in enumerate(objects):
        p, n = transform_vertices(M, m.pos, m.nrm)
        P.append(p), N.append(n), U.append(np.asarray(m.uv, np.float32))
        I.append(np.asarray(m.idx, np.uint32) + np.uint32(base))
        MM.append(np.full(len(m.idx), k, np.int32))
        rows.append(row)
        base += len(p)
    return SceneData(np.concatenate(P), np.concatenate(N), np.concatenate(U), np.concatenate(I),
                     np.concatenate(MM), make_materials(rows), lights, camera, width, height, max_bounces,
                     meta=dict(name=name))


def _normalized(v):
    """vec3::normalize (include/rt/vector.hpp:151-175): v / sqrt(sum v_i^2)."""
    v = np.asarray(v, np.float32)
    n = np.float32(0.0)
    for x in v:
        n = np.float32(n + np.float32(x * x))
    return (v / np.float32(np.sqrt(n))).astype(np.float32)


def preset_boat(models_dir, res=(1280, 720), max_bounces=3, view=0) -> SceneData:
    """boat_set_scene (include/rt/test-params.hpp:10-50): the water frame
    (scale 2, at (0, -8, 0), blue, reflectiveness 0.98, IOR 1.333) and the
    boat (same place and scale, red, shininess 1), each loaded as one mesh
    (load_obj_file_as_aggregate); a directional light toward (0.05, -1,
    -0.02), colour 0.7, ambient 0.2; the camera at one of the three stats
    views, FOV pi/4."""
    import math
    from pathlib import Path
    from . import ObjModel
    d = Path(models_dir)
    water = ObjModel(d / "frame_water1.obj")
    boat = ObjModel(d / "frame_boat1.obj")
    M = geom_matrix((0.0, -8.0, 0.0), (0.0, 0.0, 0.0), 2.0)
    objects = [(water, M, (0.0, 0.0, 1.0, 0.0, 0.98, 1.333)), (boat, M, (1.0, 0.0, 0.0, 1.0, 0.0, 0.0))]
    lights = make_lights((0.2, 0.2, 0.2), [dict(type=LIGHT_DIR, dir=tuple(_normalized((0.05, -1.0, -0.02))),
                                                color=(0.7, 0.7, 0.7))])
    cams = [((20.0186, -5.49632, 71.8718), (0.131732, 0.0845093, -0.987677)),
            ((-56.9387, -2.41959, 29.574), (0.927574, -0.22893, -0.295292)),
            ((68.6859, 5.18034, 13.6691), (-0.820819, -0.478219, -0.31235))]
    pos, dr = cams[view]
    cam = dict(pos=pos, dir=dr, up=(0.0, 1.0, 0.0), fov=math.pi / 4)
    return aggregate(objects, lights, cam, res[0], res[1], max_bounces, name="boat")
