"""Pins of the oracle's light model: POINT, SPOT and DIR lights in shading and
in shadow rays, from what geometry fixes, never through the oracle's own ray
builder.

Shading (shade_sample, src/kernel/ray-shader.cl:160-176 and the per-light
sum): for a light whose direction of travel at the hit point is Ld,
    val = ambient*diffuse
        + sum over unshadowed lights of  LD*rgb*diffuse*cos^2 + LD*rgb*spec^8*cos^2*shininess
with cos = clamp(n . -Ld), spec = clamp(reflect(d, n) . -Ld), then
*(1 - reflectiveness) and floor(x * 1e6).  Here every cosine is set up from
positions (a light at height h and offset x above the point: cos = h/sqrt(h^2
+ x^2)), so a wrong sign of Ld, a missing normalisation, a wrong power or a
light summed in the wrong slot changes the pixel by far more than the
tolerance of a few fixed-point units that float32 rounding needs.

SPOT (ray-shader.cl:168-176): LD = clamp((Ld . spot_dir - angle) / (1 - angle)),
with Ld = normalize(hit - pos): the cone's axis gives 1, the cut-off cosine 0.

Shadow rays (shadow_trace_single, src/kernel/shadow-trace-bvh.cl:375-418; POINT
is the configs' extension): origin the hit point, tMin 0.01; DIR toward
-dir with tMax 1e37; SPOT toward pos with tMax 1e37, and a point outside the
cone (dot(-dir_to_light, spot_dir) < angle) is shadowed untraced; POINT
toward pos with tMax = distance.  The checker below builds these rays itself
in float64 and intersects every triangle (Moller-Trumbore, t in (tMin,
tMax)), so a wrong direction, tMax or cone rule in the oracle fails it.
"""
import numpy as np
import pytest

import oracle as O

TOL = 3  # fixed-point units of 1e-6: float32 rounding of the closed forms


def mats(rows):
    m = np.zeros(len(rows), [("diffuse", "<f4", 4), ("shininess", "<f4"), ("reflectiveness", "<f4"),
                             ("refractive_index", "<f4"), ("texture", "<i4")])
    for i, (r, g, bb, sh, rf, ior) in enumerate(rows):
        m[i] = ((r, g, bb, 0), sh, rf, ior, -1)
    return m


def lights(specs, ambient=(0.1, 0.1, 0.1)):
    """specs: dicts with type, pos, dir, color, angle."""
    L = np.zeros(1, O.LIGHTS_DTYPE)
    L["ambient"] = ambient
    L["count"] = len(specs)
    for i, s in enumerate(specs):
        L["l"]["type"][0, i] = s["type"]
        L["l"]["pos"][0, i] = s.get("pos", (0, 0, 0))
        L["l"]["dir"][0, i] = s.get("dir", (0, 0, 0))
        L["l"]["color"][0, i] = s.get("color", (1, 1, 1))
        L["l"]["angle"][0, i] = s.get("angle", 0.0)
    return L


def shade1(material, L, n, d, p=(0, 0, 0), flags=1, contrib=1.0):
    h = O.empty_hits(1)
    h["flags"][0] = flags
    h["id"][0] = 0
    h["n4"][0, :3] = n
    h["p4"][0, :3] = p
    acc = np.zeros((1, 3), np.uint32)
    d4 = np.array([[*d, 1e36]], np.float32)
    O.shade(material, np.zeros(1, np.int32), L, d4, np.zeros(1, np.int32), np.array([contrib], np.float32), h, acc)
    return acc[0].astype(np.int64)


def fixed(v):
    return np.floor(np.clip(np.asarray(v, np.float64), 0.0, 1.0) * 1e6).astype(np.int64)


def close(acc, expect):
    e = fixed(expect)
    assert np.all(np.abs(acc - e) <= TOL), (acc.tolist(), e.tolist())


DIFF = (0.25, 0.5, 0.75)
SHIN = 0.3
AMB = 0.1


def lit(cos2, spec8, LD=1.0, rgb=(1, 1, 1), diff=DIFF, shin=SHIN, amb=AMB):
    diff = np.asarray(diff, np.float64)
    rgb = np.asarray(rgb, np.float64) * LD
    return amb * diff + rgb * diff * cos2 + rgb * spec8 * cos2 * shin


# ------------------------------------------------------------ POINT lights --
def test_point_light_overhead():
    # light 4 above the point, view straight down: the mirror direction points
    # at the light, so cos = 1 and spec = 1
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    L = lights([dict(type=O.POINT, pos=(0, 4, 0))])
    close(shade1(m, L, (0, 1, 0), (0, -1, 0)), lit(1.0, 1.0))


def test_point_light_oblique_cos_squared():
    # light at (3, 4, 0) from the point: cos = 4/5 toward the normal; the view
    # ray arrives along the mirror of the light direction, so spec = 1
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    L = lights([dict(type=O.POINT, pos=(3, 4, 0))])
    close(shade1(m, L, (0, 1, 0), (0.6, -0.8, 0)), lit(0.64, 1.0))
    # the same light translated with the point: only the relative position counts
    L2 = lights([dict(type=O.POINT, pos=(13, 6, -7))])
    close(shade1(m, L2, (0, 1, 0), (0.6, -0.8, 0), p=(10, 2, -7)), lit(0.64, 1.0))


def test_point_light_specular_power_eight():
    # view straight down: the mirror ray is +y, at cos 4/5 from the light
    # direction, so spec = 0.8^8 (a power of 2 or 16 would be far off)
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    L = lights([dict(type=O.POINT, pos=(3, 4, 0))])
    close(shade1(m, L, (0, 1, 0), (0, -1, 0)), lit(0.64, 0.8 ** 8))


def test_point_light_no_falloff():
    # the configs' POINT light has LD = 1 at any distance
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    near = shade1(m, lights([dict(type=O.POINT, pos=(0, 1, 0))]), (0, 1, 0), (0, -1, 0))
    far = shade1(m, lights([dict(type=O.POINT, pos=(0, 1000, 0))]), (0, 1, 0), (0, -1, 0))
    assert np.array_equal(near, far)


def test_point_light_below_surface_only_ambient():
    # a light under the surface: n . -Ld < 0 clamps to 0, so both terms vanish
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    L = lights([dict(type=O.POINT, pos=(3, -4, 0))])
    close(shade1(m, L, (0, 1, 0), (0.6, -0.8, 0)), AMB * np.asarray(DIFF))


def test_two_point_lights_summed_in_order_and_masked_by_their_bits():
    # red light overhead (cos 1, spec 1), green light oblique (cos^2 .64,
    # spec .8^8); bit 8 masks light 0, bit 9 light 1
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    L = lights([dict(type=O.POINT, pos=(0, 4, 0), color=(1, 0, 0)),
                dict(type=O.POINT, pos=(3, 4, 0), color=(0, 1, 0))])
    a = AMB * np.asarray(DIFF)
    red = lit(1.0, 1.0, rgb=(1, 0, 0)) - a
    green = lit(0.64, 0.8 ** 8, rgb=(0, 1, 0)) - a
    close(shade1(m, L, (0, 1, 0), (0, -1, 0)), a + red + green)
    close(shade1(m, L, (0, 1, 0), (0, -1, 0), flags=1 | 1 << 8), a + green)
    close(shade1(m, L, (0, 1, 0), (0, -1, 0), flags=1 | 1 << 9), a + red)
    close(shade1(m, L, (0, 1, 0), (0, -1, 0), flags=1 | 3 << 8), a)


def test_point_light_reflectiveness_and_contribution():
    m = mats([(*DIFF, SHIN, 0.5, 0.0)])
    L = lights([dict(type=O.POINT, pos=(3, 4, 0))])
    close(shade1(m, L, (0, 1, 0), (0.6, -0.8, 0), contrib=0.5), np.minimum(lit(0.64, 1.0), 1.0) * 0.5 * 0.5)


# ------------------------------------------------------------- SPOT lights --
@pytest.mark.parametrize("angle,LD", [(0.6, 0.5), (0.4, 2.0 / 3.0), (0.8, 0.0), (0.9, 0.0)])
def test_spot_light_cone_falloff(angle, LD):
    # spot 4 above, pointing straight down; the point is 3 off the axis, so
    # the angle from the axis has cosine 4/5: LD = clamp((0.8 - angle)/(1 - angle))
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    L = lights([dict(type=O.SPOT, pos=(3, 4, 0), dir=(0, -1, 0), angle=angle)])
    close(shade1(m, L, (0, 1, 0), (0.6, -0.8, 0)), lit(0.64, 1.0, LD=LD))


def test_spot_light_on_axis_full_strength():
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    L = lights([dict(type=O.SPOT, pos=(0, 4, 0), dir=(0, -1, 0), angle=0.7)])
    close(shade1(m, L, (0, 1, 0), (0, -1, 0)), lit(1.0, 1.0, LD=1.0))


def test_dir_light_oblique():
    # DIR travels along dir: cos = n . -dir
    m = mats([(*DIFF, SHIN, 0.0, 0.0)])
    L = lights([dict(type=O.DIR, dir=(-0.6, -0.8, 0))])
    close(shade1(m, L, (0, 1, 0), (0, -1, 0)), lit(0.64, 0.8 ** 8))


# ------------------------------------------------------------ shadow rays --
def _mt(o, d, tri):
    """float64 Moller-Trumbore of one ray against (T, 3, 3) triangles: t, u,
    v and whether the determinant passes the |a| > 1e-26 test."""
    v0, v1, v2 = tri[:, 0], tri[:, 1], tri[:, 2]
    e1, e2 = v1 - v0, v2 - v0
    h = np.cross(d, e2)
    a = np.einsum("ij,ij->i", e1, h)
    ok = np.abs(a) > 1e-26
    f = np.where(ok, 1.0 / np.where(ok, a, 1.0), 0.0)
    s = o - v0
    u = f * np.einsum("ij,ij->i", s, h)
    q = np.cross(s, e1)
    v = f * (q @ d)
    t = f * np.einsum("ij,ij->i", e2, q)
    return t, u, v, ok


def brute_shadow(pos, idx, L, points, eps=1e-5):
    """Expected shadow bit per (point, light) and an 'ambiguous' mask where
    float32 arithmetic may decide otherwise: no hit lies clearly inside a
    triangle and clearly inside (tMin, tMax), but some candidate lies within
    eps of an edge or of tMin/tMax, or the point is on the cone's boundary."""
    tri = pos.astype(np.float64)[idx]
    nl = int(L["count"][0])
    shadow = np.zeros((len(points), nl), bool)
    amb = np.zeros((len(points), nl), bool)
    for k, p in enumerate(points.astype(np.float64)):
        for li in range(nl):
            ty = int(L["l"]["type"][0, li])
            lpos = L["l"]["pos"][0, li].astype(np.float64)
            ldir = L["l"]["dir"][0, li].astype(np.float64)
            tmax = 1e37
            if ty == O.DIR:
                d = -ldir
            else:
                to = lpos - p
                dist = np.sqrt(to @ to)
                d = to / dist
                if ty == O.POINT:
                    tmax = dist
                else:
                    c = (-d) @ ldir
                    ang = float(L["l"]["angle"][0, li])
                    if abs(c - ang) < 1e-5:
                        amb[k, li] = True
                    if c < ang:
                        shadow[k, li] = True
                        continue
            t, u, v, ok = _mt(p, d, tri)
            inside = ok & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1)
            hit = inside & (t > 0.01) & (t < tmax)
            shadow[k, li] = bool(hit.any())
            rel = eps * max(1.0, tmax if tmax < 1e30 else 1.0)
            clear_in = ok & (u >= eps) & (v >= eps) & (u + v <= 1 - eps)
            near_in = ok & (u >= -eps) & (v >= -eps) & (u + v <= 1 + eps) & (u <= 1 + eps)
            t_clear = (t > 0.01 + eps) & (t < tmax - rel)
            t_near = (t > 0.01 - eps) & (t < tmax + rel)
            if (clear_in & t_clear).any():
                continue
            amb[k, li] |= bool((near_in & t_near & ~(clear_in & t_clear)).any())
    return shadow, amb


def oracle_shadow(pos, idx, L, points):
    b = O.Build(pos, idx, None, 30, 1)
    n = len(points)
    p4 = np.zeros((n, 4), np.float32)
    p4[:, :3] = points
    f = O.trace_shadow(b, L, np.ones(n, np.uint32), p4)
    nl = int(L["count"][0])
    got = np.stack([(f >> (8 + li)) & 1 for li in range(nl)], axis=1).astype(bool)
    fb = O.trace_shadow_brute(b, L, np.ones(n, np.uint32), p4)
    assert np.array_equal(f, fb), "oracle BVH any-hit differs from its brute force"
    return got


def quad(y, x0, x1, z0, z1):
    """Two triangles of the rectangle [x0,x1] x [z0,z1] at height y."""
    p = np.array([[x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]], np.float32)
    return p, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def merge(*parts):
    pos, idx, off = [], [], 0
    for p, i in parts:
        pos.append(p)
        idx.append(i + off)
        off += len(p)
    return np.concatenate(pos), np.concatenate(idx)


def test_point_shadow_occluder_between_and_beyond():
    # ground points at y = 0, a point light at y = 2 over x = 0; an occluder
    # slab at y = 1 covers x in [-10, 0] and a roof at y = 3 covers everything.
    # The ray from (x, 0) to the light crosses y = 1 at x/2: shadowed iff x < 0
    # (x/2 in the slab).  The roof lies beyond the light: a POINT light's ray
    # stops at the light (tMax = distance), so the roof never shadows.
    pos, idx = merge(quad(1.0, -10, 0, -10, 10), quad(3.0, -20, 20, -20, 20))
    xs = np.array([-3.1, -1.3, -0.7, 0.7, 1.3, 3.1], np.float32)
    pts = np.stack([xs, np.zeros_like(xs), np.full_like(xs, 0.37)], axis=1)
    L = lights([dict(type=O.POINT, pos=(0, 2, 0.37))])
    got = oracle_shadow(pos, idx, L, pts)[:, 0]
    assert got.tolist() == (xs < 0).tolist()
    exp, amb = brute_shadow(pos, idx, L, pts)
    assert not amb.any() and np.array_equal(got, exp[:, 0])


def test_dir_and_spot_shadowed_by_roof_beyond_light():
    # the same roof above the light: DIR and SPOT rays run to tMax = 1e37
    # (shadow-trace-bvh.cl:412), so the roof shadows every point for them
    pos, idx = quad(3.0, -20, 20, -20, 20)
    xs = np.array([-1.3, 0.7, 1.3], np.float32)
    pts = np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], axis=1)
    L = lights([dict(type=O.POINT, pos=(0, 2, 0)),
                dict(type=O.SPOT, pos=(0, 2, 0), dir=(0, -1, 0), angle=0.1),
                dict(type=O.DIR, dir=(0, -1, 0))])
    got = oracle_shadow(pos, idx, L, pts)
    assert got[:, 0].tolist() == [False] * 3
    assert got[:, 1].tolist() == [True] * 3
    assert got[:, 2].tolist() == [True] * 3


def test_spot_outside_cone_shadowed_without_geometry_in_the_way():
    # a far-away triangle that no ray meets; the spot at (0, 4, 0) points down
    # with cut-off cosine 0.9: points at offset x have cos = 4/sqrt(16 + x^2)
    pos, idx = quad(-50.0, 100, 101, 100, 101)
    xs = np.array([0.0, 1.0, 1.5, 2.0, 3.0, 5.0], np.float32)  # cos 1, .970, .936, .894, .8, .625
    pts = np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], axis=1)
    L = lights([dict(type=O.SPOT, pos=(0, 4, 0), dir=(0, -1, 0), angle=0.9)])
    got = oracle_shadow(pos, idx, L, pts)[:, 0]
    assert got.tolist() == [False, False, False, True, True, True]


def test_shadow_tmin_skips_self_and_near_geometry():
    # an occluder 0.005 above the point lies inside tMin = 0.01; one at 0.02 does not
    for y, shadowed in [(0.005, False), (0.02, True)]:
        pos, idx = quad(y, -1, 1, -1, 1)
        L = lights([dict(type=O.POINT, pos=(0, 2, 0))])
        got = oracle_shadow(pos, idx, L, np.zeros((1, 3), np.float32))[0, 0]
        assert got == shadowed


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_shadow_flags_against_independent_brute_force(seed):
    # random triangle soup, random hit points, one light of each type: the
    # oracle's flags equal a float64 brute force over rays built here
    rng = np.random.default_rng(1000 + seed)
    T = 150
    cent = rng.uniform(-4, 4, (T, 1, 3))
    pos = (cent + rng.normal(scale=0.6, size=(T, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * T, dtype=np.uint32).reshape(T, 3)
    pts = rng.uniform(-5, 5, (400, 3)).astype(np.float32)
    L = lights([dict(type=O.POINT, pos=tuple(rng.uniform(-3, 3, 3))),
                dict(type=O.SPOT, pos=tuple(rng.uniform(-3, 3, 3)), dir=(0.0, -1.0, 0.0), angle=0.3),
                dict(type=O.DIR, dir=tuple(np.array([0.36, -0.48, 0.8])))])
    got = oracle_shadow(pos, idx, L, pts)
    exp, amb = brute_shadow(pos, idx, L, pts)
    ok = ~amb
    assert ok.mean() > 0.97
    assert np.array_equal(got[ok], exp[ok]), np.argwhere((got != exp) & ok)[:5]
    # every light type decides both ways on this soup
    for li in range(3):
        assert exp[ok[:, li], li].any() and (~exp[ok[:, li], li]).any()
