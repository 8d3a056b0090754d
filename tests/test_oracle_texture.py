"""Pins of the oracle's texture sampling and cubemap lookup (shade_sample,
src/kernel/ray-shader.cl:145-153, 187-240, 275-307).

The reference samples through OpenCL images (CLK_FILTER_LINEAR, normalized
coordinates, REPEAT for material textures, CLAMP_TO_EDGE for the cubemap)
and its vendor's texture unit; the weights' precision is implementation
defined, so the oracle follows the OpenCL 1.2 filter formulas.  They are
held here against what those formulas fix in closed form: texel centres give
the texel, midpoints the mean of two texels, a linear ramp is reproduced,
coordinates repeat with period 1, the edges wrap (REPEAT) or clamp
(CLAMP_TO_EDGE), and UNORM8 texels read as c/255.  Cubemap: axis directions
pick their face's centre, the face order of the kernel's if-chain decides
edges and corners, the face coordinates follow that chain's table, and a
miss with no face found adds black.  Shading: a constant texture shades
exactly as the same colour given as the material's diffuse.
"""
import numpy as np

import oracle as O


def _img(h, w, seed):
    rng = np.random.default_rng(seed)
    im = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    im[..., 3] = 255
    return im


def _f(im):
    return im[..., :3].astype(np.float32) / np.float32(255.0)


def test_texel_centres_and_midpoints():
    for h, w in ((8, 16), (4, 4), (32, 8)):
        im = _img(h, w, h * w)
        ts = O.TexSet([im])
        jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        st = np.stack([(ii + 0.5) / w, (jj + 0.5) / h], -1).reshape(-1, 2).astype(np.float32)
        got = O.sample_texture(ts, 0, st).reshape(h, w, 3)
        assert np.array_equal(got, _f(im)), "a texel centre must give the texel (power-of-two sizes: exact)"
        # midway between texels i and i+1 of a row: their mean
        st = np.stack([(ii[:, :-1] + 1.0) / w, (jj[:, :-1] + 0.5) / h], -1).reshape(-1, 2).astype(np.float32)
        got = O.sample_texture(ts, 0, st).reshape(h, w - 1, 3)
        want = 0.5 * _f(im)[:, :-1] + 0.5 * _f(im)[:, 1:]
        assert np.allclose(got, want, atol=1e-6)


def test_linear_ramp_is_reproduced():
    # R = x (texel column), G = y (row), B = 7: bilinear filtering reproduces
    # affine functions, so the sample is (u - 0.5, v - 0.5) inside the image
    h, w = 40, 50
    jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    im = np.stack([ii * 5, jj * 6, np.full_like(ii, 7), np.full_like(ii, 255)], -1).astype(np.uint8)
    ts = O.TexSet([im])
    rng = np.random.default_rng(1)
    s = rng.uniform(0.5 / w, 1 - 0.5 / w, 5000).astype(np.float32)
    t = rng.uniform(0.5 / h, 1 - 0.5 / h, 5000).astype(np.float32)
    got = O.sample_texture(ts, 0, np.stack([s, t], 1)).astype(np.float64)
    u = s.astype(np.float64) * w - 0.5
    v = t.astype(np.float64) * h - 0.5
    assert np.allclose(got[:, 0], 5 * u / 255, atol=2e-5)
    assert np.allclose(got[:, 1], 6 * v / 255, atol=2e-5)
    assert np.allclose(got[:, 2], 7 / 255, atol=1e-7)


def test_repeat_period_and_edge_wrap():
    im = _img(16, 8, 3)
    ts = O.TexSet([im])
    k = np.arange(-2048, 2048, 7, dtype=np.float32) / 1024.0  # exact in float32, as are k + 1 and k - 3
    st = np.stack([k, k[::-1]], 1)
    a = O.sample_texture(ts, 0, st)
    assert np.array_equal(a, O.sample_texture(ts, 0, st + 1.0))
    assert np.array_equal(a, O.sample_texture(ts, 0, st - 3.0))
    # s = 0: half of the last column and half of the first (texel-centre rows)
    rows = (np.arange(16) + 0.5) / 16
    got = O.sample_texture(ts, 0, np.stack([np.zeros(16), rows], 1).astype(np.float32))
    assert np.allclose(got, 0.5 * _f(im)[:, -1] + 0.5 * _f(im)[:, 0], atol=1e-6)
    # a non-finite coordinate samples as 0
    assert np.array_equal(O.sample_texture(ts, 0, np.float32([[np.nan, 0.3]])),
                          O.sample_texture(ts, 0, np.float32([[0.0, 0.3]])))


def test_clamp_to_edge():
    im = _img(8, 8, 4)
    ts = O.TexSet([im])
    rows = (np.arange(8) + 0.5) / 8
    for s, col in ((0.0, 0), (0.03, 0), (1.0, 7), (0.98, 7)):
        got = O.sample_texture(ts, 0, np.stack([np.full(8, s), rows], 1).astype(np.float32), repeat=False)
        assert np.allclose(got, _f(im)[:, col], atol=1e-6), f"s={s}"


def _faces(constant=True):
    cols = [(255, 0, 0), (0, 255, 0), (0, 0, 255), (255, 255, 0), (0, 255, 255), (255, 0, 255)]
    faces = []
    for f, c in enumerate(cols):
        if constant:
            im = np.zeros((4, 4, 4), np.uint8)
            im[..., :3] = c
        else:  # R = column ramp, G = row ramp, B = face number
            jj, ii = np.meshgrid(np.arange(64), np.arange(64), indexing="ij")
            im = np.stack([ii * 4, jj * 4, np.full_like(ii, 40 * f), np.full_like(ii, 0)], -1).astype(np.uint8)
        im[..., 3] = 255
        faces.append(im)
    return faces


def _dirs(d):
    d = np.asarray(d, np.float32)
    out = np.zeros((len(d), 4), np.float32)
    out[:, :3] = d
    return out


def test_cubemap_faces_and_order():
    ts = O.TexSet([], cubemap=_faces())
    axes = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    got = O.sample_cubemap(ts, _dirs(axes))
    want = np.float32([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1]])
    assert np.array_equal(got, want)
    # edges and corners go to the first face of the chain +x, -x, +y, -y, +z, -z
    got = O.sample_cubemap(ts, _dirs([(1, 1, 0), (-1, -1, 1), (0, 1, 1), (0, -1, -1), (1, 1, 1)]))
    assert np.array_equal(got, want[[0, 1, 2, 3, 0]])
    # a NaN direction matches no face: black
    assert np.array_equal(O.sample_cubemap(ts, _dirs([(np.nan, 0, 0)])), np.zeros((1, 3), np.float32))


def test_cubemap_face_coordinates():
    # with ramp faces the sample returns (s*64 - 0.5, t*64 - 0.5)*4/255; the
    # expected (s, t) below are the kernel's table evaluated by hand for
    # d = (a, b, c) on the face of its major axis
    ts = O.TexSet([], cubemap=_faces(constant=False))
    a, b, c = 1.0, 0.5, 0.25
    cases = [  # direction, face, s, t
        ((a, b, c), 0, (1 - c) / 2, (1 + b) / 2),
        ((-a, b, c), 1, (1 + c) / 2, (1 + b) / 2),
        ((b, a, c), 2, (1 + b) / 2, (1 - c) / 2),
        ((b, -a, c), 3, (1 + b) / 2, (1 + c) / 2),
        ((b, c, a), 4, (1 + b) / 2, (1 + c) / 2),
        ((b, c, -a), 5, (1 - b) / 2, (1 + c) / 2),
    ]
    got = O.sample_cubemap(ts, _dirs([d for d, *_ in cases]))
    for (d, face, s, t), g in zip(cases, got):
        assert abs(g[2] - 40 * face / 255) < 1e-6, f"{d}: wrong face"
        assert abs(g[0] - (s * 64 - 0.5) * 4 / 255) < 1e-5 and abs(g[1] - (t * 64 - 0.5) * 4 / 255) < 1e-5, \
            f"{d}: coordinates {g[:2] * 255 / 4} against {(s * 64 - 0.5, t * 64 - 0.5)}"


def test_constant_texture_shades_as_diffuse():
    from rtx import make_materials, make_lights, LIGHT_POINT
    rng = np.random.default_rng(7)
    n = 500
    k = (90, 160, 220)
    tex = np.zeros((8, 8, 4), np.uint8)
    tex[..., :3] = k
    tex[..., 3] = 255
    ts = O.TexSet([tex])
    diffuse = [np.float32(x) / np.float32(255.0) for x in k]
    plain = make_materials([(float(diffuse[0]), float(diffuse[1]), float(diffuse[2]), 0.7, 0.2, 0.0)])
    textured = plain.copy().view(np.int32)
    textured[7] = 0  # texture index of material 0
    textured = textured.view(np.uint8)
    lights = O.lights_from(make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(3.0, 5.0, -2.0))]))
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nrm = -d + rng.normal(size=(n, 3)).astype(np.float32) * 0.3
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    hits = {"flags": np.where(rng.uniform(size=n) < 0.5, 1, 1 | 256).astype(np.uint32),
            "id": np.zeros(n, np.int32), "n4": O.pos4(nrm.astype(np.float32)),
            "p4": O.pos4(rng.uniform(-1, 1, (n, 3)).astype(np.float32)),
            # texel centres (shifted by whole periods): the sample is the texel itself
            "uv2": ((rng.integers(0, 8, (n, 2)) + 0.5) / 8 + rng.integers(-3, 3, (n, 2))).astype(np.float32)}
    dir4 = O.pos4(d)
    px = rng.integers(0, 50, n).astype(np.int32)
    con = rng.uniform(0.2, 1.0, n).astype(np.float32)
    a = O.shade(plain, np.zeros(1, np.int32), lights, dir4, px, con, hits, np.zeros((50, 3), np.uint32))
    b = O.shade(textured, np.zeros(1, np.int32), lights, dir4, px, con, hits, np.zeros((50, 3), np.uint32), tex=ts)
    assert a.sum() > 0 and np.array_equal(a, b)
    # an index without a texture samples black: ambient and diffuse vanish, specular stays
    c = O.shade(textured, np.zeros(1, np.int32), lights, dir4, px, con, hits, np.zeros((50, 3), np.uint32))
    assert (c <= a).all() and c.sum() < a.sum()


def test_miss_adds_the_cubemap_colour():
    from rtx import make_materials, make_lights, LIGHT_POINT
    ts = O.TexSet([], cubemap=_faces())
    n = 6
    d = _dirs([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)])
    hits = {"flags": np.zeros(n, np.uint32), "id": np.full(n, -1, np.int32), "n4": np.zeros((n, 4), np.float32),
            "p4": np.zeros((n, 4), np.float32), "uv2": np.zeros((n, 2), np.float32)}
    mats = make_materials([(1, 1, 1, 0, 0, 0)])
    lights = O.lights_from(make_lights((0.1, 0.1, 0.1), [dict(type=LIGHT_POINT, pos=(0.0, 5.0, 0.0))]))
    con = np.float32([1.0, 0.5, 0.25, 1.0, 0.3, 0.7])
    acc = O.shade(mats, np.zeros(1, np.int32), lights, d, np.arange(n, dtype=np.int32), con, hits,
                  np.zeros((n, 3), np.uint32), tex=ts)
    want = np.floor(np.float32([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1]]) *
                    con[:, None] * np.float32(1e6)).astype(np.uint32)
    assert np.array_equal(acc, want)
    # without a cubemap a miss adds nothing
    assert not O.shade(mats, np.zeros(1, np.int32), lights, d, np.arange(n, dtype=np.int32), con, hits,
                       np.zeros((n, 3), np.uint32)).any()
