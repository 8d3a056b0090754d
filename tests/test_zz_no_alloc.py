"""The allocation-free LBVH build (rtx_lbvh_build_noalloc) against
rtx_lbvh_build, through the C ABI: the same sorted codes, permutation,
permuted index buffer and material map, topology, node AABBs, level counts
and global AABB, bit for bit, on two calls in a row (the second with fewer
triangles, so the shared scratch array holds the first call's contents), on
a chained rebuild, at an odd sort-pass count, and with the size and state
errors the entry point documents.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _soup(seed, T):
    rng = np.random.default_rng(seed)
    V = max(3, T // 2)
    pos = rng.uniform(-10, 10, (V, 3)).astype(np.float32)
    idx = rng.integers(0, V, (T, 3)).astype(np.uint32)
    mm = rng.integers(0, 5, T).astype(np.int32)
    return pos, idx, mm


def _tree(s):
    n, lc = s.info()
    g = s.nodes()
    lo, hi = s.global_bbox()
    return dict(codes=s.codes(), perm=s.perm(), idx=s.indices(), mm=s.matmap(), n=n, lc=lc, a=g["a"], b=g["b"],
                parent=g["parent"], leaf=g["leaf"] != 0, lo=g["lo"][:, :3], hi=g["hi"][:, :3], glo=lo, ghi=hi)


def _same(x, y, what):
    for k in x:
        assert np.array_equal(np.asarray(x[k]), np.asarray(y[k])), f"{what}: {k} differs"


def _reference(ctx, pos, idx, mm, depth, builds=1):
    import rtx
    s = rtx.Scene(ctx)
    s.set_geometry(pos, idx, None, None, mm)
    for _ in range(builds):
        s.build(depth, 1)
    t = _tree(s)
    s.destroy()
    return t


def test_noalloc_equals_build_two_calls_second_smaller(ctx):
    import rtx
    big = _soup(1, 300_000)
    small = _soup(2, 70_000)
    s = rtx.Scene(ctx)
    try:
        # reserve the tree buffers with one ordinary build, then reload the input
        s.set_geometry(big[0], big[1], None, None, big[2])
        s.build(30, 1)
        s.set_geometry(big[0], big[1], None, None, big[2])
        s.build_noalloc(30, 1)
        _same(_tree(s), _reference(ctx, *big, 30), "first call (300k triangles)")
        # second call, smaller input: the scratch keeps the first call's contents
        s.set_geometry(small[0], small[1], None, None, small[2])
        s.build_noalloc(30, 1)
        _same(_tree(s), _reference(ctx, *small, 30), "second call (70k triangles)")
        # a rebuild continues from the permuted buffers, as rtx_lbvh_build does
        s.build_noalloc(30, 1)
        _same(_tree(s), _reference(ctx, *small, 30, builds=2), "chained rebuild")
    finally:
        s.destroy()


@pytest.mark.parametrize("depth", [24, 8])
def test_noalloc_odd_pass_count(ctx, depth):
    # 3 and 1 radix passes leave the sorted pairs in the scratch's alternates
    import rtx
    soup = _soup(3 + depth, 50_000)
    s = rtx.Scene(ctx)
    try:
        s.set_geometry(soup[0], soup[1], None, None, soup[2])
        s.build(depth, 1)
        s.set_geometry(soup[0], soup[1], None, None, soup[2])
        s.build_noalloc(depth, 1)
        _same(_tree(s), _reference(ctx, *soup, depth), f"depth {depth}")
    finally:
        s.destroy()


def test_noalloc_matches_oracle_small(ctx):
    # the tree itself is the oracle's (canonical order), not only the build's
    import rtx
    import oracle as O
    pos, idx, mm = _soup(9, 5000)
    s = rtx.Scene(ctx)
    try:
        s.set_geometry(pos, idx, None, None, mm)
        s.build(30, 1)
        s.set_geometry(pos, idx, None, None, mm)
        s.build_noalloc(30, 1)
        b = O.Build(pos, idx, mm, 30, 1)
        g = s.nodes()
        assert np.array_equal(s.codes().astype(np.uint64), b.codes)
        assert np.array_equal(g["a"], b.nodes["a"]) and np.array_equal(g["b"], b.nodes["b"])
        assert np.array_equal(g["lo"][:, :3], b.nodes["lo"][:, :3])
        assert np.array_equal(g["hi"][:, :3], b.nodes["hi"][:, :3])
    finally:
        s.destroy()


def test_noalloc_errors(ctx):
    import rtx
    s = rtx.Scene(ctx)
    try:
        pos, idx, mm = _soup(4, 1000)
        s.set_geometry(pos, idx, None, None, mm)
        with pytest.raises(rtx.RtxError, match="not reserved"):
            s.build_noalloc(30, 1)  # no earlier build: tree buffers absent
        s.build(30, 1)
        pos, idx, mm = _soup(5, 2000)
        s.set_geometry(pos, idx, None, None, mm)
        with pytest.raises(rtx.RtxError, match="not reserved"):
            s.build_noalloc(30, 1)  # reserved for 1000 triangles only
        V = 1000
        pos = np.random.default_rng(6).uniform(-1, 1, (V, 3)).astype(np.float32)
        idx = np.random.default_rng(7).integers(0, V, (1_048_577, 3)).astype(np.uint32)
        s.set_geometry(pos, idx)
        with pytest.raises(rtx.RtxError, match="RTX_LBVH_NOALLOC_MAX_TRIS"):
            s.build_noalloc(30, 1)
    finally:
        s.destroy()
