"""Pins the oracle's LBVH hierarchy (process_task / find_middle,
src/kernel/bvh-builder.cl:484-671) and refit without trusting it:

* the reference's invariants (test_node, src/rt/bvh-builder.cpp:23-114):
  every primitive in exactly one leaf, every node reached once, children
  partition the parent's range, depth bounded by bvh_depth;
* an independent construction: a Karras binary radix tree over the sorted
  codes (keys made unique by position), truncated where the reference makes
  leaves (count <= min_leaf + 2, or equal first/last codes), numbered in
  level order — must equal the oracle's tree node for node;
* node AABBs against brute-force unions of the triangles in each subtree.
"""
from collections import deque

import numpy as np
import pytest

import oracle as O


def rt_delta(codes, i, j):
    T = len(codes)
    if j < 0 or j >= T:
        return -1
    x = int(codes[i]) ^ int(codes[j])
    if x:
        return 64 - x.bit_length()
    return 64 + (64 - (i ^ j).bit_length())


def karras(codes):
    """Karras 2012 radix tree: for internal node i, (j, gamma)."""
    T = len(codes)
    out = {}
    for i in range(T - 1):
        d = 1 if rt_delta(codes, i, i + 1) - rt_delta(codes, i, i - 1) > 0 else -1
        dmin = rt_delta(codes, i, i - d)
        lmax = 2
        while rt_delta(codes, i, i + lmax * d) > dmin:
            lmax *= 2
        l, t = 0, lmax // 2
        while t >= 1:
            if rt_delta(codes, i, i + (l + t) * d) > dmin:
                l += t
            t //= 2
        j = i + l * d
        dnode = rt_delta(codes, i, j)
        s, t = 0, l
        while True:
            t = (t + 1) // 2
            if rt_delta(codes, i, i + (s + t) * d) > dnode:
                s += t
            if t <= 1:
                break
        gamma = i + s * d + min(d, 0)
        out[i] = (j, gamma)
    return out


def level_order_tree(codes, min_leaf):
    """Level-order (BFS) numbering of the truncated radix tree."""
    kt = karras(codes)
    rng_to_split = {}
    for i, (j, g) in kt.items():
        rng_to_split[(min(i, j), max(i, j))] = g + 1
    T = len(codes)
    nodes = [(0, T, 0)]  # (start, end, parent)
    a, b, leaf = [], [], []
    q = deque([0])
    while q:
        n = q.popleft()
        s, e, _ = nodes[n]
        cnt = e - s
        is_leaf = ((cnt - 2) & 0xFFFFFFFF) <= min_leaf or codes[s] == codes[e - 1]
        a.append(None), b.append(None), leaf.append(is_leaf)
        if not is_leaf:
            m = rng_to_split[(s, e - 1)]
            c = len(nodes)
            nodes.append((s, m, n))
            nodes.append((m, e, n))
            q.append(c)
            q.append(c + 1)
    # assign a/b after all nodes exist
    child = {}
    for idx, (s, e, p) in enumerate(nodes[1:], start=1):
        child.setdefault(p, []).append(idx)
    A = np.zeros(len(nodes), np.uint32)
    B = np.zeros(len(nodes), np.uint32)
    P = np.zeros(len(nodes), np.uint32)
    for n, (s, e, p) in enumerate(nodes):
        P[n] = p
        if leaf[n]:
            A[n], B[n] = s, e
        else:
            A[n], B[n] = child[n]
    return A, B, P, np.array(leaf)


def scene(seed, T, V=None, dup_every=0):
    rng = np.random.default_rng(seed)
    V = V or max(3, T)
    pos = (rng.uniform(-5, 5, (V, 3)).astype(np.float32) + np.float32(0.0))
    idx = rng.integers(0, V, (T, 3)).astype(np.uint32)
    if dup_every:
        idx[::dup_every] = idx[0]
    return pos, idx


@pytest.mark.parametrize("T,seed,depth,min_leaf,dup", [
    (1, 1, 30, 1, 0), (2, 2, 30, 1, 0), (5, 3, 30, 1, 0), (64, 4, 30, 1, 0), (1000, 5, 30, 1, 0),
    (3000, 6, 30, 4, 0), (2000, 7, 12, 1, 0), (2000, 8, 30, 1, 3), (1500, 9, 9, 2, 0), (800, 10, 32, 1, 5),
])
def test_hierarchy_equals_truncated_radix_tree(T, seed, depth, min_leaf, dup):
    pos, idx = scene(seed, T, dup_every=dup)
    b = O.Build(pos, idx, None, depth, min_leaf)
    A, B, P, leaf = level_order_tree(b.codes, min_leaf)
    assert b.node_count == len(A)
    assert np.array_equal(b.nodes["a"], A) and np.array_equal(b.nodes["b"], B)
    assert np.array_equal(b.nodes["parent"][1:], P[1:])
    assert np.array_equal(b.nodes["leaf"] != 0, leaf)


@pytest.mark.parametrize("T,seed,depth,min_leaf", [(3000, 21, 30, 1), (5000, 22, 16, 3), (700, 23, 30, 8)])
def test_reference_invariants_and_bounds(T, seed, depth, min_leaf):
    pos, idx = scene(seed, T)
    b = O.Build(pos, idx, None, depth, min_leaf)
    nd = b.nodes
    N = b.node_count
    seen_prim = np.zeros(T, int)
    seen_node = np.zeros(N, int)
    tri = pos[b.idx.reshape(-1, 3)]  # sorted triangles' vertices (T, 3, 3)
    stack = [(0, 0)]
    maxdepth = 0
    while stack:
        n, dep = stack.pop()
        seen_node[n] += 1
        maxdepth = max(maxdepth, dep)
        lo, hi = nd["lo"][n][:3], nd["hi"][n][:3]
        if nd["leaf"][n]:
            s, e = nd["a"][n], nd["b"][n]
            assert s < e <= T
            seen_prim[s:e] += 1
            v = tri[s:e].reshape(-1, 3)
            assert np.array_equal(lo, v.min(0)) and np.array_equal(hi, v.max(0))
        else:
            l, r = nd["a"][n], nd["b"][n]
            assert r == l + 1 and nd["parent"][l] == n and nd["parent"][r] == n
            # children partition the parent's range: collect ranges
            stack.append((l, dep + 1))
            stack.append((r, dep + 1))
            assert np.array_equal(lo, np.minimum(nd["lo"][l][:3], nd["lo"][r][:3]))
            assert np.array_equal(hi, np.maximum(nd["hi"][l][:3], nd["hi"][r][:3]))
    assert (seen_prim == 1).all(), "every primitive exactly once"
    assert (seen_node == 1).all(), "every node reached exactly once"
    assert maxdepth <= depth, "tree depth bounded by bvh_depth"
    # level blocks: node counts per iteration sum to the node count
    assert 1 + int(b.level_counts[:depth].sum()) == N


def test_split_is_first_index_with_highest_differing_bit():
    pos, idx = scene(31, 4000)
    b = O.Build(pos, idx, None, 30, 1)
    nd, c = b.nodes, b.codes
    N = b.node_count
    rs, re_ = np.zeros(N, np.int64), np.zeros(N, np.int64)
    for n in range(N - 1, -1, -1):  # children follow their parent in level order
        if nd["leaf"][n]:
            rs[n], re_[n] = nd["a"][n], nd["b"][n]
        else:
            rs[n], re_[n] = rs[nd["a"][n]], re_[nd["b"][n]]
    for n in range(N):
        if nd["leaf"][n]:
            continue
        l, r = nd["a"][n], nd["b"][n]
        s, m, e = rs[n], rs[r], re_[n]
        assert re_[l] == m
        x = int(c[s]) ^ int(c[e - 1])
        bit = x.bit_length() - 1
        assert (int(c[m - 1]) >> bit) & 1 == 0 and (int(c[m]) >> bit) & 1 == 1
        assert s < m < e
