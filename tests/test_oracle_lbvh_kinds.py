"""Pins the oracle's point and edge LBVHs (no GPU).

* Points: equal to the box LBVH of zero-extent boxes [p, p] (an independent code path: the box
  key point (lo + hi) * 0.5 is p exactly), keys restated in numpy FP32, topology equal to the
  independent top-down builder, every node box the union of its leaves' points.
* Edges: an edge whose end points coincide is its point (the LBVH of such edges equals the point
  LBVH); in general, keys restated in numpy FP32 from the midpoint (p1 + p2) / 2, topology against
  the top-down builder, and every node box the union of its edges' end points (brute force).
"""
import numpy as np
import pytest

from gprt_amd import scenes


def _np_keys(c, lo, hi, orc):
    """Sorted (morton << 32 | primID) from FP32 key points c[n, 3] and the scene box."""
    c = c.astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):  # a flat axis: 0 / 0, which fmax takes to 0
        pt = (c - lo) / (hi - lo)
        q = np.fmin(np.fmax(pt * np.float32(1024.0), np.float32(0.0)), np.float32(1023.0)).astype(np.uint32)
    codes = np.array([orc.lib().orc_morton3(int(a), int(b), int(d)) for a, b, d in q], dtype=np.uint64)
    keys = (codes << np.uint64(32)) | np.arange(len(c), dtype=np.uint64)
    return np.sort(keys)


def _check_boxes_bruteforce(nodes, aabbs, leaf_lo, leaf_hi):
    """Every node's box = min / max over the leaf boxes of its leaf range."""
    n = (len(nodes) + 1) // 2
    for x in range(len(nodes)):
        a = b = x
        while a < n - 1:
            a = nodes[a, 0]
        while b < n - 1:
            b = nodes[b, 1]
        prims = nodes[a:b + 1, 3]
        np.testing.assert_array_equal(aabbs[x, :3], leaf_lo[prims].min(0))
        np.testing.assert_array_equal(aabbs[x, 3:], leaf_hi[prims].max(0))


def _topdown(orc, keys):
    import ctypes as C
    n = len(keys)
    nodes = np.zeros((2 * n - 1, 4), dtype=np.int32)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    orc.lib().orc_topdown(k.ctypes.data_as(C.c_void_p), n, nodes.ctypes.data_as(C.c_void_p))
    return nodes


@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000, 5000])
def test_points_lbvh(orc, n):
    rng = np.random.default_rng(n)
    p = rng.random((n, 3), dtype=np.float32)
    p[: n // 3] = p[0]  # duplicates: equal codes, ordered by primID
    keys, nodes, aabbs = orc.lbvh_build_points(p)
    bk, bn, ba = orc.lbvh_build_boxes(np.concatenate([p, p], 1))
    np.testing.assert_array_equal(keys, bk)
    np.testing.assert_array_equal(nodes, bn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ba.view(np.uint32))
    lo, hi = p.min(0), p.max(0)
    np.testing.assert_array_equal(keys, _np_keys(p, lo, hi, orc))
    np.testing.assert_array_equal(nodes, _topdown(orc, keys))
    if n <= 1000:
        _check_boxes_bruteforce(nodes, aabbs, p, p)


def test_points_t04_extremes(orc):
    # The reference's t04 range: 42 points within +-1e38 (tests/t04-swBVH/hostCode.cpp:39-68).
    rng = np.random.default_rng(4)
    p = (rng.random((42, 3)) * 2e38 - 1e38).astype(np.float32)
    keys, nodes, aabbs = orc.lbvh_build_points(p)
    np.testing.assert_array_equal(keys, _np_keys(p, p.min(0), p.max(0), orc))
    _check_boxes_bruteforce(nodes, aabbs, p, p)


@pytest.mark.parametrize("n", [1, 2, 5, 300, 4000])
def test_edges_lbvh(orc, n):
    rng = np.random.default_rng(100 + n)
    v = rng.random((n + 7, 3), dtype=np.float32)
    e = rng.integers(0, len(v), (n, 2)).astype(np.uint32)
    keys, nodes, aabbs = orc.lbvh_build_edges(v, e)
    p1, p2 = v[e[:, 0]], v[e[:, 1]]
    lo = np.minimum(p1, p2).min(0)
    hi = np.maximum(p1, p2).max(0)
    mid = (p1 + p2) / np.float32(2.0)
    np.testing.assert_array_equal(keys, _np_keys(mid, lo, hi, orc))
    np.testing.assert_array_equal(nodes, _topdown(orc, keys))
    if n <= 300:
        _check_boxes_bruteforce(nodes, aabbs, np.minimum(p1, p2), np.maximum(p1, p2))


def test_degenerate_edges_are_points(orc):
    rng = np.random.default_rng(9)
    p = rng.random((2000, 3), dtype=np.float32)
    e = np.repeat(np.arange(len(p), dtype=np.uint32)[:, None], 2, 1)
    ek, en, ea = orc.lbvh_build_edges(p, e)
    pk, pn, pa = orc.lbvh_build_points(p)
    np.testing.assert_array_equal(ek, pk)
    np.testing.assert_array_equal(en, pn)
    np.testing.assert_array_equal(ea.view(np.uint32), pa.view(np.uint32))


def test_mistaken_midpoint_is_caught(orc):
    """The numpy restatement notices a key point that is not the midpoint (a planted mistake:
    the first end point instead of the midpoint)."""
    rng = np.random.default_rng(10)
    v = rng.random((500, 3), dtype=np.float32)
    e = rng.integers(0, len(v), (400, 2)).astype(np.uint32)
    keys, _, _ = orc.lbvh_build_edges(v, e)
    p1, p2 = v[e[:, 0]], v[e[:, 1]]
    lo, hi = np.minimum(p1, p2).min(0), np.maximum(p1, p2).max(0)
    assert not np.array_equal(keys, _np_keys(p1, lo, hi, orc))
