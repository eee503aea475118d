"""CPU oracle for box primitives (AABB geometry), pinned by closed forms and brute force.

The boxes' LBVH reuses the pinned stages (stable sort, Karras, refit) with a box's corners as its
leaf box and its centre (lo + hi) * 0.5 as its key point; the closest hit is the box surface (entry
distance, or the exit distance from inside). Checks: textbook hits, keys restated in numpy FP32,
the tree against the independent top-down builder, node boxes against per-range unions, and the
brute-force trace against per-box tests.
"""
import numpy as np
import pytest

from gprt_amd import scenes

UNIT = np.array([0, 0, 0, 1, 1, 1], np.float32)
INF = np.float32(3e38)


def random_boxes(n, seed, inverted=0.1, flat=0.05):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    h = (rng.random((n, 3), dtype=np.float32) * np.float32(0.02) + np.float32(1e-3)).astype(np.float32)
    lo, hi = (c - h).astype(np.float32), (c + h).astype(np.float32)
    sw = rng.random((n, 3)) < inverted  # corners given in the wrong order on some axes
    lo2, hi2 = np.where(sw, hi, lo), np.where(sw, lo, hi)
    fl = rng.random((n, 3)) < flat  # zero extent on some axes
    hi2 = np.where(fl, lo2, hi2)
    return np.concatenate([lo2, hi2], 1).astype(np.float32)


def test_box_hit_textbook(orc):
    assert orc.box_hit([-1, 0.5, 0.5, 0, 1, 0, 0, INF], UNIT) == (True, 1.0)   # enters at x = 0
    assert orc.box_hit([0.5, 0.5, 0.5, 0, 1, 0, 0, INF], UNIT) == (True, 0.5)  # inside: the exit
    assert orc.box_hit([-1, 2, 0.5, 0, 1, 0, 0, INF], UNIT)[0] is False        # passes beside
    assert orc.box_hit([-1, 0.5, 0.5, 2.5, 1, 0, 0, INF], UNIT)[0] is False    # tmin beyond the exit
    # tmin exactly at the entry: the open interval tmin < t keeps the exit (t = 2)
    assert orc.box_hit([-1, 0.5, 0.5, 1.0, 1, 0, 0, INF], UNIT) == (True, 2.0)
    # corners given in the wrong order describe the same box
    assert orc.box_hit([-1, 0.5, 0.5, 0, 1, 0, 0, INF], [1, 1, 1, 0, 0, 0]) == (True, 1.0)
    # diagonal ray: the x, y, z slabs are entered at t = 1, 1.25, 0.5 and left at 2, 2.25, 1.5,
    # so the box is entered at 1.25; shifted so the slab intervals do not overlap, it misses
    assert orc.box_hit([-1, -1.25, -0.5, 0, 1, 1, 1, INF], UNIT) == (True, 1.25)
    assert orc.box_hit([-1, -2, -3, 0, 1, 1, 1, INF], UNIT)[0] is False
    # axis-parallel ray outside one slab never hits (|d| = 0 is clamped to 1e-30, no NaN)
    assert orc.box_hit([-1, 1.5, 0.5, 0, 1, 0, 0, INF], UNIT)[0] is False
    # a flat box (zero extent in x) is hit at its plane
    assert orc.box_hit([-1, 0.5, 0.5, 0, 1, 0, 0, INF], [0.25, 0, 0, 0.25, 1, 1]) == (True, 1.25)


def test_box_keys_closed_form(orc):
    b = random_boxes(5000, 1)
    bounds = orc.box_bounds(b)
    corners = np.concatenate([b[:, :3], b[:, 3:]], 0)
    np.testing.assert_array_equal(bounds[:3], corners.min(0))
    np.testing.assert_array_equal(bounds[3:], corners.max(0))
    keys = orc.box_keys(b, bounds)
    c = ((b[:, :3] + b[:, 3:]) * np.float32(0.5)).astype(np.float32)
    pt = ((c - bounds[:3]) / (bounds[3:] - bounds[:3])).astype(np.float32)
    q = np.clip(pt * np.float32(1024.0), 0, 1023).astype(np.uint32)
    codes = np.array([orc.lib().orc_morton3(*map(int, qq)) for qq in q], dtype=np.uint64)
    np.testing.assert_array_equal(keys, (codes << np.uint64(32)) | np.arange(len(b), dtype=np.uint64))


@pytest.mark.parametrize("n", [1, 2, 3, 57, 3000])
def test_box_lbvh_tree_and_boxes(orc, n):
    b = random_boxes(n, 10 + n)
    keys, nodes, aabbs = orc.lbvh_build_boxes(b)
    assert np.all(np.diff(keys.astype(np.float64)) > 0) or n == 1
    if n > 1:
        np.testing.assert_array_equal(nodes, orc.topdown(keys))
    _check_unions(b, nodes, aabbs)


def _check_unions(b, nodes, aabbs):
    """Every node's box = the union of the ordered leaf boxes in its leaf range."""
    n = len(b)
    lo = np.minimum(b[:, :3], b[:, 3:])
    hi = np.maximum(b[:, :3], b[:, 3:])
    leaf_prim = nodes[n - 1:, 3]
    for x in range(2 * n - 1):
        a, e = x, x
        while a < n - 1:
            a = nodes[a, 0]
        while e < n - 1:
            e = nodes[e, 1]
        prims = leaf_prim[a - (n - 1): e - (n - 1) + 1]
        np.testing.assert_array_equal(aabbs[x, :3], lo[prims].min(0))
        np.testing.assert_array_equal(aabbs[x, 3:], hi[prims].max(0))


def test_box_bruteforce_is_the_minimum_of_single_tests(orc):
    b = random_boxes(200, 3)
    rng = np.random.default_rng(4)
    o = rng.random((300, 3), dtype=np.float32) * np.float32(1.2) - np.float32(0.1)
    d = rng.normal(size=(300, 3)).astype(np.float32)
    rays = np.zeros(300, dtype=scenes.RAY_DTYPE)
    rays["ox"], rays["oy"], rays["oz"] = o[:, 0], o[:, 1], o[:, 2]
    rays["dx"], rays["dy"], rays["dz"] = d[:, 0], d[:, 1], d[:, 2]
    rays["tmin"], rays["tmax"] = 0.0, scenes.TMAX
    hits = orc.trace_boxes_bruteforce(rays, b, threads=2)
    for i in range(len(rays)):
        r8 = [rays[f][i] for f in ("ox", "oy", "oz", "tmin", "dx", "dy", "dz", "tmax")]
        best_t, best = np.float32(rays["tmax"][i]), -1
        for p in range(len(b)):
            hit, t = orc.box_hit(r8, b[p])
            if hit and (np.float32(t) < best_t or (np.float32(t) == best_t and (best < 0 or p < best))):
                best_t, best = np.float32(t), p
        assert hits["primID"][i] == best
        assert hits["t"][i] == best_t


def test_box_refit_keeps_topology_and_unions(orc):
    b = random_boxes(2000, 5)
    _, nodes, aabbs = orc.lbvh_build_boxes(b)
    np.testing.assert_array_equal(orc.box_refit(b, nodes).view(np.uint32), aabbs.view(np.uint32))
    rng = np.random.default_rng(6)
    moved = (b + np.repeat(rng.random((len(b), 3), dtype=np.float32) * np.float32(0.2), 2, axis=0)
             .reshape(-1, 6)).astype(np.float32)
    _check_unions(moved, nodes, orc.box_refit(moved, nodes))
