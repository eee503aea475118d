"""Pins the CPU oracle (no GPU). Every check here is independent of the oracle's own code:
known-answer values, closed forms of the reference's t03 patterns, numpy implementations,
brute force on small inputs, an independent tree builder, and FP64 geometry.

The reference's own tests hold no golden vectors for this path (t02/t04 draw from
std::random_device; t03 is closed-form; t05 is empty), so these checks are the pin.
"""
import numpy as np
import pytest

from gprt_amd import scenes


# ------------------------------------------------------------------------------------------------
# Sort (tests/t02-bufferSort/hostCode.cpp semantics, stable)
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("kdt,vdt", [(np.uint32, None), (np.uint32, np.uint32), (np.uint64, np.uint64),
                                     (np.uint64, None), (np.uint64, np.uint32), (np.uint32, np.uint64)])
def test_sort_matches_numpy_stable(orc, kdt, vdt):
    rng = np.random.default_rng(7)
    n = 20000
    keys = rng.integers(0, 500, n).astype(kdt)  # many duplicates: stability matters
    vals = None if vdt is None else np.arange(n).astype(vdt)
    k, v = orc.stable_sort(keys, vals)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    if vals is not None:
        np.testing.assert_array_equal(v, vals[order])


def test_sort_bruteforce_tiny(orc):
    # insertion sort (stable by construction) on every length 0..12
    rng = np.random.default_rng(3)
    for n in range(13):
        keys = rng.integers(0, 4, n).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        pairs = list(zip(keys.tolist(), vals.tolist()))
        for i in range(1, n):
            j = i
            while j > 0 and pairs[j - 1][0] > pairs[j][0]:
                pairs[j - 1], pairs[j] = pairs[j], pairs[j - 1]
                j -= 1
        k, v = orc.stable_sort(keys, vals)
        assert list(zip(k.tolist(), v.tolist())) == pairs


def test_sort_t02_unique_shuffled(orc):
    # t02 (b): keys 0..N-1 shuffled with random payloads; sorted keys are the identity.
    rng = np.random.default_rng(11)
    n = 100000
    keys = rng.permutation(n).astype(np.uint32)
    vals = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    k, v = orc.stable_sort(keys, vals)
    np.testing.assert_array_equal(k, np.arange(n, dtype=np.uint32))
    inv = np.empty(n, dtype=np.int64)
    inv[keys] = np.arange(n)
    np.testing.assert_array_equal(v, vals[inv])


# ------------------------------------------------------------------------------------------------
# Scan family (tests/t03-bufferScans/hostCode.cpp)
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("rnd", [0, 1, 37, 99])
def test_exclusive_sum_t03_all_ones(orc, rnd):
    n = 9973 * (rnd + 2)
    out, total = orc.exclusive_sum(np.ones(n, dtype=np.uint32))
    assert total == n
    np.testing.assert_array_equal(out, np.arange(n, dtype=np.uint32))


def test_exclusive_sum_wraps(orc):
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2**32, 300000, dtype=np.uint64).astype(np.uint32)
    out, total = orc.exclusive_sum(x)
    ref = np.concatenate([np.zeros(1, np.uint64), np.cumsum(x.astype(np.uint64))[:-1]]) & np.uint64(0xFFFFFFFF)
    np.testing.assert_array_equal(out, ref.astype(np.uint32))
    assert total == int(x.astype(np.uint64).sum() & 0xFFFFFFFF)


def _t03_pattern(n):
    i = np.arange(n, dtype=np.int64)
    return (i * np.where(i & 1, 1, -1)).astype(np.int32)


@pytest.mark.parametrize("sel_pos", [True, False])
def test_partition_t03_closed_form(orc, sel_pos):
    n = 9973 * 3
    x = _t03_pattern(n)
    out, total = orc.partition(x, sel_pos)
    mask = (x >= 0) if sel_pos else (x < 0)
    expect = np.concatenate([x[mask], x[~mask][::-1]])  # selected in order, rest reversed at the tail
    assert total == int(mask.sum())
    np.testing.assert_array_equal(out, expect)


@pytest.mark.parametrize("sel_pos", [True, False])
def test_select_t03_closed_form(orc, sel_pos):
    n = 9973 * 4
    x = _t03_pattern(n)
    out, total = orc.select(x, sel_pos)
    mask = (x >= 0) if sel_pos else (x < 0)
    assert total == int(mask.sum())
    np.testing.assert_array_equal(out[:total], x[mask])


# ------------------------------------------------------------------------------------------------
# Morton keys (gprt/development/gprt_lbvh.hlsl:25-38, :262-277)
# ------------------------------------------------------------------------------------------------
def test_morton_known_answers(orc):
    assert orc.lib().orc_separate_bits(0) == 0
    assert orc.lib().orc_separate_bits(1) == 1
    assert orc.lib().orc_separate_bits(2) == 8
    assert orc.lib().orc_separate_bits(0x3FF) == 0x09249249
    assert orc.lib().orc_separate_bits(0x7FF) == 0x09249249  # only 10 bits are used
    assert orc.lib().orc_morton3(1, 0, 0) == 1
    assert orc.lib().orc_morton3(0, 1, 0) == 2
    assert orc.lib().orc_morton3(0, 0, 1) == 4
    assert orc.lib().orc_morton3(1023, 1023, 1023) == 0x3FFFFFFF


def _interleave(x, y, z):
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return code


def test_morton_bitwise(orc):
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1024, (500, 3)):
        assert orc.lib().orc_morton3(int(x), int(y), int(z)) == _interleave(int(x), int(y), int(z))


def test_morton_keys_numpy(orc):
    verts, idx = scenes.uniform_scene(2000, seed=9)
    b = orc.scene_bounds(verts, idx)
    np.testing.assert_array_equal(b[:3], verts.min(0))
    np.testing.assert_array_equal(b[3:], verts.max(0))
    keys = orc.morton_keys(verts, idx, b)
    p = verts[idx]  # [n, 3, 3] float32
    c = (p[:, 0] + p[:, 1] + p[:, 2]) / np.float32(3.0)
    q = np.clip((c - b[:3]) / (b[3:] - b[:3]) * np.float32(1024.0), np.float32(0), np.float32(1023)).astype(np.uint32)
    codes = np.array([_interleave(*map(int, r)) for r in q], dtype=np.uint64)
    np.testing.assert_array_equal(keys >> np.uint64(32), codes)
    np.testing.assert_array_equal(keys & np.uint64(0xFFFFFFFF), np.arange(len(idx), dtype=np.uint64))


def test_morton_quantisation_edges(orc):
    # a centroid at the box minimum maps to 0, at the maximum to 1023 on every axis
    verts = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 2, 4], [1, 2, 4], [1, 2, 4]], dtype=np.float32)
    idx = np.array([[0, 1, 2], [3, 4, 5]], dtype=np.uint32)
    b = orc.scene_bounds(verts, idx)
    keys = orc.morton_keys(verts, idx, b)
    assert int(keys[0] >> np.uint64(32)) == 0
    assert int(keys[1] >> np.uint64(32)) == 0x3FFFFFFF


# ------------------------------------------------------------------------------------------------
# Karras topology and refit
# ------------------------------------------------------------------------------------------------
def _check_tree(nodes, n):
    # every node except the root has exactly one parent, and the parent lists it as a child
    parents = nodes[:, 2]
    assert parents[0] == -1
    for x in range(1, 2 * n - 1):
        p = parents[x]
        assert p >= 0 and x in (nodes[p, 0], nodes[p, 1])
    leaves = nodes[n - 1:, 3]
    assert sorted(leaves.tolist()) == list(range(n)) or len(set(leaves.tolist())) == n


@pytest.mark.parametrize("n", [2, 3, 5, 64, 1000, 16384])
def test_karras_equals_topdown(orc, n):
    verts, idx = scenes.uniform_scene(n, seed=100 + n)
    keys = np.sort(orc.morton_keys(verts, idx, orc.scene_bounds(verts, idx)))
    a = orc.karras(keys)
    b = orc.topdown(keys)
    np.testing.assert_array_equal(a, b)
    _check_tree(a, n)


def test_karras_with_morton_collisions(orc):
    # many triangles share a Morton code: ties are ordered by primID in the low 32 bits
    rng = np.random.default_rng(2)
    codes = rng.integers(0, 16, 3000).astype(np.uint64)
    keys = np.sort((codes << np.uint64(32)) | np.arange(3000, dtype=np.uint64))
    np.testing.assert_array_equal(orc.karras(keys), orc.topdown(keys))


def test_refit_equals_bruteforce(orc):
    verts, idx = scenes.blob_scene(4096, seed=5)
    keys, nodes, aabbs = orc.lbvh_build(verts, idx)
    ref = orc.refit(verts, idx, nodes, bruteforce=True)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ref.view(np.uint32))
    # root box equals the scene box
    b = orc.scene_bounds(verts, idx)
    np.testing.assert_array_equal(aabbs[0], b)


def test_lbvh_single_triangle(orc):
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    idx = np.array([[0, 1, 2]], dtype=np.uint32)
    keys, nodes, aabbs = orc.lbvh_build(verts, idx)
    assert nodes.tolist() == [[-1, -1, -1, 0]]
    np.testing.assert_array_equal(aabbs[0], [0, 0, 0, 1, 1, 0])


# ------------------------------------------------------------------------------------------------
# Ray / triangle and traversal
# ------------------------------------------------------------------------------------------------
def _ray(o, d, tmin=0.0, tmax=1e30):
    return np.array([*o, tmin, *d, tmax], dtype=np.float32)


def test_triangle_known_hits(orc):
    v0, v1, v2 = [-1, -1, 1], [3, -1, 1], [-1, 3, 1]
    ok, t, u, v = orc.intersect(_ray([0, 0, -1], [0, 0, 1]), v0, v1, v2)
    assert ok and t == 2.0 and abs(u - 0.25) < 1e-7 and abs(v - 0.25) < 1e-7
    # reversed direction: behind the origin -> miss (t must exceed tmin)
    ok, *_ = orc.intersect(_ray([0, 0, -1], [0, 0, -1]), v0, v1, v2)
    assert not ok
    # outside the triangle
    ok, *_ = orc.intersect(_ray([2.5, 2.5, -1], [0, 0, 1]), v0, v1, v2)
    assert not ok
    # open interval: a hit exactly at tmin is rejected
    ok, *_ = orc.intersect(_ray([0, 0, -1], [0, 0, 1], tmin=2.0), v0, v1, v2)
    assert not ok
    # vertex hit (edge function exactly zero -> FP64 re-evaluation) is accepted
    ok, t, *_ = orc.intersect(_ray([-1, -1, -1], [0, 0, 1]), v0, v1, v2)
    assert ok and t == 2.0
    # barycentric weights: hitting v1 exactly gives u = 1, v = 0
    ok, t, u, v = orc.intersect(_ray([3, -1, 0], [0, 0, 1]), v0, v1, v2)
    assert ok and u == 1.0 and v == 0.0


def test_triangle_t_against_fp64(orc):
    n = 256
    cent = scenes.rand01(4, 0, 3 * n).reshape(n, 3)
    off = scenes.rand01(4, 1, 9 * n) * 0.4 - 0.2  # large triangles: many hits
    verts = (cent[:, None, :] + off.reshape(n, 3, 3)).astype(np.float32).reshape(-1, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    rays = scenes.random_rays(300, seed=8, lo=0.2, hi=0.8)
    r8 = rays.view(np.float32).reshape(-1, 8)
    errs = []
    for i in range(len(r8)):
        for p in range(len(idx)):
            v = verts[idx[p]]
            ok, t, u, w = orc.intersect(r8[i], v[0], v[1], v[2])
            if ok:
                t64 = orc.plane_t_fp64(r8[i], v[0], v[1], v[2])
                errs.append(abs(t - t64) / abs(t64))
                assert u >= 0 and w >= 0 and u + w <= 1 + 1e-6
    errs = np.array(errs)
    assert len(errs) > 20
    # FP32 is well conditioned for typical hits; grazing hits lose a few more digits.
    assert np.median(errs) < 1e-6, np.median(errs)
    assert errs.max() < 1e-3, errs.max()


def test_bvh_traversal_equals_bruteforce(orc):
    verts, idx = scenes.uniform_scene(16384, seed=1)
    keys, nodes, aabbs = orc.lbvh_build(verts, idx)
    rays = scenes.pinhole_rays(64, 64)
    brute = orc.trace_bruteforce(rays, verts, idx)
    bvh, counters = orc.trace_bvh(rays, verts, idx, nodes, aabbs)
    np.testing.assert_array_equal(bvh["primID"], brute["primID"])
    np.testing.assert_array_equal(bvh["t"].view(np.uint32), brute["t"].view(np.uint32))
    assert (brute["primID"] >= 0).mean() > 0.1
    assert counters[0] > 0 and counters[1] > 0


def test_bvh_traversal_random_rays(orc):
    verts, idx = scenes.blob_scene(8192, seed=12, offset=0.01)
    keys, nodes, aabbs = orc.lbvh_build(verts, idx)
    rays = scenes.random_rays(3000, seed=13)
    brute = orc.trace_bruteforce(rays, verts, idx)
    bvh, _ = orc.trace_bvh(rays, verts, idx, nodes, aabbs, threads=4)
    np.testing.assert_array_equal(bvh["primID"], brute["primID"])
    np.testing.assert_array_equal(bvh["t"].view(np.uint32), brute["t"].view(np.uint32))
