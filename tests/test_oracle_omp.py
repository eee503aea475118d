"""The all-cores CPU baseline (oracle/baseline_omp.cpp) computes exactly what the serial oracle
computes: the parallel radix sort equals std::stable_sort (and numpy's stable argsort), the
blocked scan equals the serial scan, the parallel LBVH builds (triangles and boxes) equal the
serial builds (keys, topology, every box), the sphere bounds equal numpy's, the per-mesh batch equals one build per mesh, and the multi-mesh trace
equals the minimum of (t, mesh, primID) over per-mesh traces. Thread counts that do not divide
the input are included."""
import numpy as np
import pytest

from gprt_amd import scenes


@pytest.mark.parametrize("n", [0, 1, 2, 7, 1000, 100_003])
@pytest.mark.parametrize("kbits,vbits", [(32, 32), (64, 0), (64, 64), (32, 0), (64, 32)])
def test_par_radix_sort_equals_stable_sort(orc, n, kbits, vbits):
    rng = np.random.default_rng(n + kbits + vbits)
    kt = np.uint32 if kbits == 32 else np.uint64
    keys = rng.integers(0, 1 << 20, n, dtype=np.uint64).astype(kt)  # many duplicates: stability
    if kbits == 64:
        keys = (keys << np.uint64(40)) | rng.integers(0, 4, n, dtype=np.uint64)
    vals = None if vbits == 0 else np.arange(n, dtype=np.uint32 if vbits == 32 else np.uint64)
    for threads in (1, 3, 8):
        k, v = orc.par_radix_sort(keys, vals, threads=threads)
        rk, rv = orc.stable_sort(keys, vals)
        np.testing.assert_array_equal(k, rk)
        if vals is not None:
            np.testing.assert_array_equal(v, rv)
            np.testing.assert_array_equal(v, np.argsort(keys, kind="stable").astype(vals.dtype))


@pytest.mark.parametrize("n", [0, 1, 5, 9973 * 3, 1 << 20])
def test_par_exclusive_sum_equals_serial(orc, n):
    x = (scenes.rand_u64(5, 1, n) >> np.uint64(33)).astype(np.uint32)  # large values: wraps mod 2^32
    for threads in (1, 5, 8):
        out, total = orc.par_exclusive_sum(x, threads=threads)
        ref, rtotal = orc.exclusive_sum(x)
        np.testing.assert_array_equal(out, ref)
        assert total == rtotal


@pytest.mark.parametrize("n,kind", [(1, "u"), (2, "u"), (3, "u"), (1000, "u"), (50_000, "blob"), (200_001, "blob")])
def test_par_lbvh_build_equals_serial(orc, n, kind):
    v, idx = scenes.uniform_scene(n, seed=31) if kind == "u" else scenes.blob_scene(n, seed=31)
    rk, rn, ra = orc.lbvh_build(v, idx)
    for threads in (1, 7):
        k, nodes, aabbs = orc.par_lbvh_build(v, idx, threads=threads)
        np.testing.assert_array_equal(k, rk)
        np.testing.assert_array_equal(nodes, rn)
        np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))
    # The refit alone over the kept topology, after the vertices move.
    moved = (v + scenes.rand01(9, 0, v.size).reshape(v.shape).astype(np.float32) * np.float32(0.01)).astype(np.float32)
    np.testing.assert_array_equal(orc.par_refit(moved, idx, rn, threads=5).view(np.uint32),
                                  orc.refit(moved, idx, rn).view(np.uint32))


@pytest.mark.parametrize("n", [1, 2, 3, 1000, 65_537])
def test_par_box_lbvh_build_equals_serial(orc, n):
    # Boxes given with some corners swapped (lo > hi on an axis): the leaf box is the ordered one.
    rng = np.random.default_rng(n)
    c = rng.random((n, 3), dtype=np.float32)
    e = (rng.random((n, 3), dtype=np.float32) * np.float32(0.01)).astype(np.float32)
    boxes = np.concatenate([c - e, c + e], 1).astype(np.float32)
    boxes[::5, [0, 3]] = boxes[::5, [3, 0]]
    rk, rn, ra = orc.lbvh_build_boxes(boxes)
    for threads in (1, 6):
        k, nodes, aabbs = orc.par_lbvh_build_boxes(boxes, threads=threads)
        np.testing.assert_array_equal(k, rk)
        np.testing.assert_array_equal(nodes, rn)
        np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))


def test_par_sphere_bounds_equals_numpy(orc):
    rng = np.random.default_rng(3)
    centres = (rng.random((10_001, 3), dtype=np.float32) * np.float32(100.0)).astype(np.float32)
    radii = rng.random(10_001, dtype=np.float32)
    ref = np.concatenate([centres - radii[:, None], centres + radii[:, None]], 1)
    for threads in (1, 7):
        np.testing.assert_array_equal(orc.par_sphere_bounds(centres, radii, threads=threads).view(np.uint32),
                                      ref.view(np.uint32))


def test_par_lbvh_build_duplicate_codes(orc):
    # Many triangles share one Morton code (a flat, tiny cluster): ties are ordered by primID.
    v, idx = scenes.uniform_scene(5000, seed=32)
    v = v.copy()
    v[: 3 * 2500] *= np.float32(1e-6)
    rk, rn, ra = orc.lbvh_build(v, idx)
    k, nodes, aabbs = orc.par_lbvh_build(v, idx, threads=6)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))


def test_build_many_and_multi_trace(orc):
    sizes = np.array([64, 1, 3000, 777, 20000], dtype=np.int64)
    meshes = [scenes.uniform_scene(int(n), seed=40 + i) for i, n in enumerate(sizes)]
    verts = np.concatenate([m[0] for m in meshes])
    idx = np.concatenate([m[1] for m in meshes])
    keys, nodes, aabbs = orc.lbvh_build_many(verts, idx, sizes, threads=4)
    first = np.concatenate([[0], np.cumsum(sizes)])
    for m, (v, ix) in enumerate(meshes):
        rk, rn, ra = orc.lbvh_build(v, ix)
        o = 2 * first[m] - m
        np.testing.assert_array_equal(keys[first[m]: first[m + 1]], rk)
        np.testing.assert_array_equal(nodes[o: o + 2 * sizes[m] - 1], rn)
        np.testing.assert_array_equal(aabbs[o: o + 2 * sizes[m] - 1].view(np.uint32), ra.view(np.uint32))
    rays = np.concatenate([scenes.pinhole_rays(48, 48), scenes.random_rays(2000, 41)])
    hits, inst = orc.trace_bvh_multi(rays, verts, idx, sizes, nodes, aabbs, threads=5)
    # Reference: brute force per mesh, then the minimum of (t, mesh, primID).
    best = np.zeros(len(rays), dtype=orc.HIT_DTYPE)
    best["t"] = rays["tmax"]
    best["primID"] = -1
    bi = np.full(len(rays), -1, np.int32)
    for m, (v, ix) in enumerate(meshes):
        h = orc.trace_bruteforce(rays, v, ix)
        better = (h["primID"] >= 0) & ((best["primID"] < 0) | (h["t"] < best["t"]))
        best[better] = h[better]
        bi[better] = m
    np.testing.assert_array_equal(inst, bi)
    np.testing.assert_array_equal(hits.view(np.uint32).reshape(-1, 4), best.view(np.uint32).reshape(-1, 4))
    assert (inst >= 0).mean() > 0.1 and len(np.unique(inst)) >= 5


def test_host_info(orc):
    info = orc.host_info()
    assert info["nproc"] >= 1 and info["omp_max_threads"] >= 1
