"""GPU parity of the LBVH build and closest-hit traversal against the CPU oracle (C-ABI calls).

LBVH: sorted keys (morton30<<32 | primID), topology in the reference's int4 node layout
(gprt/development/gprt_lbvh.h:28-55) and every node box must be bit-identical to the oracle's
restatement of gprt/development/gprt_lbvh.hlsl. Trace: primID exact and t / u / v bit-identical
to the oracle's brute force (every triangle tested) on the same FP32 rays.
"""
import numpy as np
import pytest

from gprt_amd import api, scenes

pytestmark = pytest.mark.gpu


def _build(ctx, verts, idx, mode=api.GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE):
    acc = api.TriangleAccel(ctx, verts, idx)
    acc.build(mode)
    assert api.last_error() is None, api.last_error()
    return acc


def _check_lbvh(ctx, orc, verts, idx):
    acc = _build(ctx, verts, idx)
    keys, nodes, aabbs = acc.export()
    rk, rn, ra = orc.lbvh_build(verts, idx)
    np.testing.assert_array_equal(keys, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))
    return acc


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 4097, 16384])
def test_lbvh_uniform(ctx, orc, n):
    verts, idx = scenes.uniform_scene(n, seed=1 if n == 16384 else 50 + n)
    _check_lbvh(ctx, orc, verts, idx).destroy()


def test_lbvh_blobs(ctx, orc):
    verts, idx = scenes.blob_scene(200_000, seed=2)
    _check_lbvh(ctx, orc, verts, idx).destroy()


def test_lbvh_morton_collisions(ctx, orc):
    # 3000 copies of 5 triangles: identical centroids, keys differ only in primID
    base, _ = scenes.uniform_scene(5, seed=3)
    verts = np.tile(base.reshape(5, 3, 3), (600, 1, 1)).reshape(-1, 3)
    idx = np.arange(len(verts), dtype=np.uint32).reshape(-1, 3)
    _check_lbvh(ctx, orc, verts, idx).destroy()


def test_lbvh_shared_vertices_and_flat_axis(ctx, orc):
    # indexed grid in the z=0.5 plane: shared vertices, zero extent on z (0/0 quantisation)
    g = 40
    ys, xs = np.meshgrid(np.arange(g + 1), np.arange(g + 1), indexing="ij")
    verts = np.stack([xs.ravel() / g, ys.ravel() / g, np.full(xs.size, 0.5)], 1).astype(np.float32)
    tris = []
    for y in range(g):
        for x in range(g):
            a = y * (g + 1) + x
            tris += [[a, a + 1, a + g + 1], [a + 1, a + g + 2, a + g + 1]]
    idx = np.array(tris, dtype=np.uint32)
    _check_lbvh(ctx, orc, verts, idx).destroy()


def _permuted(verts, idx, seed):
    """The same triangles with the vertex array shuffled and the indices remapped."""
    perm = np.random.default_rng(seed).permutation(len(verts))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    return verts[perm], inv[idx].astype(np.uint32)


def test_lbvh_non_identity_indices(ctx, orc):
    # Triangle i indexing vertices 3i..3i+2 takes a path without the index gather; shuffled
    # vertices, and a mesh that breaks the pattern in its last triangle only, take the general one.
    verts, idx = scenes.blob_scene(200_000, seed=5)
    pv, pi = _permuted(verts, idx, 6)
    _check_lbvh(ctx, orc, pv, pi).destroy()
    last = idx.copy()
    last[-1] = last[-1][::-1]
    _check_lbvh(ctx, orc, verts, last).destroy()
    accs = [api.TriangleAccel(ctx, verts, idx), api.TriangleAccel(ctx, pv, pi), api.TriangleAccel(ctx, verts, last)]
    api.build_batch(ctx, accs)
    assert api.last_error() is None, api.last_error()
    rk, rn, ra = orc.lbvh_build(verts, idx)
    for acc, (v, ix) in zip(accs, [(verts, idx), (pv, pi), (verts, last)]):
        keys, nodes, aabbs = acc.export()
        if ix is not idx:
            rk, rn, ra = orc.lbvh_build(v, ix)
        np.testing.assert_array_equal(keys, rk)
        np.testing.assert_array_equal(nodes, rn)
        np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))
        acc.destroy()


def test_lbvh_update_refit(ctx, orc):
    # 200,000 triangles: the refit also runs the region and top hierarchy levels
    verts, idx = scenes.blob_scene(200_000, seed=6)
    acc = _build(ctx, verts, idx, api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    keys0, nodes0, _ = acc.export()
    moved = (verts + scenes.rand01(7, 0, verts.size).reshape(verts.shape).astype(np.float32) * np.float32(2e-3)
             - np.float32(1e-3)).astype(np.float32)
    acc.vbuf.write(moved)
    acc.update()
    keys1, nodes1, aabbs1 = acc.export()
    np.testing.assert_array_equal(nodes1, nodes0)  # topology kept
    ref = orc.refit(moved, idx, nodes0)
    np.testing.assert_array_equal(aabbs1.view(np.uint32), ref.view(np.uint32))
    acc.destroy()


def test_update_requires_update_mode(ctx):
    verts, idx = scenes.uniform_scene(100, seed=4)
    acc = _build(ctx, verts, idx)
    acc.update()
    assert "AND_UPDATE" in api.last_error()
    ctx.lib.gprtClearLastError()
    acc.destroy()


def _trace(ctx, acc, rays):
    rb = ctx.device_buffer(rays)
    hb = ctx.device_buffer(nbytes=len(rays) * 16)
    acc.trace(rb, hb, len(rays))
    assert api.last_error() is None, api.last_error()
    hits = hb.read(api.HIT_DTYPE)
    rb.destroy()
    hb.destroy()
    return hits


def _assert_hits_equal(g, r):
    np.testing.assert_array_equal(g["primID"], r["primID"])
    np.testing.assert_array_equal(g["t"].view(np.uint32), r["t"].view(np.uint32))
    np.testing.assert_array_equal(g["u"].view(np.uint32), r["u"].view(np.uint32))
    np.testing.assert_array_equal(g["v"].view(np.uint32), r["v"].view(np.uint32))


def test_trace_uniform_primary_all_rays(ctx, orc):
    # 16,384 triangles, 256x256 pinhole rays: every ray against brute force
    verts, idx = scenes.uniform_scene(16384, seed=1)
    acc = _build(ctx, verts, idx)
    rays = scenes.pinhole_rays(256, 256)
    g = _trace(ctx, acc, rays)
    r = orc.trace_bruteforce(rays, verts, idx)
    _assert_hits_equal(g, r)
    assert (g["primID"] >= 0).sum() > 5000
    # bounce rays from the primary hits
    b = scenes.bounce_rays(rays, r, verts, idx, 8192, seed=3)
    _assert_hits_equal(_trace(ctx, acc, b), orc.trace_bruteforce(b, verts, idx))
    acc.destroy()


def test_trace_random_rays_inside_scene(ctx, orc):
    verts, idx = scenes.blob_scene(30_000, seed=8, offset=0.01)
    acc = _build(ctx, verts, idx)
    rays = scenes.random_rays(20_000, seed=9)
    _assert_hits_equal(_trace(ctx, acc, rays), orc.trace_bruteforce(rays, verts, idx))
    acc.destroy()


def test_trace_blob_scene_sampled(ctx, orc):
    # a 1M-triangle blob scene; a seeded 2,048-ray subsample (primary + bounce) against brute force
    verts, idx = scenes.blob_scene(1 << 20, seed=2)
    acc = _build(ctx, verts, idx)
    prim = scenes.pinhole_rays(512, 512)
    gp = _trace(ctx, acc, prim)
    sel = np.random.default_rng(0).choice(len(prim), 1024, replace=False)
    _assert_hits_equal(gp[sel], orc.trace_bruteforce(prim[sel], verts, idx))
    b = scenes.bounce_rays(prim, gp, verts, idx, 1024, seed=3)
    _assert_hits_equal(_trace(ctx, acc, b), orc.trace_bruteforce(b, verts, idx))
    acc.destroy()


def test_trace_t_error_vs_fp64(ctx, orc):
    verts, idx = scenes.uniform_scene(16384, seed=1)
    acc = _build(ctx, verts, idx)
    rays = scenes.pinhole_rays(128, 128)
    g = _trace(ctx, acc, rays)
    r8 = rays.view(np.float32).reshape(-1, 8)
    errs = []
    for i in np.nonzero(g["primID"] >= 0)[0][:3000]:
        v = verts[idx[g["primID"][i]]]
        t64 = orc.plane_t_fp64(r8[i], v[0], v[1], v[2])
        errs.append(abs(float(g["t"][i]) - t64) / abs(t64))
    errs = np.array(errs)
    print(f"hit t vs FP64 plane distance: max rel err {errs.max():.3e}, median {np.median(errs):.3e}")
    assert np.median(errs) < 1e-6 and errs.max() < 1e-3
    acc.destroy()


def test_trace_single_triangle_and_misses(ctx, orc):
    verts = np.array([[-1, -1, 1], [3, -1, 1], [-1, 3, 1]], dtype=np.float32)
    idx = np.array([[0, 1, 2]], dtype=np.uint32)
    acc = _build(ctx, verts, idx)
    rays = np.zeros(3, dtype=api.RAY_DTYPE)
    rays["oz"] = -1
    rays["dz"] = [1, -1, 1]
    rays["tmin"] = [0, 0, 2.0]  # third: hit exactly at tmin is excluded (open interval)
    rays["tmax"] = 1e30
    g = _trace(ctx, acc, rays)
    assert g["primID"].tolist() == [0, -1, -1]
    assert g["t"][0] == 2.0
    acc.destroy()


def test_refit_follows_moved_vertex_buffer(ctx, orc):
    """An update reads the geometry where it is now: after the vertex buffer is resized (the
    allocation moves) or the geometry is pointed at another buffer, the refit boxes equal the
    oracle's refit of the new vertices over the kept topology."""
    verts, idx = scenes.blob_scene(70_000, seed=8)
    acc = _build(ctx, verts, idx, api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    _, nodes0, _ = acc.export()
    moved = (verts + np.float32(1e-3)).astype(np.float32)
    ctx.lib.gprtBufferResize(ctx.h, acc.vbuf.h, 1, 2 * verts.nbytes, False, 0)
    acc.vbuf.write(moved)
    acc.update()
    assert api.last_error() is None, api.last_error()
    _, nodes1, aabbs1 = acc.export()
    np.testing.assert_array_equal(nodes1, nodes0)
    np.testing.assert_array_equal(aabbs1.view(np.uint32), orc.refit(moved, idx, nodes0).view(np.uint32))
    other = ctx.device_buffer((verts - np.float32(2e-3)).astype(np.float32))
    ctx.lib.gprtTrianglesSetVertices(acc.geom, other.h, len(verts), 12, 0)
    acc.update()
    _, _, aabbs2 = acc.export()
    ref = orc.refit((verts - np.float32(2e-3)).astype(np.float32), idx, nodes0)
    np.testing.assert_array_equal(aabbs2.view(np.uint32), ref.view(np.uint32))
    acc.destroy()
    other.destroy()


def test_refit_rechecks_index_layout(ctx, orc):
    """A refit runs no bounds pass, so it checks again whether every triangle i indexes vertices
    3i, 3i+1, 3i+2 (a triangle soup, whose vertices it then reads without gathering the indices):
    after the index buffer leaves that layout and returns to it, the refit boxes equal the
    oracle's refit of the triangles the current indices name."""
    verts, idx = scenes.blob_scene(70_000, seed=9)
    assert np.array_equal(idx.reshape(-1), np.arange(idx.size))  # the blob scene is a soup
    acc = _build(ctx, verts, idx, api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    _, nodes0, _ = acc.export()
    idx2 = np.ascontiguousarray(np.roll(idx, -1, axis=0))  # triangle i takes triangle i+1's vertices
    acc.ibuf.write(idx2)
    acc.update()
    assert api.last_error() is None, api.last_error()
    _, nodes1, aabbs1 = acc.export()
    np.testing.assert_array_equal(nodes1, nodes0)
    np.testing.assert_array_equal(aabbs1.view(np.uint32), orc.refit(verts, idx2, nodes0).view(np.uint32))
    moved = (verts + np.float32(1e-3)).astype(np.float32)
    acc.ibuf.write(idx)
    acc.vbuf.write(moved)
    acc.update()
    _, _, aabbs2 = acc.export()
    np.testing.assert_array_equal(aabbs2.view(np.uint32), orc.refit(moved, idx, nodes0).view(np.uint32))
    acc.destroy()


def test_plain_and_general_trace_forms_agree(ctx):
    """A plain trace (no proposal, no fetch counts) runs a specialised kernel; asking for fetch
    counts runs the general one. Both must give bit-identical hits, on a triangle accel, an AABB
    accel and an instance accel with transformed instances."""
    from test_oracle_boxes import random_boxes

    def both(acc, rays):
        n = len(rays)
        rb = ctx.device_buffer(rays)
        h1, h2 = ctx.device_buffer(nbytes=16 * n), ctx.device_buffer(nbytes=16 * n)
        i1, i2 = ctx.device_buffer(nbytes=4 * n), ctx.device_buffer(nbytes=4 * n)
        vb = ctx.device_buffer(nbytes=4 * n)
        api.trace_detail(ctx, acc, rb, h1, n, i1)              # plain form
        api.trace_detail(ctx, acc, rb, h2, n, i2, visits_buf=vb)  # general form
        assert api.last_error() is None, api.last_error()
        a, b = h1.read(api.HIT_DTYPE), h2.read(api.HIT_DTYPE)
        _assert_hits_equal(a, b)
        np.testing.assert_array_equal(i1.read(np.int32), i2.read(np.int32))
        assert (vb.read(np.uint32) > 0).any()
        hit = float((a["primID"] >= 0).mean())
        for b_ in (rb, h1, h2, i1, i2, vb):
            b_.destroy()
        return hit

    verts, idx = scenes.blob_scene(100_000, seed=12)
    tri = _build(ctx, verts, idx)
    prim = scenes.pinhole_rays(160, 120)
    ph = _trace(ctx, tri, prim)
    rays = np.concatenate([prim, scenes.bounce_rays(prim, ph, verts, idx, 20000, seed=13)])
    hits = [both(tri, rays)]

    boxes = random_boxes(20000, 9)
    box = api.AABBAccel(ctx, boxes)
    box.build()
    hits.append(both(box, scenes.random_rays(30000, seed=14)))

    meshes = [scenes.uniform_scene(n, seed=70 + k) for k, n in enumerate([3000, 5000, 800])]
    blas = [_build(ctx, v, ix) for v, ix in meshes]
    moves = [np.eye(3, 4, dtype=np.float32) for _ in blas]
    for k, m in enumerate(moves):
        m[:, 3] = np.float32(0.15 * k)  # translated instances: traced in object space
    tlas = api.InstanceAccel(ctx, api.instance_records(blas, transforms=moves))
    tlas.build()
    hits.append(both(tlas, scenes.random_rays(30000, seed=15)))
    print(f"hit rates (triangles, boxes, instances): {hits}", flush=True)
    assert min(hits) > 0.02  # the comparison covers hits, not only misses
    for a in [tlas, box, tri] + blas:
        a.destroy()
