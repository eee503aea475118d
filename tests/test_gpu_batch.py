"""Batched builds (gprtAccelBuildBatch): every mesh of one graph-captured batch launch must equal
the oracle's LBVH of that mesh alone (keys, topology, boxes bit-exact), across graph replays with
vertex updates in between, and traces against batch members must match brute force."""
import numpy as np
import pytest

from gprt_amd import api, scenes

pytestmark = pytest.mark.gpu

# 65536 fills one hierarchy region exactly; 150001 spans three (edges through the region and top levels)
SIZES = [1, 2, 3, 64, 100, 777, 2048, 4095, 4096, 4097, 8191, 8193, 20000, 65536, 3, 1000, 150001]


def _meshes(seed0):
    out = []
    for i, n in enumerate(SIZES):
        v, idx = scenes.blob_scene(n, seed=seed0 + i) if n >= 64 else scenes.uniform_scene(n, seed=seed0 + i)
        out.append((v, idx))
    return out


def _check(acc, orc, v, idx):
    k, nodes, aabbs = acc.export()
    rk, rn, ra = orc.lbvh_build(v, idx)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))


@pytest.mark.parametrize("bits,tile", [(8, 4096), (6, 2048), (4, 8192), (8, 8192)])
def test_batch_equals_oracle_per_mesh(ctx, orc, bits, tile):
    ctx.lib.gprtSetSortConfig(ctx.h, bits, tile, 0)
    try:
        meshes = _meshes(1000)
        accs = [api.TriangleAccel(ctx, v, idx) for v, idx in meshes]
        api.build_batch(ctx, accs)
        assert api.last_error() is None, api.last_error()
        for acc, (v, idx) in zip(accs, meshes):
            _check(acc, orc, v, idx)
        for a in accs:
            a.destroy()
    finally:
        ctx.lib.gprtSetSortConfig(ctx.h, 8, 8192, 0)


def test_batch_graph_replays_with_moving_vertices(ctx, orc):
    meshes = _meshes(2000)
    accs = [api.TriangleAccel(ctx, v, idx) for v, idx in meshes]
    for frame in range(3):
        if frame:
            moved = []
            for j, (acc, (v, idx)) in enumerate(zip(accs, meshes)):
                d = scenes.rand01(7000 + frame, j, v.size).reshape(v.shape) * 2e-3 - 1e-3
                v2 = (v + d).astype(np.float32)
                acc.vbuf.write(v2)
                moved.append((v2, idx))
            meshes = moved
        api.build_batch(ctx, accs)  # frame 0 captures the graph, later frames replay it
        assert api.last_error() is None, api.last_error()
        for acc, (v, idx) in zip(accs, meshes):
            _check(acc, orc, v, idx)
    # a trace against one batch member
    v, idx = meshes[-3]
    rays = scenes.pinhole_rays(64, 64)
    rb, hb = ctx.device_buffer(rays), ctx.device_buffer(nbytes=len(rays) * 16)
    accs[-3].trace(rb, hb, len(rays))
    g = hb.read(api.HIT_DTYPE)
    r = orc.trace_bruteforce(rays, v, idx)
    np.testing.assert_array_equal(g["primID"], r["primID"])
    np.testing.assert_array_equal(g["t"].view(np.uint32), r["t"].view(np.uint32))
    for a in accs:
        a.destroy()


def test_rebuild_alone_after_batch(ctx, orc):
    meshes = _meshes(3000)[:5]
    accs = [api.TriangleAccel(ctx, v, idx) for v, idx in meshes]
    api.build_batch(ctx, accs)
    accs[2].build()  # leaves the batch store
    for acc, (v, idx) in zip(accs, meshes):
        _check(acc, orc, v, idx)
    for a in accs:
        a.destroy()


def test_batch_rejects_duplicates(ctx):
    v, idx = scenes.uniform_scene(10, seed=1)
    a = api.TriangleAccel(ctx, v, idx)
    api.build_batch(ctx, [a, a])
    assert "twice" in api.last_error()
    ctx.lib.gprtClearLastError()
    a.destroy()
