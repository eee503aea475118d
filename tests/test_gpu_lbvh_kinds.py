"""The development LBVH interface (gprtPointsLBVHCreate / gprtEdgesLBVHCreate /
gprtTriangleLBVHCreate + gprtLBVHBuild, gprt/development/gprt_lbvh.h:28-211): its four result
buffers — sorted Morton codes, sorted primIDs, int4 nodes and float3 box pairs — bit-exact
against the oracle's point, edge and triangle LBVHs, across the builder's tile, sub-tile and
region edges, for the reference's t04 range (points within +-1e38), after the inputs change, and
with malformed input rejected before any launch."""
import numpy as np
import pytest

from gprt_amd import api, scenes

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 1000, 8193, 70_000, 300_001]


def _check(lb, ref):
    codes, ids, nodes, aabbs = lb.read()
    rk, rn, ra = ref
    np.testing.assert_array_equal(codes, (rk >> np.uint64(32)).astype(np.uint32))
    np.testing.assert_array_equal(ids, (rk & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))


@pytest.mark.parametrize("n", SIZES)
def test_points_lbvh(ctx, orc, n):
    c = scenes.blob_scene(n, seed=60 + n)[0][::3] if n >= 64 else np.random.default_rng(n).random((n, 3), np.float32)
    c = np.ascontiguousarray(c[:n], dtype=np.float32)
    lb = api.LBVH(ctx, c, kind="points")
    print(f"points LBVH: {n}", flush=True)
    lb.build()
    assert api.last_error() is None, api.last_error()
    _check(lb, orc.lbvh_build_points(c))
    lb.destroy()


@pytest.mark.parametrize("n", SIZES)
def test_edges_lbvh(ctx, orc, n):
    rng = np.random.default_rng(200 + n)
    v = rng.random((n // 2 + 3, 3), dtype=np.float32)
    e = rng.integers(0, len(v), (n, 2)).astype(np.uint32)
    lb = api.LBVH(ctx, v, e, kind="edges")
    print(f"edges LBVH: {n}", flush=True)
    lb.build()
    assert api.last_error() is None, api.last_error()
    _check(lb, orc.lbvh_build_edges(v, e))
    lb.destroy()


@pytest.mark.parametrize("n", SIZES)
def test_triangles_lbvh(ctx, orc, n):
    v, idx = scenes.blob_scene(n, seed=300 + n) if n >= 64 else scenes.uniform_scene(n, seed=300 + n)
    lb = api.LBVH(ctx, v, idx, kind="triangles")
    print(f"triangle LBVH: {n}", flush=True)
    lb.build()
    assert api.last_error() is None, api.last_error()
    _check(lb, orc.lbvh_build(v, idx))
    lb.destroy()


def test_points_t04_range_and_rebuild(ctx, orc):
    # tests/t04-swBVH/hostCode.cpp:39-68: 42 points within +-1e38.
    rng = np.random.default_rng(4)
    p = (rng.random((42, 3)) * 2e38 - 1e38).astype(np.float32)
    lb = api.LBVH(ctx, p, kind="points")
    lb.build()
    _check(lb, orc.lbvh_build_points(p))
    # New positions in the same buffer: a second build follows them.
    q = (p * np.float32(0.5) + np.float32(3.0)).astype(np.float32)
    lb.vbuf.write(q)
    lb.build()
    _check(lb, orc.lbvh_build_points(q))
    lb.destroy()


def test_lbvh_rejects_malformed_input(ctx):
    lib = ctx.lib
    v = ctx.device_buffer(np.zeros((10, 3), np.float32))
    e = ctx.device_buffer(np.zeros((4, 2), np.uint32))
    bad = lib.gprtPointsLBVHCreate(ctx.h, v.h, 11)  # more points than the buffer holds
    assert not bad.impl and "exceed" in (api.last_error() or "")
    lib.gprtClearLastError()
    bad = lib.gprtEdgesLBVHCreate(ctx.h, v.h, e.h, 5)  # more edges than the index buffer holds
    assert not bad.impl and "exceed" in (api.last_error() or "")
    lib.gprtClearLastError()
    bad = lib.gprtTriangleLBVHCreate(ctx.h, v.h, None, 3)
    assert not bad.impl and "null" in (api.last_error() or "")
    lib.gprtClearLastError()
    ok = lib.gprtEdgesLBVHCreate(ctx.h, v.h, e.h, 4)
    assert ok.impl
    ctx.lib.gprtBufferResize(ctx.h, e.h, 1, 8, False, 0)  # shrunk below 4 edges after create
    lib.gprtLBVHBuild(ctx.h, api.C.byref(ok))
    assert "exceed" in (api.last_error() or "")
    lib.gprtClearLastError()
    lib.gprtLBVHDestroy(api.C.byref(ok))
    assert not ok.impl
    v.destroy()
    e.destroy()
