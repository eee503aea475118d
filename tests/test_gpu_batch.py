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


@pytest.fixture(scope="module")
def batch_meshes(orc):
    """The meshes with their oracle LBVHs (the same for every configuration)."""
    return [(v, idx, orc.lbvh_build(v, idx)) for v, idx in _meshes(1000)]


# Every configuration the autotuner may pick: digit width x tile x sort-pass occupancy cap.
@pytest.mark.parametrize("bits", [8, 6, 4])
@pytest.mark.parametrize("tile", [2048, 4096, 8192, 16384])
@pytest.mark.parametrize("waves", [0, 8, 16])
def test_batch_equals_oracle_per_mesh(ctx, orc, batch_meshes, bits, tile, waves):
    ctx.lib.gprtSetSortConfig(ctx.h, bits, tile, waves)
    try:
        accs = [api.TriangleAccel(ctx, v, idx) for v, idx, _ in batch_meshes]
        print(f"batch: {len(accs)} meshes, digits {bits}, tile {tile}, waves/CU {waves}", flush=True)
        api.build_batch(ctx, accs)
        assert api.last_error() is None, api.last_error()
        for acc, (v, idx, (rk, rn, ra)) in zip(accs, batch_meshes):
            k, nodes, aabbs = acc.export()
            np.testing.assert_array_equal(k, rk)
            np.testing.assert_array_equal(nodes, rn)
            np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))
            lo = v[idx.ravel()].min(0)
            hi = v[idx.ravel()].max(0)
            np.testing.assert_array_equal(acc.bounds(), np.concatenate([lo, hi]))
        for a in accs:
            a.destroy()
    finally:
        ctx.lib.gprtSetSortConfig(ctx.h, 8, 0, 0)


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


def _shrink(ctx, buf, nbytes):
    ctx.lib.gprtBufferResize(ctx.h, buf.h, 1, nbytes, True, 0)


@pytest.mark.parametrize("which", ["vertices", "indices"])
def test_build_rejects_buffers_shrunk_after_set(ctx, which):
    """A vertex or index buffer resized below what gprtTrianglesSet* declared fails the build, the
    batch build and the update before any kernel reads it."""
    v, idx = scenes.uniform_scene(1000, seed=3)
    a = api.TriangleAccel(ctx, v, idx)
    a.build(api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    assert api.last_error() is None
    buf = a.vbuf if which == "vertices" else a.ibuf
    _shrink(ctx, buf, buf.nbytes // 2)
    assert api.last_error() is None
    for call in (lambda: a.update(), lambda: a.build(), lambda: api.build_batch(ctx, [a])):
        call()
        err = api.last_error()
        assert err is not None and f"exceed the {'vertex' if which == 'vertices' else 'index'} buffer" in err, err
        ctx.lib.gprtClearLastError()
    a.destroy()


def test_box_build_rejects_buffer_shrunk_after_set(ctx):
    rng = np.random.default_rng(3)
    c = rng.random((500, 3), dtype=np.float32)
    boxes = np.concatenate([c - 0.01, c + 0.01], 1).astype(np.float32)
    a = api.AABBAccel(ctx, boxes)
    _shrink(ctx, a.vbuf, 24 * 499)
    a.build()
    assert "exceed the box buffer" in (api.last_error() or "")
    ctx.lib.gprtClearLastError()
    a.destroy()


def test_autotune_small_batch(ctx, orc):
    """gprtAccelBuildBatchAutotune: every configuration of the sweep is reported, the returned index
    names the fastest one, gprtGetSortConfig reads it back, and the accels it leaves built equal the
    oracle's LBVH of each mesh. A batch naming an accel twice is rejected."""
    import json
    import ctypes as C
    meshes = _meshes(4000)[:9]
    accs = [api.TriangleAccel(ctx, v, idx) for v, idx in meshes]
    arr = (C.c_void_p * len(accs))(*[a.h for a in accs])
    rep = C.create_string_buffer(1 << 14)
    try:
        best = ctx.lib.gprtAccelBuildBatchAutotune(ctx.h, arr, len(accs), 1, rep, len(rep))
        assert api.last_error() is None, api.last_error()
        r = json.loads(rep.value.decode())
        cfgs = r["configs"]
        assert len(cfgs) == 30
        assert {(c["digitBits"], c["tileKeys"], c["wavesPerCU"]) for c in cfgs} == {
            (b, t, w) for b in (4, 6, 8) for t in (2048, 4096, 8192) for w in (0, 8, 16)} | {
            (b, 16384, 0) for b in (4, 6, 8)}
        ms = [c["ms"] for c in cfgs]
        assert best == int(np.argmin(ms))
        for k in ("digitBits", "tileKeys", "wavesPerCU"):
            assert r["best"][k] == cfgs[best][k]
        got = [C.c_uint32(0) for _ in range(3)]
        ctx.lib.gprtGetSortConfig(ctx.h, *[C.byref(g) for g in got])
        assert [g.value for g in got] == [r["best"]["digitBits"], r["best"]["tileKeys"], r["best"]["wavesPerCU"]]
        for acc, (v, idx) in zip(accs, meshes):
            _check(acc, orc, v, idx)
        dup = (C.c_void_p * 2)(accs[0].h, accs[0].h)
        ctx.lib.gprtAccelBuildBatchAutotune(ctx.h, dup, 2, 1, rep, len(rep))
        assert "twice" in (api.last_error() or "")
        ctx.lib.gprtClearLastError()
    finally:
        ctx.lib.gprtSetSortConfig(ctx.h, 8, 0, 0)
        for a in accs:
            a.destroy()
