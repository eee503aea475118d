"""Scenes of several accels: gprtAccelTraceClosestHitMulti against the oracle's per-mesh brute
force combined under the order (t, instance, primID), and gprtSynthBounceRays against the numpy
generator scenes.bounce_rays_multi."""
import numpy as np
import pytest

from gprt_amd import api, scenes

pytestmark = pytest.mark.gpu


def _meshes(sizes, seed0):
    out = []
    for i, n in enumerate(sizes):
        v, idx = scenes.blob_scene(n, seed=seed0 + i) if n >= 64 else scenes.uniform_scene(n, seed=seed0 + i)
        out.append((v, idx))
    return out


def _oracle_multi(orc, rays, meshes):
    """Per-mesh brute-force closest hit, then the lexicographic minimum of (t, instance, primID)."""
    n = len(rays)
    best_t = rays["tmax"].astype(np.float32).copy()
    best_p = np.full(n, -1, np.int32)
    best_i = np.full(n, -1, np.int32)
    for k, (v, idx) in enumerate(meshes):
        h = orc.trace_bruteforce(rays, v, idx)
        hit = h["primID"] >= 0
        better = hit & ((best_p < 0) | (h["t"] < best_t))
        best_t[better] = h["t"][better]
        best_p[better] = h["primID"][better]
        best_i[better] = k
    return best_t, best_p, best_i


def _rays(n_side, seed):
    prim = scenes.pinhole_rays(n_side, n_side)
    rnd = scenes.random_rays(n_side * n_side // 2, seed)
    return np.concatenate([prim, rnd])


@pytest.mark.parametrize("sizes", [[5000, 1, 3000, 64, 20000, 7], [40] * 70])
def test_multi_trace_matches_bruteforce(ctx, orc, sizes):
    meshes = _meshes(sizes, 300 + len(sizes))
    accs = [api.TriangleAccel(ctx, v, idx) for v, idx in meshes]
    api.build_batch(ctx, accs)
    rays = _rays(96, 11)
    n = len(rays)
    rb = ctx.device_buffer(rays)
    hb = ctx.device_buffer(nbytes=16 * n)
    ib = ctx.device_buffer(nbytes=4 * n)
    print(f"multi trace: {len(sizes)} instances, {n} rays", flush=True)
    api.trace_multi(ctx, accs, rb, hb, n, ib)
    assert api.last_error() is None, api.last_error()
    hits = hb.read(api.HIT_DTYPE, n)
    inst = ib.read(np.int32, n)
    t, p, i = _oracle_multi(orc, rays, meshes)
    np.testing.assert_array_equal(hits["primID"], p)
    np.testing.assert_array_equal(inst, i)
    hit = p >= 0
    assert hit.mean() > 0.05
    np.testing.assert_array_equal(hits["t"][hit].view(np.uint32), t[hit].view(np.uint32))
    # A single-instance multi trace equals the plain trace.
    hb1 = ctx.device_buffer(nbytes=16 * n)
    accs[0].trace(rb, hb1, n)
    api.trace_multi(ctx, accs[:1], rb, hb, n, None)
    np.testing.assert_array_equal(hb.read(api.HIT_DTYPE, n).view(np.uint32), hb1.read(api.HIT_DTYPE, n).view(np.uint32))
    for b in (rb, hb, ib, hb1):
        b.destroy()
    for a in accs:
        a.destroy()


def test_bounce_rays_match_numpy(ctx, orc):
    meshes = _meshes([20000, 5000, 64], 500)
    accs = [api.TriangleAccel(ctx, v, idx) for v, idx in meshes]
    api.build_batch(ctx, accs)
    rays = scenes.pinhole_rays(128, 128)
    n = len(rays)
    rb = ctx.device_buffer(rays)
    hb = ctx.device_buffer(nbytes=16 * n)
    ib = ctx.device_buffer(nbytes=4 * n)
    api.trace_multi(ctx, accs, rb, hb, n, ib)
    hits = hb.read(api.HIT_DTYPE, n)
    inst = ib.read(np.int32, n)
    out_n = 3 * n + 17
    ob = ctx.device_buffer(nbytes=32 * out_n)
    api.bounce_rays(ctx, accs, rb, hb, ib, n, ob, out_n, seed=9, tmin=1e-4)
    assert api.last_error() is None, api.last_error()
    got = ob.read(api.RAY_DTYPE, out_n)
    ref = scenes.bounce_rays_multi(rays, hits, inst, meshes, out_n, seed=9, tmin=1e-4)
    assert len(ref) == out_n
    g = got.view(np.float32).reshape(out_n, 8)
    r = ref.view(np.float32).reshape(out_n, 8)
    # FP64 cos/sin/sqrt on the device and in libm may differ in the last FP64 bit, which can move a
    # rounded FP32 value by one ulp; origins and the interval are exact.
    np.testing.assert_array_equal(g[:, [0, 1, 2, 3, 7]], r[:, [0, 1, 2, 3, 7]])
    ulps = np.abs(g.view(np.int32)[:, 4:7].astype(np.int64) - r.view(np.int32)[:, 4:7].astype(np.int64))
    assert ulps.max() <= 1
    assert (ulps == 0).mean() > 0.999
    # No hit at all: every bounce ray is an empty interval.
    far = rays.copy()
    far["oz"] = -100.0
    far["dz"] = -1.0
    rb.write(far)
    api.trace_multi(ctx, accs, rb, hb, n, ib)
    api.bounce_rays(ctx, accs, rb, hb, ib, n, ob, 64, seed=9)
    e = ob.read(api.RAY_DTYPE, 64)
    assert (e["tmax"] < e["tmin"]).all()
    for b in (rb, hb, ib, ob):
        b.destroy()
    for a in accs:
        a.destroy()
