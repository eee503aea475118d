"""Instance accels (gprtInstanceAccelCreate over gprt::Instance records): closest hits against the
oracle's per-instance brute force in each instance's object space, combined under the order
(t, instance index, primID); hidden, invalid and destroyed instances; the learned proposal over
top-level leaves; and bounce rays over instances against the numpy generator."""
import math

import numpy as np
import pytest

from gprt_amd import api, scenes

pytestmark = pytest.mark.gpu


def _meshes(sizes, seed0, blob=False):
    """Uniform meshes (triangles of about 0.02, so that small meshes still catch many rays), or
    blob meshes as in the dynamic scene."""
    out = []
    for i, n in enumerate(sizes):
        if blob and n >= 64:
            v, idx = scenes.blob_scene(n, seed=seed0 + i)
        else:
            v, idx = scenes.uniform_scene(n, seed=seed0 + i)
        out.append((v, idx))
    return out


def _transform(k):
    """A rotation about a tilted axis, a non-uniform scale and a translation (FP32 3x4)."""
    ang = 0.3 + 0.7 * k
    ax = np.array([1.0, 2.0 + k, 0.5])
    ax /= np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)
    S = np.diag([1.0 + 0.1 * k, 0.8, 1.2 - 0.05 * k])
    out = np.zeros((3, 4), dtype=np.float32)
    out[:, :3] = R @ S
    out[:, 3] = [0.1 * k - 0.2, 0.05 * k, -0.1]
    return out


def _oracle(orc, rays, meshes, transforms=None, enabled=None):
    """Per-instance brute force (object-space rays), then the minimum of (t, instance, primID)."""
    n = len(rays)
    best_t = rays["tmax"].astype(np.float32).copy()
    best_p = np.full(n, -1, np.int32)
    best_i = np.full(n, -1, np.int32)
    for k, (v, idx) in enumerate(meshes):
        if enabled is not None and not enabled[k]:
            continue
        r = rays if transforms is None else scenes.object_rays(rays, transforms[k])
        h = orc.trace_bruteforce(r, v, idx)
        hit = h["primID"] >= 0
        better = hit & ((best_p < 0) | (h["t"] < best_t))
        best_t[better] = h["t"][better]
        best_p[better] = h["primID"][better]
        best_i[better] = k
    return best_t, best_p, best_i


def _rays(n_side, seed):
    return np.concatenate([scenes.pinhole_rays(n_side, n_side), scenes.random_rays(n_side * n_side // 2, seed)])


class Scene:
    def __init__(self, ctx, sizes, seed0, transforms=None, masks=None, blob=False):
        self.ctx = ctx
        self.meshes = _meshes(sizes, seed0, blob)
        self.blas = [api.TriangleAccel(ctx, v, idx) for v, idx in self.meshes]
        api.build_batch(ctx, self.blas)
        self.records = api.instance_records(self.blas, transforms, masks)
        self.tlas = api.InstanceAccel(ctx, self.records)
        self.tlas.build()
        assert api.last_error() is None, api.last_error()

    def trace(self, rays, proposal=None):
        ctx, n = self.ctx, len(rays)
        rb = ctx.device_buffer(rays)
        hb, ib, lb = (ctx.device_buffer(nbytes=16 * n), ctx.device_buffer(nbytes=4 * n),
                      ctx.device_buffer(nbytes=4 * n))
        print(f"instance trace: {len(self.blas)} instances, {n} rays, proposal={proposal is not None}", flush=True)
        api.trace_detail(ctx, self.tlas, rb, hb, n, ib, lb, proposal=proposal)
        assert api.last_error() is None, api.last_error()
        out = hb.read(api.HIT_DTYPE, n), ib.read(np.int32, n), lb.read(np.int32, n)
        for b in (rb, hb, ib, lb):
            b.destroy()
        return out

    def destroy(self):
        self.tlas.destroy()
        for a in self.blas:
            a.destroy()


def _check(hits, inst, ref):
    t, p, i = ref
    np.testing.assert_array_equal(hits["primID"], p)
    np.testing.assert_array_equal(inst, i)
    hit = p >= 0
    np.testing.assert_array_equal(hits["t"][hit].view(np.uint32), t[hit].view(np.uint32))
    return hit.mean()


@pytest.mark.parametrize("sizes", [[5000, 1, 3000, 64, 20000, 7], [40] * 70, [3] * 300])
def test_identity_instances_match_bruteforce(ctx, orc, sizes):
    sc = Scene(ctx, sizes, 300 + len(sizes))
    rays = _rays(96, 11)
    hits, inst, leaf = sc.trace(rays)
    # Hit rates are about 19%, 3% and 1% for the three scenes.
    assert _check(hits, inst, _oracle(orc, rays, sc.meshes)) > 0.003
    # The top level holds one leaf per triangle, instance after instance: a hit's leaf lies in its
    # instance's range.
    L = int(ctx.lib.gprtAccelGetPrimCount(sc.tlas.h))
    assert L == sum(sizes)
    first = np.concatenate([[0], np.cumsum(sizes)])
    hit = hits["primID"] >= 0
    np.testing.assert_array_equal(np.searchsorted(first, leaf[hit], side="right") - 1, inst[hit])
    assert (leaf[~hit] == -1).all()
    sc.destroy()


def test_single_instance_equals_triangle_trace(ctx):
    sc = Scene(ctx, [30000], 77)
    rays = _rays(64, 5)
    n = len(rays)
    hits, inst, _ = sc.trace(rays)
    rb, hb = ctx.device_buffer(rays), ctx.device_buffer(nbytes=16 * n)
    sc.blas[0].trace(rb, hb, n)
    np.testing.assert_array_equal(hits.view(np.uint32), hb.read(api.HIT_DTYPE, n).view(np.uint32))
    assert ((inst == 0) == (hits["primID"] >= 0)).all()
    rb.destroy()
    hb.destroy()
    sc.destroy()


def test_transformed_instances_match_bruteforce(ctx, orc):
    sizes = [4000, 2500, 64, 9000]
    tf = [_transform(k) for k in range(len(sizes))]
    sc = Scene(ctx, sizes, 900, transforms=tf)
    rays = _rays(96, 12)
    hits, inst, _ = sc.trace(rays)
    assert _check(hits, inst, _oracle(orc, rays, sc.meshes, tf)) > 0.05
    sc.destroy()


def test_hidden_invalid_and_destroyed_instances(ctx, orc):
    sizes = [3000, 3000, 3000, 3000]
    masks = [0xFF, 0, 0xFF, 0xFF]
    sc = Scene(ctx, sizes, 1300, masks=masks)
    rays = _rays(64, 13)
    hits, inst, _ = sc.trace(rays)
    _check(hits, inst, _oracle(orc, rays, sc.meshes, enabled=[True, False, True, True]))
    # An address that names no descriptor slot disables its instance (checked on the device
    # before any read through it).
    rec = sc.records.copy()
    rec["accelAddress"][2] = 0x1000
    sc.tlas.set_records(rec)
    sc.tlas.build()
    hits, inst, _ = sc.trace(rays)
    _check(hits, inst, _oracle(orc, rays, sc.meshes, enabled=[True, False, False, True]))
    # Destroying a triangle accel disables it in an instance accel built earlier.
    sc.tlas.set_records(sc.records)
    sc.tlas.build()
    sc.blas[3].destroy()
    hits, inst, _ = sc.trace(rays)
    _check(hits, inst, _oracle(orc, rays, sc.meshes, enabled=[True, False, True, False]))
    sc.destroy()


def test_rebuild_follows_moving_vertices(ctx, orc):
    sizes = [6000, 5000, 4000]
    sc = Scene(ctx, sizes, 1700)
    rays = _rays(64, 14)
    for frame in range(1, 3):
        moved = []
        for j, (acc, (v, idx)) in enumerate(zip(sc.blas, sc.meshes)):
            d = scenes.rand01(7000 + frame, j, v.size).reshape(v.shape) * 2e-3 - 1e-3
            v2 = (v + d).astype(np.float32)
            acc.vbuf.write(v2)
            moved.append((v2, idx))
        sc.meshes = moved
        api.build_batch(ctx, sc.blas)  # graph replay
        sc.tlas.build()
        hits, inst, _ = sc.trace(rays)
        _check(hits, inst, _oracle(orc, rays, sc.meshes))
    sc.destroy()


def test_proposal_over_instances_is_exact(ctx):
    sizes = [20000] * 8
    sc = Scene(ctx, sizes, 2100, blob=True)
    rays = _rays(96, 15)
    n = len(rays)
    hits, inst, leaf = sc.trace(rays)
    prop = api.Proposal(ctx, max_batch=1024, seed=3)
    classes = prop.bind(sc.tlas)
    # Classes are fixed groups of each instance's triangles: 64 per instance while
    # instances * 64 <= 4096.
    assert classes == 64 * len(sizes)
    rb, lbuf = ctx.device_buffer(rays), ctx.device_buffer(leaf)
    loss, acc = prop.train_labels(rb, lbuf, n, epochs=4, lr=2e-3)
    assert api.last_error() is None, api.last_error()
    assert np.isfinite(loss)
    h2, i2, l2 = sc.trace(rays, proposal=prop)
    np.testing.assert_array_equal(h2.view(np.uint32), hits.view(np.uint32))
    np.testing.assert_array_equal(i2, inst)
    np.testing.assert_array_equal(l2, leaf)
    rb.destroy()
    lbuf.destroy()
    prop.destroy()
    sc.destroy()


@pytest.mark.parametrize("transformed", [False, True])
def test_bounce_rays_over_instances_match_numpy(ctx, transformed):
    sizes = [20000, 5000, 64]
    tf = [_transform(k) for k in range(len(sizes))] if transformed else None
    sc = Scene(ctx, sizes, 500, transforms=tf)
    rays = scenes.pinhole_rays(128, 128)
    n = len(rays)
    rb = ctx.device_buffer(rays)
    hb, ib = ctx.device_buffer(nbytes=16 * n), ctx.device_buffer(nbytes=4 * n)
    api.trace_detail(ctx, sc.tlas, rb, hb, n, ib)
    hits, inst = hb.read(api.HIT_DTYPE, n), ib.read(np.int32, n)
    out_n = 3 * n + 17
    ob = ctx.device_buffer(nbytes=32 * out_n)
    api.bounce_rays(ctx, sc.tlas, rb, hb, ib, n, ob, out_n, seed=9, tmin=1e-4)
    assert api.last_error() is None, api.last_error()
    got = ob.read(api.RAY_DTYPE, out_n)
    ref = scenes.bounce_rays_multi(rays, hits, inst, sc.meshes, out_n, seed=9, tmin=1e-4, transforms=tf)
    g = got.view(np.float32).reshape(out_n, 8)
    r = ref.view(np.float32).reshape(out_n, 8)
    # FP64 cos / sin / sqrt on the device and in libm may differ in the last FP64 bit, which can
    # move a rounded FP32 direction by one ulp; origins and the interval are exact.
    np.testing.assert_array_equal(g[:, [0, 1, 2, 3, 7]], r[:, [0, 1, 2, 3, 7]])
    ulps = np.abs(g.view(np.int32)[:, 4:7].astype(np.int64) - r.view(np.int32)[:, 4:7].astype(np.int64))
    assert ulps.max() <= 1
    assert (ulps == 0).mean() > 0.999
    # No hit at all: every bounce ray is an empty interval.
    far = rays.copy()
    far["oz"] = -100.0
    far["dz"] = -1.0
    rb.write(far)
    api.trace_detail(ctx, sc.tlas, rb, hb, n, ib)
    api.bounce_rays(ctx, sc.tlas, rb, hb, ib, n, ob, 64, seed=9)
    e = ob.read(api.RAY_DTYPE, 64)
    assert (e["tmax"] < e["tmin"]).all()
    for b in (rb, hb, ib, ob):
        b.destroy()
    sc.destroy()


def test_bounce_rays_single_accel_match_numpy(ctx):
    v, idx = scenes.blob_scene(30000, seed=44)
    acc = api.TriangleAccel(ctx, v, idx)
    acc.build()
    rays = scenes.pinhole_rays(128, 128)
    n = len(rays)
    rb, hb = ctx.device_buffer(rays), ctx.device_buffer(nbytes=16 * n)
    acc.trace(rb, hb, n)
    hits = hb.read(api.HIT_DTYPE, n)
    ob = ctx.device_buffer(nbytes=32 * n)
    api.bounce_rays(ctx, acc, rb, hb, None, n, ob, n, seed=3)
    got = ob.read(api.RAY_DTYPE, n).view(np.float32).reshape(n, 8)
    ref = scenes.bounce_rays(rays, hits, v, idx, n, seed=3).view(np.float32).reshape(n, 8)
    np.testing.assert_array_equal(got[:, [0, 1, 2, 3, 7]], ref[:, [0, 1, 2, 3, 7]])
    ulps = np.abs(got.view(np.int32)[:, 4:7].astype(np.int64) - ref.view(np.int32)[:, 4:7].astype(np.int64))
    assert ulps.max() <= 1
    for b in (rb, hb, ob):
        b.destroy()
    acc.destroy()


@pytest.mark.parametrize("change", ["refit", "rebuild"])
def test_proposal_skipped_after_triangle_accel_changes(ctx, orc, change):
    """An instance accel traces copies of its instances' triangles taken at its own build. When a
    triangle accel is refit or rebuilt (same size, same storage) afterwards without rebuilding the
    instance accel, the proposal step must not walk the changed triangles: traces with and
    without a proposal stay identical, and both equal the brute force over the geometry the
    instance accel was built from."""
    sizes = [20000, 20000, 20000]
    meshes = _meshes(sizes, 2600, blob=True)
    blas = [api.TriangleAccel(ctx, v, idx) for v, idx in meshes]
    for a in blas:
        a.build(api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    tlas = api.InstanceAccel(ctx, api.instance_records(blas))
    tlas.build()
    rays = _rays(96, 16)
    n = len(rays)
    rb = ctx.device_buffer(rays)
    hb, ib, lb = ctx.device_buffer(nbytes=16 * n), ctx.device_buffer(nbytes=4 * n), ctx.device_buffer(nbytes=4 * n)
    api.trace_detail(ctx, tlas, rb, hb, n, ib, lb)
    hits0, inst0, leaf0 = hb.read(api.HIT_DTYPE, n), ib.read(np.int32, n), lb.read(np.int32, n)
    assert _check(hits0, inst0, _oracle(orc, rays, meshes)) > 0.01
    prop = api.Proposal(ctx, max_batch=1024, seed=5)
    prop.bind(tlas)
    prop.train_labels(rb, lb, n, epochs=4, lr=2e-3)
    # Move every triangle of instance 0 towards the camera: the moved geometry would be hit first.
    v0, idx0 = meshes[0]
    moved = v0.copy()
    moved[:, 2] -= np.float32(0.3)
    blas[0].vbuf.write(moved)
    if change == "refit":
        blas[0].update()
    else:
        blas[0].build(api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    assert api.last_error() is None, api.last_error()
    h1, i1, l1 = (ctx.device_buffer(nbytes=16 * n), ctx.device_buffer(nbytes=4 * n),
                  ctx.device_buffer(nbytes=4 * n))
    print(f"proposal after a {change} of a triangle accel: {n} rays", flush=True)
    api.trace_detail(ctx, tlas, rb, hb, n, ib, lb)
    api.trace_detail(ctx, tlas, rb, h1, n, i1, l1, proposal=prop)
    ctx.lib.gprtAccelTraceClosestHitEx(ctx.h, tlas.h, prop.h, rb.h, h1.h, n, None)
    assert api.last_error() is None, api.last_error()
    plain = hb.read(api.HIT_DTYPE, n)
    np.testing.assert_array_equal(plain.view(np.uint32), hits0.view(np.uint32))
    np.testing.assert_array_equal(ib.read(np.int32, n), inst0)
    np.testing.assert_array_equal(h1.read(api.HIT_DTYPE, n).view(np.uint32), plain.view(np.uint32))
    np.testing.assert_array_equal(i1.read(np.int32, n), inst0)
    np.testing.assert_array_equal(l1.read(np.int32, n), leaf0)
    # After the instance accel is rebuilt it traces the moved triangles, with or without proposal.
    tlas.build()
    api.trace_detail(ctx, tlas, rb, hb, n, ib, lb)
    api.trace_detail(ctx, tlas, rb, h1, n, i1, l1, proposal=prop)
    now = [(moved, idx0)] + meshes[1:]
    _check(hb.read(api.HIT_DTYPE, n), ib.read(np.int32, n), _oracle(orc, rays, now))
    np.testing.assert_array_equal(h1.read(api.HIT_DTYPE, n).view(np.uint32), hb.read(api.HIT_DTYPE, n).view(np.uint32))
    for b in (rb, hb, ib, lb, h1, i1, l1):
        b.destroy()
    prop.destroy()
    tlas.destroy()
    for a in blas:
        a.destroy()
