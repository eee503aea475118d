"""GPU parity of box-primitive accels (gprtAABBsSetPositions + gprtAABBAccelCreate) against the
CPU oracle: sorted keys, topology and node boxes bit-exact (single builds, batches mixed with
triangle meshes, refits), and closest hits (primID exact, t bit-identical) against the oracle's
brute force over every box, plainly and with a learned proposal."""
import numpy as np
import pytest

from gprt_amd import api, scenes

from test_oracle_boxes import random_boxes

pytestmark = pytest.mark.gpu


def _check(acc, orc, boxes):
    k, nodes, aabbs = acc.export()
    rk, rn, ra = orc.lbvh_build_boxes(boxes)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.random((n, 3), dtype=np.float32) * np.float32(1.4) - np.float32(0.2)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    r = np.zeros(n, dtype=scenes.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = o[:, 0], o[:, 1], o[:, 2]
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["tmin"], r["tmax"] = 0.0, scenes.TMAX
    return r


def _trace(ctx, acc, rays):
    rb = ctx.device_buffer(rays)
    hb = ctx.device_buffer(nbytes=16 * len(rays))
    acc.trace(rb, hb, len(rays))
    h = hb.read(api.HIT_DTYPE)
    rb.destroy()
    hb.destroy()
    return h


def _same_hits(h, ref):
    np.testing.assert_array_equal(h["primID"], ref["primID"])
    np.testing.assert_array_equal(h["t"].view(np.uint32), ref["t"].view(np.uint32))


@pytest.mark.parametrize("n", [1, 2, 3, 100, 4097, 150001])
def test_box_lbvh_matches_oracle(ctx, orc, n):
    boxes = random_boxes(n, 100 + n)
    acc = api.AABBAccel(ctx, boxes)
    acc.build()
    assert api.last_error() is None, api.last_error()
    _check(acc, orc, boxes)
    acc.destroy()


def test_box_trace_matches_bruteforce(ctx, orc):
    boxes = random_boxes(20000, 7)
    acc = api.AABBAccel(ctx, boxes)
    acc.build()
    rays = _rays(4096, 8)
    h = _trace(ctx, acc, rays)
    ref = orc.trace_boxes_bruteforce(rays, boxes)
    _same_hits(h, ref)
    assert (h["primID"] >= 0).mean() > 0.5  # most rays hit something (many start inside a box)
    assert np.all(h["u"] == 0) and np.all(h["v"] == 0)
    acc.destroy()


def test_box_batch_mixed_with_triangles(ctx, orc):
    b1, b2 = random_boxes(5000, 21), random_boxes(70000, 22)
    v, idx = scenes.blob_scene(30000, seed=23)
    accs = [api.AABBAccel(ctx, b1), api.TriangleAccel(ctx, v, idx), api.AABBAccel(ctx, b2)]
    api.build_batch(ctx, accs)
    assert api.last_error() is None, api.last_error()
    _check(accs[0], orc, b1)
    _check(accs[2], orc, b2)
    k, nodes, aabbs = accs[1].export()
    rk, rn, ra = orc.lbvh_build(v, idx)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))
    rays = _rays(2048, 24)
    _same_hits(_trace(ctx, accs[2], rays), orc.trace_boxes_bruteforce(rays, b2))
    for a in accs:
        a.destroy()


def test_box_update_refit(ctx, orc):
    boxes = random_boxes(100000, 31)
    acc = api.AABBAccel(ctx, boxes)
    acc.build(api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    _, nodes0, _ = acc.export()
    moved = (boxes + np.repeat(scenes.rand01(32, 0, len(boxes) * 3).reshape(-1, 3).astype(np.float32)
                               * np.float32(2e-3) - np.float32(1e-3), 2, axis=0).reshape(-1, 6)).astype(np.float32)
    acc.vbuf.write(moved)
    acc.update()
    _, nodes1, aabbs1 = acc.export()
    np.testing.assert_array_equal(nodes1, nodes0)  # topology kept
    np.testing.assert_array_equal(aabbs1.view(np.uint32), orc.box_refit(moved, nodes0).view(np.uint32))
    rays = _rays(2048, 33)
    _same_hits(_trace(ctx, acc, rays), orc.trace_boxes_bruteforce(rays, moved))
    acc.destroy()


def test_box_proposal_is_exact(ctx, orc):
    boxes = random_boxes(50000, 41)
    acc = api.AABBAccel(ctx, boxes)
    acc.build()
    train = _rays(8192, 42)
    rays = _rays(4096, 43)
    prop = api.Proposal(ctx, max_batch=1024, seed=5)
    classes = prop.bind(acc, 8)
    assert classes > 1
    tb = ctx.device_buffer(train)
    th = ctx.device_buffer(_trace(ctx, acc, train))
    prop.train(tb, th, len(train), epochs=1)
    rb = ctx.device_buffer(rays)
    plain = _trace(ctx, acc, rays)
    hb = ctx.device_buffer(nbytes=16 * len(rays))
    ctx.lib.gprtAccelTraceClosestHitEx(ctx.h, acc.h, prop.h, rb.h, hb.h, len(rays), None)
    learned = hb.read(api.HIT_DTYPE)
    _same_hits(learned, plain)
    _same_hits(plain, orc.trace_boxes_bruteforce(rays, boxes))
    for b in (tb, th, rb, hb):
        b.destroy()
    prop.destroy()
    acc.destroy()
