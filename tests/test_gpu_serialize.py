"""Serialised accels (gprtAccelSerialize / gprtAccelDeserialize): a built triangle or AABB accel
written to bytes and loaded into a new accel over the same geometry is the same accel (export
bit-identical to the original and to the oracle, identical traces, identical refits, usable under
an instance accel); header fields and the CRC are as documented; damaged or mismatched data is
rejected and leaves the target as it was."""
import zlib

import numpy as np
import pytest

from gprt_amd import api, scenes

from test_oracle_boxes import random_boxes

pytestmark = pytest.mark.gpu

HDR = np.dtype([("magic", "S8"), ("version", "<u4"), ("hb", "<u4"), ("kind", "<u4"), ("prims", "<u4"),
                ("nodeBytes", "<u4"), ("recordBytes", "<u4"), ("mode", "<u4"), ("crc", "<u4"), ("res", "<u4", 6)])


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


def _same_export(a, b):
    for x, y in zip(a.export(), b.export()):
        np.testing.assert_array_equal(x.view(np.uint8), y.view(np.uint8))


def test_triangle_accel_round_trip(ctx, orc):
    verts, idx = scenes.blob_scene(50_000, seed=61)
    a = api.TriangleAccel(ctx, verts, idx)
    a.build(api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    blob = a.serialize()
    n = len(idx)
    h = np.frombuffer(blob[:64], HDR)[0]
    assert h["magic"] == b"GPRTBVH" and h["version"] == 1 and h["hb"] == 64 and h["kind"] == 0
    assert h["prims"] == n and h["nodeBytes"] == 64 and h["recordBytes"] == 48 and not h["res"].any()
    assert h["mode"] == api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE
    assert len(blob) == 64 + (n - 1) * 64 + n * 48 + n * 8 + 64
    assert h["crc"] == zlib.crc32(blob[64:])

    b = api.TriangleAccel(ctx, verts, idx)
    assert b.deserialize(blob), api.last_error()
    _same_export(a, b)
    k, nodes, boxes = b.export()
    rk, rn, ra = orc.lbvh_build(verts, idx)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(boxes.view(np.uint32), ra.view(np.uint32))
    rays = _rays(4096, 62)
    ha, hb = _trace(ctx, a, rays), _trace(ctx, b, rays)
    np.testing.assert_array_equal(ha.view(np.uint32), hb.view(np.uint32))
    assert b.serialize() == blob

    # the loaded accel refits from its geometry like the original
    moved = (verts + scenes.rand01(63, 0, verts.size).reshape(verts.shape).astype(np.float32) * np.float32(2e-3)
             - np.float32(1e-3)).astype(np.float32)
    a.vbuf.write(moved)
    b.vbuf.write(moved)
    a.update()
    b.update()
    assert api.last_error() is None, api.last_error()
    _same_export(a, b)
    np.testing.assert_array_equal(b.export()[2].view(np.uint32), orc.refit(moved, idx, nodes).view(np.uint32))

    # and serves an instance accel as a built one does
    recs = api.instance_records([b])
    ia = api.InstanceAccel(ctx, recs)
    ia.build()
    hi = _trace(ctx, ia, rays)
    hb2 = _trace(ctx, b, rays)
    np.testing.assert_array_equal(hi["primID"], hb2["primID"])
    np.testing.assert_array_equal(hi["t"].view(np.uint32), hb2["t"].view(np.uint32))
    ia.destroy()
    a.destroy()
    b.destroy()


def test_aabb_accel_round_trip(ctx, orc):
    boxes = random_boxes(30_000, 64)
    a = api.AABBAccel(ctx, boxes)
    a.build()
    blob = a.serialize()
    assert np.frombuffer(blob[:64], HDR)[0]["kind"] == 1
    b = api.AABBAccel(ctx, boxes)
    assert b.deserialize(blob), api.last_error()
    _same_export(a, b)
    rk, rn, ra = orc.lbvh_build_boxes(boxes)
    k, nodes, bb = b.export()
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(bb.view(np.uint32), ra.view(np.uint32))
    rays = _rays(2048, 65)
    hb = _trace(ctx, b, rays)
    ref = orc.trace_boxes_bruteforce(rays, boxes)
    np.testing.assert_array_equal(hb["primID"], ref["primID"])
    np.testing.assert_array_equal(hb["t"].view(np.uint32), ref["t"].view(np.uint32))
    a.destroy()
    b.destroy()


def test_serialised_data_is_checked(ctx):
    lib = ctx.lib
    verts, idx = scenes.uniform_scene(3000, seed=66)
    a = api.TriangleAccel(ctx, verts, idx)
    a.build()
    blob = a.serialize()
    target = api.TriangleAccel(ctx, verts, idx)
    target.build()
    before = target.serialize()

    def rejected(data, words):
        lib.gprtClearLastError()
        ok = target.deserialize(data)
        err = api.last_error() or ""
        lib.gprtClearLastError()
        return not ok and words in err

    bad = bytearray(blob)
    bad[64 + 5000] ^= 0x10
    assert rejected(bytes(bad), "checksum")
    assert rejected(b"NOTABVH!" + blob[8:], "magic")
    assert rejected(blob[:-8], "expected")
    assert rejected(blob[:32], "header")
    v2 = bytearray(blob)
    v2[8] = 9
    assert rejected(bytes(v2), "version")
    small_v, small_i = scenes.uniform_scene(2999, seed=67)
    other = api.TriangleAccel(ctx, small_v, small_i)
    lib.gprtClearLastError()
    assert not other.deserialize(blob) and "primitives" in (api.last_error() or "")
    lib.gprtClearLastError()
    box = api.AABBAccel(ctx, random_boxes(3000, 68))
    assert not box.deserialize(blob)
    assert "triangle" in (api.last_error() or "")
    lib.gprtClearLastError()
    assert target.serialize() == before  # rejected data changed nothing
    fresh = api.TriangleAccel(ctx, verts, idx)
    assert lib.gprtAccelSerialize(ctx.h, fresh.h, None, 0) == 0
    assert "not built" in (api.last_error() or "")
    lib.gprtClearLastError()
    ia = api.InstanceAccel(ctx, api.instance_records([a]))
    ia.build()
    assert lib.gprtAccelSerialize(ctx.h, ia.h, None, 0) == 0
    assert "instance" in (api.last_error() or "")
    lib.gprtClearLastError()
    for x in (ia, fresh, box, other, target, a):
        x.destroy()
