"""Buffers at the sizes of the reference's own tests: full and offset copies of 10^8 u32
(tests/t00-bufferCopy/hostCode.cpp:30-99) and resizes from 0.5 x 10^9 to 10^9 u32 (4 GB) of host
and device buffers, with and without preserving the contents
(tests/t01-bufferResize/hostCode.cpp:29-146); and the reference's point LBVH test
(tests/t04-swBVH/hostCode.cpp:46-61: 42 points uniform in +-1e38, which it builds without checking)
as an AABB accel of degenerate boxes, checked against the oracle."""
import ctypes as C

import numpy as np
import pytest

from gprt_amd import api

pytestmark = pytest.mark.gpu


def _host_buffer(ctx, count):
    h = ctx.lib.gprtHostBufferCreate(ctx.h, 4, count, None, 16)
    assert h, api.last_error()
    return api.Buffer(ctx, h)


def _device_buffer(ctx, count):
    h = ctx.lib.gprtDeviceBufferCreate(ctx.h, 4, count, None, 16)
    assert h, api.last_error()
    return api.Buffer(ctx, h)


def test_copy_full_and_offset_t00(ctx):
    n = 100_000_000
    seq = np.arange(n, dtype=np.uint32)
    a, b = _device_buffer(ctx, n), _device_buffer(ctx, n)
    a.write(seq)
    ctx.lib.gprtBufferCopy(ctx.h, a.h, b.h, 0, 0, 4, n, 0, 0)
    np.testing.assert_array_equal(b.read(np.uint32), seq)
    b.write(np.zeros(n, np.uint32))
    ctx.lib.gprtBufferCopy(ctx.h, a.h, b.h, 1, 0, 4, n - 1, 0, 0)  # source offset 1, n - 1 elements
    got = b.read(np.uint32)
    np.testing.assert_array_equal(got[: n - 1], seq[1:])
    assert got[n - 1] == 0
    ctx.lib.gprtClearLastError()
    ctx.lib.gprtBufferCopy(ctx.h, a.h, b.h, 2, 0, 4, n - 1, 0, 0)  # one element past the source
    assert "out of bounds" in api.last_error()
    ctx.lib.gprtClearLastError()
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("kind", ["host", "device"])
@pytest.mark.parametrize("preserve", [False, True])
def test_resize_t01(ctx, kind, preserve):
    half, full = 500_000_000, 1_000_000_000
    buf = (_host_buffer if kind == "host" else _device_buffer)(ctx, half)
    seq = np.arange(half, dtype=np.uint32)
    buf.write(seq)
    ctx.lib.gprtBufferResize(ctx.h, buf.h, 4, full, preserve, 0)
    assert api.last_error() is None, api.last_error()
    assert buf.nbytes == full * 4
    if preserve:
        got = buf.read(np.uint32, half)
        np.testing.assert_array_equal(got, seq)
    del seq
    # the whole resized buffer is usable
    tail = np.full(full - half, 7, np.uint32)
    both = np.concatenate([np.zeros(half, np.uint32), tail])
    buf.write(both)
    np.testing.assert_array_equal(buf.read(np.uint32)[half:], tail)
    buf.destroy()


def test_point_lbvh_t04(ctx, orc):
    rng = np.random.default_rng(74)
    pts = (rng.random((42, 3)) * 2e38 - 1e38).astype(np.float32)
    boxes = np.concatenate([pts, pts], 1)  # a point is a box of zero extent
    acc = api.AABBAccel(ctx, boxes)
    acc.build()
    assert api.last_error() is None, api.last_error()
    k, nodes, aabbs = acc.export()
    rk, rn, ra = orc.lbvh_build_boxes(boxes)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))
    assert np.all(np.isfinite(aabbs))
    acc.destroy()
