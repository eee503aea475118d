"""gprtBufferExclusiveSumNoAlloc (the scan that allocates nothing) through the C-ABI: output and
total identical to gprtBufferExclusiveSum on two calls in a row, the second smaller than the
first (the static state keeps the previous call's contents and must be cleared), and to numpy;
inputs beyond its limit are rejected before any GPU work."""
import numpy as np
import pytest

from gprt_amd import api, scenes

pytestmark = pytest.mark.gpu


def test_no_alloc_scan_matches_exclusive_sum(ctx):
    lib = ctx.lib
    for call, n in enumerate((5_000_017, 70_001)):
        x = (scenes.rand_u64(50 + call, 0, n) >> np.uint64(40)).astype(np.uint32)
        print(f"no-alloc scan: call {call}, {n} items", flush=True)
        inp = ctx.device_buffer(x)
        o1, o2 = ctx.device_buffer(nbytes=4 * n), ctx.device_buffer(nbytes=4 * n)
        t1 = lib.gprtBufferExclusiveSum(ctx.h, inp.h, o1.h, None)
        t2 = lib.gprtBufferExclusiveSumNoAlloc(ctx.h, inp.h, o2.h)
        assert api.last_error() is None, api.last_error()
        a, b = o1.read(np.uint32), o2.read(np.uint32)
        np.testing.assert_array_equal(b, a)
        assert t2 == t1
        c = np.cumsum(x, dtype=np.uint64)
        assert t2 == int(c[-1]) & 0xFFFFFFFF
        np.testing.assert_array_equal(b[1:], (c[:-1] & np.uint64(0xFFFFFFFF)).astype(np.uint32))
        assert b[0] == 0
        for buf in (inp, o1, o2):
            buf.destroy()


def test_no_alloc_scan_limits(ctx):
    lib = ctx.lib
    empty, small = ctx.device_buffer(nbytes=0), ctx.device_buffer(nbytes=4)
    assert lib.gprtBufferExclusiveSumNoAlloc(ctx.h, empty.h, small.h) == 0
    assert api.last_error() is None
    big = ctx.device_buffer(nbytes=4 * ((1 << 31) + 1))
    lib.gprtBufferExclusiveSumNoAlloc(ctx.h, big.h, small.h)
    err = api.last_error()
    assert err is not None and "limit" in err, err
    lib.gprtClearLastError()
    for buf in (empty, small, big):
        buf.destroy()
