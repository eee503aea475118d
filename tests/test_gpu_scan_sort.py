"""GPU parity of the scan family and the radix sort against the CPU oracle, through the C-ABI.

Mirrors the reference tests: t03 (tests/t03-bufferScans/hostCode.cpp: exclusive sum of all
ones over N = 9973*(r+2), partition / select of i*(+-1)) and t02 (tests/t02-bufferSort/
hostCode.cpp: key-only random keys, key-value with unique shuffled keys), plus stability,
wrap-around, edge sizes and the full-size microbenchmarks. Bar: bit-exact.
"""
import ctypes as C

import numpy as np
import pytest

from gprt_amd import api

pytestmark = pytest.mark.gpu


def _scan(ctx, x, mode, sel_pos=True, inplace=False, scratch="buffer"):
    lib = ctx.lib
    inp = ctx.device_buffer(x)
    out = inp if inplace else ctx.device_buffer(nbytes=x.nbytes)
    scr = ctx.device_buffer(nbytes=4) if scratch == "buffer" else None
    sh = scr.h if scr else None
    if mode == "sum":
        total = lib.gprtBufferExclusiveSum(ctx.h, inp.h, out.h, sh)
    elif mode == "partition":
        total = lib.gprtBufferPartition(ctx.h, inp.h, sel_pos, out.h, sh)
    else:
        total = lib.gprtBufferSelect(ctx.h, inp.h, sel_pos, out.h, sh)
    assert api.last_error() is None, api.last_error()
    res = out.read(x.dtype)
    for b in {id(inp): inp, id(out): out}.values():
        b.destroy()
    if scr:
        scr.destroy()
    return res, int(total)


@pytest.mark.parametrize("rnd", [0, 1, 10, 50, 99])
def test_exclusive_sum_t03(ctx, orc, rnd):
    n = 9973 * (rnd + 2)
    x = np.ones(n, dtype=np.uint32)
    out, total = _scan(ctx, x, "sum")
    assert total == n
    np.testing.assert_array_equal(out, np.arange(n, dtype=np.uint32))


# Edge sizes around a lane's 4-item group, a wave (64 lanes), a wave's 2,816-item chunk and the
# sum's 45,056-item tile (16 waves x 64 lanes x 44 items).
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 63, 64, 65, 2815, 2816, 2817, 4095, 4096, 4097, 8191, 8192, 8193,
                               45055, 45056, 45057, 2 * 45056 + 1, 1000003])
def test_exclusive_sum_random(ctx, orc, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)  # wraps modulo 2^32
    out, total = _scan(ctx, x, "sum")
    ref, rtotal = orc.exclusive_sum(x)
    assert total == rtotal
    np.testing.assert_array_equal(out, ref)


def test_exclusive_sum_inplace_and_null_scratch(ctx, orc):
    x = np.random.default_rng(1).integers(0, 1000, 777777).astype(np.uint32)
    ref, rtotal = orc.exclusive_sum(x)
    out, total = _scan(ctx, x, "sum", inplace=True)
    assert total == rtotal
    np.testing.assert_array_equal(out, ref)
    out, total = _scan(ctx, x, "sum", scratch=None)
    assert total == rtotal
    np.testing.assert_array_equal(out, ref)


def _t03(n):
    i = np.arange(n, dtype=np.int64)
    return (i * np.where(i & 1, 1, -1)).astype(np.int32)


@pytest.mark.parametrize("mode", ["partition", "select"])
@pytest.mark.parametrize("sel_pos", [True, False])
@pytest.mark.parametrize("n", [1, 2, 9973 * 2, 9973 * 50, 4096 * 3 + 1])
def test_partition_select_t03(ctx, orc, mode, sel_pos, n):
    x = _t03(n)
    out, total = _scan(ctx, x, mode, sel_pos)
    ref, rtotal = (orc.partition if mode == "partition" else orc.select)(x, sel_pos)
    assert total == rtotal
    if mode == "partition":
        np.testing.assert_array_equal(out, ref)
    else:
        np.testing.assert_array_equal(out[:total], ref[:total])


@pytest.mark.parametrize("mode", ["partition", "select"])
def test_partition_select_random(ctx, orc, mode):
    rng = np.random.default_rng(4)
    x = rng.integers(-2**31, 2**31, 2_000_001, dtype=np.int64).astype(np.int32)
    for sel_pos in (True, False):
        out, total = _scan(ctx, x, mode, sel_pos)
        ref, rtotal = (orc.partition if mode == "partition" else orc.select)(x, sel_pos)
        assert total == rtotal
        np.testing.assert_array_equal(out[:total] if mode == "select" else out,
                                      ref[:total] if mode == "select" else ref)


def test_exclusive_sum_full_size(ctx):
    # 2^28 u32 uniform in [0, 255]: the exclusive-scan microbenchmark size; the sum wraps.
    n = 1 << 28
    rng = np.random.default_rng(5)
    x = rng.integers(0, 256, n, dtype=np.uint8).astype(np.uint32)
    out, total = _scan(ctx, x, "sum")
    c = np.cumsum(x, dtype=np.uint64)
    assert total == int(c[-1]) & 0xFFFFFFFF
    assert out[0] == 0
    np.testing.assert_array_equal(out[1:], (c[:-1] & np.uint64(0xFFFFFFFF)).astype(np.uint32))


# ------------------------------------------------------------------------------------------------
# Sort
# ------------------------------------------------------------------------------------------------
def _sort_kv(ctx, keys, vals=None, count=None):
    lib = ctx.lib
    kb = ctx.device_buffer(keys)
    vb = ctx.device_buffer(vals) if vals is not None else None
    n = len(keys) if count is None else count
    lib.gprtBufferSortKV(ctx.h, kb.h, vb.h if vb else None, n, keys.itemsize * 8,
                         0 if vals is None else vals.itemsize * 8, None)
    assert api.last_error() is None, api.last_error()
    k = kb.read(keys.dtype)
    v = vb.read(vals.dtype) if vb else None
    kb.destroy()
    if vb:
        vb.destroy()
    return k, v


def test_sort_t02_keys_only_u64(ctx):
    # t02 (a) through gprtBufferSort (64-bit keys, as the reference reads them)
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 2**64, 1_000_000, dtype=np.uint64)
    kb = ctx.device_buffer(keys)
    ctx.lib.gprtBufferSort(ctx.h, kb.h, None)
    out = kb.read(np.uint64)
    np.testing.assert_array_equal(out, np.sort(keys))


def test_sort_t02_payload_unique_shuffled(ctx):
    # t02 (b) through gprtBufferSortPayload: unique keys 0..N-1 shuffled, random payloads
    rng = np.random.default_rng(3)
    n = 1_000_000
    keys = rng.permutation(n).astype(np.uint64)
    vals = rng.integers(0, 2**64, n, dtype=np.uint64)
    kb, vb, sb = ctx.device_buffer(keys), ctx.device_buffer(vals), ctx.device_buffer(nbytes=8)
    ctx.lib.gprtBufferSortPayload(ctx.h, kb.h, vb.h, sb.h)
    k, v = kb.read(np.uint64), vb.read(np.uint64)
    np.testing.assert_array_equal(k, np.arange(n, dtype=np.uint64))
    inv = np.empty(n, dtype=np.int64)
    inv[keys.astype(np.int64)] = np.arange(n)
    np.testing.assert_array_equal(v, vals[inv])
    assert sb.nbytes >= n * 16  # undersized scratch was resized in place


@pytest.mark.parametrize("bits", [4, 6, 8])
@pytest.mark.parametrize("tile", [2048, 4096, 8192, 16384])
def test_sort_stable_all_configs(ctx, orc, bits, tile):
    ctx.lib.gprtSetSortConfig(ctx.h, bits, tile, 0)
    try:
        rng = np.random.default_rng(bits * 10000 + tile)
        n = 300_007
        keys = rng.integers(0, 3000, n).astype(np.uint32)  # heavy duplicates
        vals = np.arange(n, dtype=np.uint32)
        k, v = _sort_kv(ctx, keys, vals)
        rk, rv = orc.stable_sort(keys, vals)
        np.testing.assert_array_equal(k, rk)
        np.testing.assert_array_equal(v, rv)
    finally:
        ctx.lib.gprtSetSortConfig(ctx.h, 8, 0, 0)


@pytest.mark.parametrize("kdt,vdt", [(np.uint32, None), (np.uint64, None), (np.uint64, np.uint32),
                                     (np.uint32, np.uint64), (np.uint64, np.uint64)])
def test_sort_widths(ctx, orc, kdt, vdt):
    rng = np.random.default_rng(9)
    n = 123_457
    hi = 2**32 if kdt == np.uint32 else 2**64
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(kdt)
    keys[::7] = keys[3]  # duplicates
    vals = None if vdt is None else np.arange(n).astype(vdt)
    k, v = _sort_kv(ctx, keys, vals)
    rk, rv = orc.stable_sort(keys, vals)
    np.testing.assert_array_equal(k, rk)
    if vals is not None:
        np.testing.assert_array_equal(v, rv)


@pytest.mark.parametrize("n", [2, 3, 63, 64, 65, 4095, 4096, 4097, 8192, 8193])
def test_sort_edge_sizes(ctx, orc, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = _sort_kv(ctx, keys, vals)
    rk, rv = orc.stable_sort(keys, vals)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(v, rv)


def test_sort_constant_and_single_digit_keys(ctx, orc):
    n = 50_000
    # every digit constant: all passes skipped, data untouched
    keys = np.full(n, 0xABCDEF01, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = _sort_kv(ctx, keys, vals)
    np.testing.assert_array_equal(k, keys)
    np.testing.assert_array_equal(v, vals)
    # only the top byte varies: one active pass (odd count -> copy back)
    keys = (np.random.default_rng(1).integers(0, 256, n).astype(np.uint32) << np.uint32(24)) | np.uint32(5)
    k, v = _sort_kv(ctx, keys, vals)
    rk, rv = orc.stable_sort(keys, vals)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(v, rv)


def test_sort_partial_count(ctx, orc):
    # count smaller than the buffer: the tail is left untouched
    keys = np.random.default_rng(8).integers(0, 2**32, 10_000, dtype=np.uint64).astype(np.uint32)
    k, _ = _sort_kv(ctx, keys, None, count=6000)
    np.testing.assert_array_equal(k[:6000], np.sort(keys[:6000]))
    np.testing.assert_array_equal(k[6000:], keys[6000:])


@pytest.mark.parametrize("payload", [False, True])
@pytest.mark.parametrize("n", [(1 << 22) - 1, (1 << 22) + 16_383])
def test_sort_automatic_tile_threshold(ctx, orc, payload, n):
    """Buffer sorts take 16,384-key tiles from 2^22 keys and 8,192-key tiles below: both sides of
    the threshold, ragged last tiles, keys only and with payloads, duplicates in the masked digits."""
    rng = np.random.default_rng(n + payload)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) & np.uint32(0xF0F0FFFF)
    vals = np.arange(n, dtype=np.uint32) if payload else None
    k, v = _sort_kv(ctx, keys, vals)
    rk, rv = orc.stable_sort(keys, vals)
    np.testing.assert_array_equal(k, rk)
    if payload:
        np.testing.assert_array_equal(v, rv)


def test_sort_full_size_microbench(ctx):
    # 2^26 u32 keys uniform over 32 bits with u32 payload = index (sort microbenchmark size)
    n = 1 << 26
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    k, v = _sort_kv(ctx, keys, vals)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, order.astype(np.uint32))


@pytest.mark.parametrize("bits,payload,n", [(32, False, 300_001), (64, False, 300_001), (32, True, 300_001),
                                            (64, True, 300_001), (32, True, 1), (64, True, 8193)])
def test_width_explicit_sorts(ctx, bits, payload, n):
    """gprtBufferSort32/64 and gprtBufferSortPayload32/64 (include/gprt_hip_ext.h): the count is the
    key buffer's size over the key width; duplicated keys keep their payload order (stable)."""
    rng = np.random.default_rng(bits + n + payload)
    kt = np.uint32 if bits == 32 else np.uint64
    keys = rng.integers(0, 1 << 20, n, dtype=np.uint64).astype(kt) * kt(3 if bits == 32 else (1 << 40) + 1)
    vals = np.arange(n, dtype=kt)
    kb = ctx.device_buffer(keys)
    vb = ctx.device_buffer(vals) if payload else None
    fn = getattr(ctx.lib, f"gprtBufferSort{'Payload' if payload else ''}{bits}")
    if payload:
        fn(ctx.h, kb.h, vb.h, None)
    else:
        fn(ctx.h, kb.h, None)
    assert api.last_error() is None, api.last_error()
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(kb.read(kt), keys[order])
    if payload:
        np.testing.assert_array_equal(vb.read(kt), vals[order])
        vb.destroy()
    kb.destroy()
