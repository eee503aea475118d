"""Runtime behaviour on the GPU: the host-wait deadline (the reference's syncTDR) reports work that
outlives it and recovers once the work is done; builds and sorts are bit-identical run to run (a
race in the look-back or the hierarchy levels would show up as a difference)."""
import ctypes as C

import numpy as np
import pytest

from gprt_amd import api, scenes

pytestmark = pytest.mark.gpu


def test_sync_timeout_reports_long_work_then_recovers(ctx):
    lib = ctx.lib
    n = 1 << 24
    rng = np.random.default_rng(71)
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    bufs = [ctx.device_buffer(keys) for _ in range(3)]
    ctx.sync()
    lib.gprtClearLastError()
    lib.gprtSetSyncTimeout(ctx.h, 1e-6)
    for b in bufs:  # tens of milliseconds of queued sorts: no bug, only longer than the deadline
        lib.gprtBufferSort(ctx.h, b.h, None)
    lib.gprtComputeSynchronize(ctx.h)
    err = api.last_error()
    lib.gprtClearLastError()
    lib.gprtSetSyncTimeout(ctx.h, 6.0)
    lib.gprtComputeSynchronize(ctx.h)
    assert api.last_error() is None, api.last_error()
    assert err is not None and "sync timeout" in err
    want = np.sort(keys)
    for b in bufs:
        np.testing.assert_array_equal(b.read(np.uint64), want)
        b.destroy()
    lib.gprtSetSyncTimeout(ctx.h, -1.0)
    assert "not a duration" in api.last_error()
    lib.gprtClearLastError()


def test_builds_and_sorts_repeat_bit_for_bit(ctx):
    verts, idx = scenes.blob_scene(300_000, seed=72)
    acc = api.TriangleAccel(ctx, verts, idx)
    runs = []
    for _ in range(3):
        acc.build()
        runs.append([x.copy() for x in acc.export()])
    for r in runs[1:]:
        for x, y in zip(runs[0], r):
            np.testing.assert_array_equal(x.view(np.uint8), y.view(np.uint8))
    acc.destroy()
    sizes = [1000, 70000, 200000, 5]
    accs = [api.TriangleAccel(ctx, *scenes.blob_scene(s, seed=80 + i)) for i, s in enumerate(sizes)]
    first = None
    for _ in range(3):
        api.build_batch(ctx, accs)
        got = [[x.copy() for x in a.export()] for a in accs]
        if first is None:
            first = got
        else:
            for ga, fa in zip(got, first):
                for x, y in zip(ga, fa):
                    np.testing.assert_array_equal(x.view(np.uint8), y.view(np.uint8))
    for a in accs:
        a.destroy()
    rng = np.random.default_rng(73)
    keys = rng.integers(0, 1 << 32, 1 << 22, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(len(keys), dtype=np.uint32)
    outs = []
    for _ in range(2):
        kb, vb = ctx.device_buffer(keys), ctx.device_buffer(vals)
        ctx.lib.gprtBufferSortKV(ctx.h, kb.h, vb.h, len(keys), 32, 32, None)
        outs.append((kb.read(np.uint32), vb.read(np.uint32)))
        kb.destroy()
        vb.destroy()
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(outs[0][1], order.astype(np.uint32))
