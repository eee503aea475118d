"""GPU checks of the learned treelet proposal: bf16-MFMA forward and backward against the FP64
CPU reference (oracle/mlp_ref.py, with bf16 rounding at the same points), Adam against the
closed form, the depth-12 treelet cut against the exported LBVH, proposal-first traces
bit-identical to plain traces, and the gain of a trained network over an untrained one."""
import ctypes as C

import numpy as np
import pytest

from gprt_amd import api, scenes
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(ctx):
    verts, idx = scenes.blob_scene(1 << 16, seed=21)
    acc = api.TriangleAccel(ctx, verts, idx)
    acc.build()
    yield verts, idx, acc
    acc.destroy()


def _logits(ctx, prop, rays):
    rb = ctx.device_buffer(rays)
    lb = ctx.device_buffer(nbytes=len(rays) * 4096 * 4)
    ctx.lib.gprtProposalLogits(ctx.h, prop.h, rb.h, len(rays), lb.h)
    z = lb.read(np.float32).reshape(len(rays), 4096)
    rb.destroy()
    lb.destroy()
    return z


def test_forward_matches_cpu(ctx):
    prop = api.Proposal(ctx, max_batch=512, seed=1)
    p = M.he_init(7).astype(np.float32)
    prop.set_params(p)
    rays = np.concatenate([scenes.random_rays(200, seed=3), scenes.pinhole_rays(8, 7)])
    z = _logits(ctx, prop, rays)
    ref, _, _ = M.forward(p.astype(np.float64), M.encode(rays), bf16=True)
    err = np.abs(z - ref).max()
    assert err <= 0.02 * np.abs(ref).max() + 1e-3, (err, np.abs(ref).max())
    assert np.mean(np.argmax(z, 1) == np.argmax(ref, 1)) > 0.95
    prop.destroy()


def test_fused_top1_matches_logits(ctx, scene):
    """gprtProposalInfer (fused: logits never leave registers) picks the top logit of the GEMM path.
    The fused kernel compares logits with their low 12 mantissa bits replaced by the class (relative
    resolution 2^-11) and accumulates in another order, so a near-tie may resolve either way."""
    prop = api.Proposal(ctx, max_batch=1024, seed=1)
    prop.set_params(M.he_init(11).astype(np.float32))
    rays = np.concatenate([scenes.random_rays(700, seed=13), scenes.pinhole_rays(17, 19)])  # 1023 rays
    n = len(rays)
    z = _logits(ctx, prop, rays)
    rb, cb = ctx.device_buffer(rays), ctx.device_buffer(nbytes=4 * n)
    for classes in (4096, None):
        if classes is None:
            classes = prop.bind(scene[2], 10)  # fewer than 4096 treelets: the argmax is restricted
        ctx.lib.gprtProposalInfer(ctx.h, prop.h, rb.h, n, cb.h)
        c = cb.read(np.int32, n)
        assert ((c >= 0) & (c < classes)).all()
        zz = z[:, :classes]
        top = zz.max(1)
        picked = zz[np.arange(n), c]
        assert (top - picked <= 1e-3 * np.maximum(1.0, np.abs(top))).all()
        assert np.mean(c == np.argmax(zz, 1)) > 0.97
    rb.destroy()
    cb.destroy()
    prop.destroy()


def test_gradients_match_cpu(ctx):
    prop = api.Proposal(ctx, max_batch=512, seed=1)
    p = M.he_init(8).astype(np.float32)
    prop.set_params(p)
    rays = scenes.random_rays(384, seed=4)
    labels = np.random.default_rng(1).integers(0, 4096, len(rays)).astype(np.int32)
    rb, lb = ctx.device_buffer(rays), ctx.device_buffer(labels)
    g = np.zeros(prop.n_params, dtype=np.float32)
    ctx.lib.gprtProposalGradients(ctx.h, prop.h, rb.h, lb.h, len(rays), g.ctypes.data_as(C.c_void_p), len(g))
    assert api.last_error() is None, api.last_error()
    _, ref = M.loss_and_grad(p.astype(np.float64), M.encode(rays), labels, bf16=True)
    for layer, (w0, shape, b0, bn) in enumerate(M.offsets()):
        for lo, hi in ((w0, w0 + shape[0] * shape[1]), (b0, b0 + bn)):
            gg, rr = g[lo:hi].astype(np.float64), ref[lo:hi]
            rel = np.linalg.norm(gg - rr) / max(np.linalg.norm(rr), 1e-12)
            assert rel < 3e-2, (layer, lo, rel)
    rb.destroy()
    lb.destroy()
    prop.destroy()


def test_treelet_cut(ctx, scene):
    verts, idx, acc = scene
    prop = api.Proposal(ctx, max_batch=256, seed=1)
    nc = prop.bind(acc, 12)
    assert 1 <= nc <= 4096
    crb, pcb = ctx.device_buffer(nbytes=4), ctx.device_buffer(nbytes=4)
    ctx.lib.gprtProposalExportTreelets(ctx.h, prop.h, crb.h, pcb.h)
    roots = crb.read(np.int32, nc)
    pclass = pcb.read(np.int32, len(idx))
    keys, nodes, _ = acc.export()
    n = len(idx)
    depth = np.full(2 * n - 1, -1)
    depth[0] = 0
    order = [0]
    for x in order:
        if x < n - 1:
            for c in nodes[x, :2]:
                depth[c] = depth[x] + 1
                order.append(c)
    ref_ids = [(r if r >= 0 else (n - 1) + ~r) for r in roots]
    for r in ref_ids:
        assert depth[r] == 12 or (r >= n - 1 and depth[r] <= 12)
    assert len(set(ref_ids)) == nc
    # every primitive's treelet root is an ancestor-or-self of its leaf
    leaf_of = np.empty(n, dtype=np.int64)
    leaf_of[nodes[n - 1:, 3]] = np.arange(n - 1, 2 * n - 1)
    assert np.all(pclass >= 0) and np.all(pclass < nc)
    for prim in np.random.default_rng(0).choice(n, 300, replace=False):
        x = leaf_of[prim]
        target = ref_ids[pclass[prim]]
        while x != target and x != 0:
            x = nodes[x, 2]
        assert x == target
    prop.destroy()


def _trace(ctx, acc, rays, prop=None):
    rb = ctx.device_buffer(rays)
    hb = ctx.device_buffer(nbytes=len(rays) * 16)
    vb = ctx.device_buffer(nbytes=len(rays) * 4)
    ctx.lib.gprtAccelTraceClosestHitEx(ctx.h, acc.h, prop.h if prop else None, rb.h, hb.h, len(rays), vb.h)
    assert api.last_error() is None, api.last_error()
    h, v = hb.read(api.HIT_DTYPE), vb.read(np.uint32)
    for b in (rb, hb, vb):
        b.destroy()
    return h, v


def _same(a, b):
    np.testing.assert_array_equal(a["primID"], b["primID"])
    np.testing.assert_array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    np.testing.assert_array_equal(a["u"].view(np.uint32), b["u"].view(np.uint32))


def test_proposal_trace_identical_and_trained_gain(ctx, scene):
    verts, idx, acc = scene
    prop = api.Proposal(ctx, max_batch=1024, seed=5)
    nc = prop.bind(acc, 12)
    test = scenes.pinhole_rays(128, 128)
    plain, v_plain = _trace(ctx, acc, test)
    test = np.concatenate([test, scenes.bounce_rays(test, plain, verts, idx, 8192, seed=9)])
    plain, v_plain = _trace(ctx, acc, test)
    hits0, v0 = _trace(ctx, acc, test, prop)  # untrained network
    _same(hits0, plain)
    # training rays: jittered primary + bounce from a different seed, labelled by plain traces
    tr = scenes.jittered_pinhole_rays(40000, seed=6)
    th, _ = _trace(ctx, acc, tr)
    tr = np.concatenate([tr, scenes.bounce_rays(tr, th, verts, idx, 20000, seed=7)])
    th, _ = _trace(ctx, acc, tr)
    rb, hb = ctx.device_buffer(tr), ctx.device_buffer(th)
    loss0, acc0 = prop.train(rb, hb, len(tr), epochs=1, lr=2e-3)
    loss, accuracy = prop.train(rb, hb, len(tr), epochs=8, lr=2e-3)
    assert np.isfinite(loss) and loss < loss0
    hits1, v1 = _trace(ctx, acc, test, prop)
    _same(hits1, plain)
    # held-out top-1 accuracy of the trained network vs chance
    cb = ctx.device_buffer(nbytes=len(test) * 4)
    tb = ctx.device_buffer(test)
    ctx.lib.gprtProposalInfer(ctx.h, prop.h, tb.h, len(test), cb.h)
    cls = cb.read(np.int32)
    pcb, crb = ctx.device_buffer(nbytes=4), ctx.device_buffer(nbytes=4)
    ctx.lib.gprtProposalExportTreelets(ctx.h, prop.h, crb.h, pcb.h)
    pclass = pcb.read(np.int32, len(idx))
    hit = plain["primID"] >= 0
    right = np.zeros(len(test), dtype=bool)
    right[hit] = cls[hit] == pclass[plain["primID"][hit]]
    top1 = right[hit].mean()
    print(f"treelets {nc}; train loss {loss0:.3f} -> {loss:.3f}; held-out top-1 {top1:.3f} (chance {1 / nc:.4f}); "
          f"inner nodes per ray: plain {v_plain.mean():.1f}, untrained {v0.mean():.1f}, trained {v1.mean():.1f}; "
          f"on correctly proposed rays: plain {v_plain[right].mean():.1f} -> proposal {v1[right].mean():.1f}")
    assert top1 > 20.0 / nc
    # the gain where the proposal is right: the candidate's hit prunes the exact traversal
    assert v1[right].mean() < v_plain[right].mean()
    for b in (rb, hb, cb, tb, pcb, crb):
        b.destroy()
    prop.destroy()


def test_adam_step_matches_closed_form(ctx, scene):
    verts, idx, acc = scene
    prop = api.Proposal(ctx, max_batch=4096, seed=11)
    prop.bind(acc, 12)
    p0 = prop.get_params()
    rays = scenes.jittered_pinhole_rays(3000, seed=12)
    hits, _ = _trace(ctx, acc, rays)
    pcb, crb = ctx.device_buffer(nbytes=4), ctx.device_buffer(nbytes=4)
    ctx.lib.gprtProposalExportTreelets(ctx.h, prop.h, crb.h, pcb.h)
    pclass = pcb.read(np.int32, len(idx))
    sel = hits["primID"] >= 0
    lab_rays = rays[sel]
    labels = pclass[hits["primID"][sel]].astype(np.int32)
    rb, lb = ctx.device_buffer(lab_rays), ctx.device_buffer(labels)
    g = np.zeros(prop.n_params, dtype=np.float32)
    ctx.lib.gprtProposalGradients(ctx.h, prop.h, rb.h, lb.h, len(lab_rays), g.ctypes.data_as(C.c_void_p), len(g))
    rb2, hb2 = ctx.device_buffer(rays), ctx.device_buffer(hits)
    prop.train(rb2, hb2, len(rays), epochs=1, lr=1e-3)  # one batch: all labelled rays fit
    p1 = prop.get_params()
    ref, _, _ = M.adam_step(p0.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, lr=1e-3)
    # colsum atomics make bias gradients order-dependent at FP32 rounding: compare with slack
    moved = np.abs(g) > 1e-4  # sign of the first Adam step is stable there
    np.testing.assert_allclose(p1[moved], ref[moved], rtol=0, atol=2e-5)
    for b in (rb, lb, rb2, hb2, pcb, crb):
        b.destroy()
    prop.destroy()


def test_checkpoint_round_trip(ctx, scene):
    """Weight checkpoints: header fields, the CRC against zlib's crc32, parameters and Adam state
    restored exactly (a re-serialised checkpoint is byte-identical), inference unchanged, and
    damaged or foreign data rejected."""
    import zlib
    verts, idx, acc = scene
    a = api.Proposal(ctx, max_batch=4096, seed=31)
    a.bind(acc, 12)
    rays = scenes.jittered_pinhole_rays(3000, seed=32)
    hits, _ = _trace(ctx, acc, rays)
    rb, hb = ctx.device_buffer(rays), ctx.device_buffer(hits)
    a.train(rb, hb, len(rays), epochs=2, lr=1e-3)
    blob = a.serialize(True)
    hdr = np.frombuffer(blob[:64], dtype=np.dtype([("magic", "S8"), ("version", "<u4"), ("hb", "<u4"),
                                                   ("params", "<u8"), ("dims", "<u4", 4), ("step", "<u4"),
                                                   ("flags", "<u4"), ("crc", "<u4"), ("res", "<u4", 3)]))[0]
    assert hdr["magic"] == b"GPRTMLP" and hdr["version"] == 1 and hdr["hb"] == 64
    assert hdr["params"] == a.n_params and list(hdr["dims"]) == [102, 128, 3, 4096]
    assert hdr["step"] > 0 and hdr["flags"] == 1 and not hdr["res"].any()
    assert len(blob) == 64 + 3 * 4 * a.n_params
    assert hdr["crc"] == zlib.crc32(blob[64:])
    np.testing.assert_array_equal(np.frombuffer(blob[64:64 + 4 * a.n_params], np.float32), a.get_params())

    b = api.Proposal(ctx, max_batch=4096, seed=33)
    b.bind(acc, 12)
    assert b.deserialize(blob), api.last_error()
    assert b.serialize(True) == blob  # parameters, both moments and the step
    np.testing.assert_array_equal(b.get_params(), a.get_params())
    n = 2048
    cls_a, cls_b = ctx.device_buffer(nbytes=4 * n), ctx.device_buffer(nbytes=4 * n)
    ctx.lib.gprtProposalInfer(ctx.h, a.h, rb.h, n, cls_a.h)
    ctx.lib.gprtProposalInfer(ctx.h, b.h, rb.h, n, cls_b.h)
    np.testing.assert_array_equal(cls_a.read(np.int32), cls_b.read(np.int32))

    plain = a.serialize(False)  # parameters only: the optimizer state starts over
    assert len(plain) == 64 + 4 * a.n_params
    assert b.deserialize(plain), api.last_error()
    again = b.serialize(True)
    assert np.frombuffer(again[40:44], "<u4")[0] == 0  # step reset
    assert not np.frombuffer(again[64 + 4 * a.n_params:], np.float32).any()  # moments zero
    np.testing.assert_array_equal(b.get_params(), a.get_params())

    ctx.lib.gprtClearLastError()
    bad = bytearray(blob)
    bad[64 + 1000] ^= 0x40
    assert not b.deserialize(bytes(bad)) and "checksum" in api.last_error()
    ctx.lib.gprtClearLastError()
    assert not b.deserialize(b"NOTAMLP!" + blob[8:]) and "magic" in api.last_error()
    ctx.lib.gprtClearLastError()
    assert not b.deserialize(blob[:-4]) and "expected" in api.last_error()
    ctx.lib.gprtClearLastError()
    wrong = bytearray(blob)
    wrong[8] = 2  # version
    assert not b.deserialize(bytes(wrong)) and "version" in api.last_error()
    ctx.lib.gprtClearLastError()
    assert not b.deserialize(blob[:10]) and "header" in api.last_error()
    ctx.lib.gprtClearLastError()
    np.testing.assert_array_equal(b.get_params(), a.get_params())  # rejected data changed nothing
    for buf in (rb, hb, cls_a, cls_b):
        buf.destroy()
    a.destroy()
    b.destroy()
