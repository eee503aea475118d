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


# Bounds for results computed with bf16 operands and FP32 accumulation, against the reference
# that rounds to bf16 at the same points (oracle/mlp_ref.py):
#  * per element, 2^-6 of its condition sum (the same sum with every term in absolute value):
#    FP32 accumulation order moves an element by at most ~K 2^-24 of it (K <= 4096: 2.4e-4), and
#    a bf16 rounding flip that the order causes (an FP32 value within its error of a bf16
#    rounding boundary) moves one stored term by at most one bf16 ulp, 2^-7 of itself; the bound
#    leaves room for two such flips in one element;
#  * per tensor, the relative Frobenius error at most 2^-12: flips are rare and unsigned, so they
#    do not add up; a misplaced bias or operand column moves the whole tensor;
#  * per output class, the mean error over the rays at most 2^-10 of the mean condition sum: a
#    wrong bias or weight column shifts every ray the same way, flips average out.
# Each bound is shown to reject the mistakes it is there for (planted in the reference, on the
# CPU, below). A float32 emulation of the device's arithmetic on the CPU stays below 1e-7 (logits)
# and 6e-3 (worst gradient element, 3e-5 relative norm) of these sums.
ELEM_TOL = 2.0 ** -6
NORM_TOL = 2.0 ** -12
MEAN_TOL = 2.0 ** -10


def _worst(got, ref, cond):
    return float((np.abs(got.astype(np.float64) - ref) / (cond + 1e-30)).max())


def _relnorm(got, ref):
    return float(np.linalg.norm(got.astype(np.float64) - ref) / max(np.linalg.norm(ref), 1e-300))


def _colmean(got, ref, cond):
    return float((np.abs((got.astype(np.float64) - ref).mean(0)) / (cond.mean(0) + 1e-30)).max())


def _logits_ok(z, ref, cond):
    return _worst(z, ref, cond) <= ELEM_TOL and _relnorm(z, ref) <= NORM_TOL and _colmean(z, ref, cond) <= MEAN_TOL


def test_forward_matches_cpu(ctx):
    prop = api.Proposal(ctx, max_batch=512, seed=1)
    p = M.he_init(7).astype(np.float32)
    prop.set_params(p)
    rays = np.concatenate([scenes.random_rays(200, seed=3), scenes.pinhole_rays(8, 7)])
    z = _logits(ctx, prop, rays)
    p64 = p.astype(np.float64)
    X = M.encode(rays)
    ref, acts, _ = M.forward(p64, X, bf16=True)
    cond = M.logit_cond(p64, acts)
    print(f"logits: worst element {_worst(z, ref, cond):.3e} of its condition sum, relative norm "
          f"{_relnorm(z, ref):.3e}, worst class mean {_colmean(z, ref, cond):.3e}", flush=True)
    assert _worst(z, ref, cond) <= ELEM_TOL
    assert _relnorm(z, ref) <= NORM_TOL
    assert _colmean(z, ref, cond) <= MEAN_TOL
    assert np.mean(np.argmax(z, 1) == np.argmax(ref, 1)) > 0.95
    # Mistakes planted in the reference fail the same checks: one output bias off by 0.05 ...
    bad = ref.copy()
    bad[:, 1234] += 0.05
    assert not _logits_ok(z, bad, cond)
    # ... two 16-class fragment columns of the output layer exchanged (rows 16..31 and 32..47 of
    # W4t) ...
    w0, shape, _, _ = M.offsets()[3]
    q = p64.copy()
    W4 = q[w0:w0 + shape[0] * shape[1]].reshape(shape)
    W4[16:32], W4[32:48] = W4[32:48].copy(), W4[16:32].copy()
    assert not _logits_ok(z, M.forward(q, X, bf16=True)[0], cond)
    # ... and one hidden bias moved to the next lane.
    q = p64.copy()
    hb0 = M.offsets()[1][2]
    q[hb0 + 5], q[hb0 + 6] = q[hb0 + 6], q[hb0 + 5]
    assert not _logits_ok(z, M.forward(q, X, bf16=True)[0], cond)
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
    X = M.encode(rays)
    _, ref, cond = M.loss_and_grad(p.astype(np.float64), X, labels, bf16=True, bf16_grads=True, with_cond=True)
    tensors = []
    for layer, (w0, shape, b0, bn) in enumerate(M.offsets()):
        tensors += [(f"W{layer + 1}", slice(w0, w0 + shape[0] * shape[1])), (f"b{layer + 1}", slice(b0, b0 + bn))]
    for name, sl in tensors:
        worst, rel = _worst(g[sl], ref[sl], cond[sl]), _relnorm(g[sl], ref[sl])
        print(f"{name}: worst element {worst:.3e} of its condition sum, relative norm {rel:.3e}", flush=True)
        assert worst <= ELEM_TOL and rel <= NORM_TOL, (name, worst, rel)
    # Mistakes planted in the reference fail: the output-layer bias gradient shifted by one lane,
    # and a hidden layer's weight gradient with two 16-column fragments exchanged.
    w0, shape, b0, bn = M.offsets()[3]
    sb = slice(b0, b0 + bn)
    bad = np.roll(ref[sb], 1)
    assert _worst(g[sb], bad, cond[sb]) > ELEM_TOL or _relnorm(g[sb], bad) > NORM_TOL
    w0, shape, _, _ = M.offsets()[1]
    sw = slice(w0, w0 + shape[0] * shape[1])
    bad = ref[sw].reshape(shape).copy()
    bad[:, 0:16], bad[:, 16:32] = bad[:, 16:32].copy(), bad[:, 0:16].copy()
    assert _worst(g[sw], bad.ravel(), cond[sw]) > ELEM_TOL or _relnorm(g[sw], bad.ravel()) > NORM_TOL
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
    # Classes are numbered breadth-first, left to right within a level (the queue order above), so
    # the numbering depends on the tree alone: binding again gives the same table.
    assert ref_ids == [x for x in order if depth[x] == 12 or (x >= n - 1 and depth[x] <= 12)]
    again = api.Proposal(ctx, max_batch=256, seed=2)
    assert again.bind(acc, 12) == nc
    crb2, pcb2 = ctx.device_buffer(nbytes=4), ctx.device_buffer(nbytes=4)
    ctx.lib.gprtProposalExportTreelets(ctx.h, again.h, crb2.h, pcb2.h)
    np.testing.assert_array_equal(crb2.read(np.int32, nc), roots)
    np.testing.assert_array_equal(pcb2.read(np.int32, len(idx)), pclass)
    again.destroy()
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


def test_confidence_gate(ctx, scene):
    """The confidence gate: with +inf no ray keeps its proposal (node fetches equal the plain
    trace's), with -inf every ray does, and a gate calibrated on training rays keeps the sample's
    total node fetches at most the plain trace's; hits are identical to the plain trace in every
    case, and gprtProposalInfer is not gated."""
    verts, idx, acc = scene
    prop = api.Proposal(ctx, max_batch=1024, seed=41)
    prop.bind(acc, 12)
    tr = scenes.jittered_pinhole_rays(30000, seed=42)
    th, _ = _trace(ctx, acc, tr)
    tr = np.concatenate([tr, scenes.bounce_rays(tr, th, verts, idx, 15000, seed=43)])
    th, _ = _trace(ctx, acc, tr)
    rb, hb = ctx.device_buffer(tr), ctx.device_buffer(th)
    prop.train(rb, hb, len(tr), epochs=6, lr=2e-3)
    test = scenes.pinhole_rays(96, 96)
    plain, v_plain = _trace(ctx, acc, test)
    test = np.concatenate([test, scenes.bounce_rays(test, plain, verts, idx, 6000, seed=44)])
    plain, v_plain = _trace(ctx, acc, test)
    lib = ctx.lib
    assert lib.gprtProposalGetGate(prop.h) == -np.inf
    h_all, v_all = _trace(ctx, acc, test, prop)
    _same(h_all, plain)
    lib.gprtProposalSetGate(ctx.h, prop.h, np.inf)
    # A gate that admits no ray skips the inference as well: the trace is the plain one.
    lib.gprtKernelTimerEnable(ctx.h, True)
    lib.gprtKernelTimerReset(ctx.h)
    h_none, v_none = _trace(ctx, acc, test, prop)
    ctx.sync()
    launches = C.c_uint32(0)
    lib.gprtKernelTimerQuery(ctx.h, b"mlp_infer", C.byref(launches))
    assert launches.value == 0
    lib.gprtKernelTimerQuery(ctx.h, b"trace_closest_proposal", C.byref(launches))
    assert launches.value == 0
    lib.gprtKernelTimerQuery(ctx.h, b"trace_closest", C.byref(launches))
    assert launches.value == 1
    lib.gprtKernelTimerEnable(ctx.h, False)
    _same(h_none, plain)
    np.testing.assert_array_equal(v_none, v_plain)
    # gprtProposalInfer ignores the gate.
    tb, cb = ctx.device_buffer(test), ctx.device_buffer(nbytes=4 * len(test))
    lib.gprtProposalInfer(ctx.h, prop.h, tb.h, len(test), cb.h)
    assert (cb.read(np.int32) >= 0).all()
    # Calibrated on (a sample of) the training rays.
    n_cal = 20000
    keep = lib.gprtProposalCalibrateGate(ctx.h, prop.h, acc.h, rb.h, n_cal)
    assert api.last_error() is None, api.last_error()
    gate = lib.gprtProposalGetGate(prop.h)
    assert 0.0 <= keep <= 1.0 and not np.isnan(gate)
    _, vc_plain = _trace(ctx, acc, tr[:n_cal])
    _, vc_gated = _trace(ctx, acc, tr[:n_cal], prop)
    print(f"gate {gate:.3f} keeps {keep:.3f} of the calibration rays: node fetches plain {vc_plain.sum()} gated "
          f"{vc_gated.sum()}; test rays: plain {v_plain.sum()}, all proposed {v_all.sum()}", flush=True)
    # A ray's fetch count depends a little on the rays that share its wave (leaf tests are batched
    # across the wave), so the gated total is compared with a 1% margin.
    assert vc_gated.sum() <= 1.01 * vc_plain.sum()
    h_gated, _ = _trace(ctx, acc, test, prop)
    _same(h_gated, plain)
    for b in (rb, hb, tb, cb):
        b.destroy()
    prop.destroy()


def test_overlapped_proposal_trace(ctx, scene):
    """Proposal-first traces with inference overlapped on the context's second stream (chunks of
    the rays, fewer traversal workgroups per CU): hits identical to the plain trace and node
    fetches identical to the trace without overlap; a call right after it (a map) sees the whole
    result."""
    verts, idx, acc = scene
    prop = api.Proposal(ctx, max_batch=1024, seed=51)
    prop.bind(acc, 12)
    tr = scenes.jittered_pinhole_rays(20000, seed=52)
    th, _ = _trace(ctx, acc, tr)
    rb, hb = ctx.device_buffer(tr), ctx.device_buffer(th)
    prop.train(rb, hb, len(tr), epochs=4, lr=2e-3)
    test = scenes.pinhole_rays(512, 512)
    plain, _ = _trace(ctx, acc, test)
    test = np.concatenate([test, scenes.bounce_rays(test, plain, verts, idx, 150_001, seed=53)])
    plain, _ = _trace(ctx, acc, test)
    serial, v_serial = _trace(ctx, acc, test, prop)
    _same(serial, plain)
    for chunks, blocks in ((4, 8), (7, 6), (64, 0), (2, 14)):
        ctx.lib.gprtProposalSetOverlap(ctx.h, prop.h, chunks, blocks)
        print(f"overlapped proposal trace: {len(test)} rays, {chunks} chunks, {blocks} workgroups/CU", flush=True)
        got, v = _trace(ctx, acc, test, prop)
        np.testing.assert_array_equal(got.view(np.uint32).reshape(-1, 4), plain.view(np.uint32).reshape(-1, 4))
        np.testing.assert_array_equal(v, v_serial)
    ctx.lib.gprtProposalSetOverlap(ctx.h, prop.h, 0, 0)
    for b in (rb, hb):
        b.destroy()
    prop.destroy()
