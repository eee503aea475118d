"""The benchmark workloads at their full sizes, checked against the oracle.

* Large scene: 4,194,304 blob triangles (seed 2). Sorted keys, topology and every node box
  bit-exact against the oracle's LBVH; all 8,388,608 rays (2048^2 pinhole primary + 4,194,304
  cosine-weighted bounce rays from their hits) traced, a seeded sample of 131,072 compared with
  the oracle's BVH traversal over its own tree (primID, t, u, v bits).
* Learned proposal on that scene: trained on 16,777,216 rays (jittered primary, seed 6, + their
  bounce rays, seed 16) labelled by the plain traversal, cut at depth 12; the proposal-first
  trace of all 8,388,608 rays must equal the plain trace bit for bit.
* Batched builds: 4,096 device-generated blob meshes (sizes log-uniform 64..262,144, seed 3;
  mesh i from seed 1000 + i) through the autotuner; every mesh's root box equals the box of its
  vertices, 64 sampled meshes (the largest and the smallest among them) equal the oracle's LBVH.
* Dynamic scene: 64 x 65,536 triangles rebuilt per frame under an instance accel, 2,097,152 rays
  per frame, proposal fine-tuned every 16 frames on the frame's own hits; 32 frames, learned ==
  plain every frame, and one frame's sample against the oracle per instance.
"""
import ctypes as C
import json

import numpy as np
import pytest

from gprt_amd import api, scenes
from gprt_amd.dynamic import DynamicScene

pytestmark = pytest.mark.gpu

N_TRIS = 4_194_304
IMG = 2048
ORACLE_THREADS = 16


def _log(msg):
    print(f"[configs] {msg}", flush=True)


@pytest.fixture(scope="module")
def scene(ctx):
    _log("large scene: generating 4,194,304 triangles")
    verts, idx = scenes.blob_scene(N_TRIS, seed=2)
    acc = api.TriangleAccel(ctx, verts, idx)
    acc.build()
    assert api.last_error() is None, api.last_error()
    prim = scenes.pinhole_rays(IMG, IMG)
    rb = ctx.device_buffer(prim)
    hb = ctx.device_buffer(nbytes=16 * len(prim))
    acc.trace(rb, hb, len(prim))
    ph = hb.read(api.HIT_DTYPE)
    rb.destroy()
    hb.destroy()
    rays = np.concatenate([prim, scenes.bounce_rays(prim, ph, verts, idx, N_TRIS, seed=3)])
    state = {"verts": verts, "idx": idx, "acc": acc, "rays": rays}
    yield state
    acc.destroy()


@pytest.fixture(scope="module")
def oracle_tree(scene, orc):
    _log("large scene: oracle LBVH")
    return orc.lbvh_build(scene["verts"], scene["idx"])


def test_large_scene_lbvh_bit_exact(scene, oracle_tree):
    keys, nodes, aabbs = scene["acc"].export()
    rk, rn, ra = oracle_tree
    np.testing.assert_array_equal(keys, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))
    np.testing.assert_array_equal(api.accel_bounds(scene["acc"].ctx, scene["acc"]).view(np.uint32),
                                  ra[0].view(np.uint32))


@pytest.fixture(scope="module")
def plain_hits(ctx, scene):
    rays = scene["rays"]
    n = len(rays)
    rb, hb = ctx.device_buffer(rays), ctx.device_buffer(nbytes=16 * n)
    _log(f"large scene: plain trace of {n} rays")
    scene["acc"].trace(rb, hb, n)
    hits = hb.read(api.HIT_DTYPE)
    assert api.last_error() is None, api.last_error()
    yield rb, hb, hits
    rb.destroy()
    hb.destroy()


def test_large_scene_trace_sample(scene, oracle_tree, plain_hits, orc):
    rays = scene["rays"]
    hits = plain_hits[2]
    assert len(hits) == 8_388_608
    # Every hit lies inside its ray's open interval; misses report tmax.
    hit = hits["primID"] >= 0
    assert (hits["t"][hit] > rays["tmin"][hit]).all() and (hits["t"][hit] < rays["tmax"][hit]).all()
    assert 0.3 < hit.mean() < 0.9
    sel = np.sort(np.random.default_rng(20).choice(len(rays), 131_072, replace=False))
    rk, rn, ra = oracle_tree
    ref, _ = orc.trace_bvh(rays[sel], scene["verts"], scene["idx"], rn, ra, threads=ORACLE_THREADS)
    got = hits[sel]
    np.testing.assert_array_equal(got["primID"], ref["primID"])
    np.testing.assert_array_equal(got.view(np.uint32).reshape(-1, 4), ref.view(np.uint32).reshape(-1, 4))
    # A few rays against every triangle (brute force), primary and bounce.
    few = np.concatenate([sel[:16], sel[-16:]])
    bf = orc.trace_bruteforce(rays[few], scene["verts"], scene["idx"], threads=ORACLE_THREADS)
    np.testing.assert_array_equal(hits[few].view(np.uint32).reshape(-1, 4), bf.view(np.uint32).reshape(-1, 4))


def test_large_scene_learned_proposal_identical(ctx, scene, plain_hits):
    acc, verts, idx = scene["acc"], scene["verts"], scene["idx"]
    rb, _, hits = plain_hits
    n = len(hits)
    # Training data: 16,777,216 rays of the same distribution with other seeds, labelled by the
    # plain traversal.
    tr = scenes.jittered_pinhole_rays(n, seed=6)
    tb, thb = ctx.device_buffer(tr), ctx.device_buffer(nbytes=16 * len(tr))
    acc.trace(tb, thb, len(tr))
    th = thb.read(api.HIT_DTYPE)
    tb.destroy()
    thb.destroy()
    tr = np.concatenate([tr, scenes.bounce_rays(tr, th, verts, idx, n, seed=16)])
    assert len(tr) == 16_777_216
    tb, thb = ctx.device_buffer(tr), ctx.device_buffer(nbytes=16 * len(tr))
    acc.trace(tb, thb, len(tr))
    prop = api.Proposal(ctx, max_batch=4096, seed=1)
    classes = prop.bind(acc, 12)
    assert 2048 < classes <= 4096
    _log(f"proposal: {classes} treelets, training on {len(tr)} rays x 2 epochs")
    loss, top1 = prop.train(tb, thb, len(tr), epochs=2, lr=1e-3)
    assert api.last_error() is None, api.last_error()
    assert np.isfinite(loss) and top1 > 0.3, (loss, top1)
    tb.destroy()
    thb.destroy()
    hb2 = ctx.device_buffer(nbytes=16 * n)
    _log(f"proposal-first trace of {n} rays")
    ctx.lib.gprtAccelTraceClosestHitEx(ctx.h, acc.h, prop.h, rb.h, hb2.h, n, None)
    got = hb2.read(api.HIT_DTYPE)
    assert api.last_error() is None, api.last_error()
    np.testing.assert_array_equal(got.view(np.uint32).reshape(-1, 4), hits.view(np.uint32).reshape(-1, 4))
    hb2.destroy()
    prop.destroy()


def test_batched_4096_meshes_autotuned(ctx, orc):
    sizes = scenes.batch_sizes(4096, seed=3)
    total = int(sizes.sum())
    _log(f"batch: {len(sizes)} meshes, {total} triangles")
    # One vertex and one index buffer per mesh (geometry offsets are 32-bit, as in the reference's
    # gprtTrianglesSetVertices, and the batch's 4.7 GB of vertices would not fit one buffer's).
    accs, vbs, ibs = [], [], []
    for i, n in enumerate(sizes):
        vb, ib = api.synth_blob_meshes(ctx, sizes[i:i + 1], 1000 + i, 0.002)
        accs.append(api.TriangleAccel.from_buffers(ctx, vb, ib, 3 * int(n), int(n)))
        vbs.append(vb)
        ibs.append(ib)
    arr = (C.c_void_p * len(accs))(*[a.h for a in accs])
    rep = C.create_string_buffer(1 << 14)
    try:
        best = ctx.lib.gprtAccelBuildBatchAutotune(ctx.h, arr, len(accs), 1, rep, len(rep))
        assert api.last_error() is None, api.last_error()
        r = json.loads(rep.value.decode())
        assert len(r["configs"]) == 30 and best == int(np.argmin([c["ms"] for c in r["configs"]]))
        got = [C.c_uint32(0) for _ in range(3)]
        ctx.lib.gprtGetSortConfig(ctx.h, *[C.byref(g) for g in got])
        assert [g.value for g in got] == [r["best"][k] for k in ("digitBits", "tileKeys", "wavesPerCU")]
        _log(f"batch: autotuned {r['best']}; reading back {36 * total >> 20} MiB of vertices")
        verts = np.concatenate([b.read(np.float32) for b in vbs]).reshape(-1, 3)
        idx = np.concatenate([b.read(np.uint32) for b in ibs]).reshape(-1, 3)
        first = np.concatenate([[0], np.cumsum(sizes)])
        # Root box of every mesh = the box of its vertices (3 per triangle, all referenced).
        lo = np.minimum.reduceat(verts, 3 * first[:-1], axis=0)
        hi = np.maximum.reduceat(verts, 3 * first[:-1], axis=0)
        roots = np.stack([api.accel_bounds(ctx, a) for a in accs])
        want = np.concatenate([lo, hi], 1)
        bad = np.nonzero((roots.view(np.uint32) != want.view(np.uint32)).any(1))[0]
        rng = np.random.default_rng(21)
        sample = set(rng.choice(len(sizes), 62, replace=False).tolist())
        sample |= {int(np.argmax(sizes)), int(np.argmin(sizes))}
        if len(bad):
            _log(f"batch: {len(bad)} root boxes differ, e.g. meshes {bad[:8].tolist()} of sizes "
                 f"{sizes[bad[:8]].tolist()}: {roots[bad[0]].tolist()} vs {want[bad[0]].tolist()}")
            sample |= set(bad[:4].tolist())
        _log(f"batch: {len(sample)} meshes against the oracle ({int(sizes[list(sample)].sum())} triangles)")
        for m in sorted(sample):
            v = verts[3 * first[m]: 3 * first[m + 1]]
            ix = idx[first[m]: first[m + 1]]
            k, nodes, aabbs = accs[m].export()
            rk, rn, ra = orc.lbvh_build(v, ix)
            np.testing.assert_array_equal(k, rk, err_msg=f"mesh {m}")
            np.testing.assert_array_equal(nodes, rn, err_msg=f"mesh {m}")
            np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32), err_msg=f"mesh {m}")
        assert len(bad) == 0, f"{len(bad)} root boxes differ from the boxes of their vertices"
    finally:
        ctx.lib.gprtSetSortConfig(ctx.h, 8, 0, 0)
        for a in accs:
            a.destroy()


def _oracle_instances(orc, rays, trees):
    """Identity instances: the minimum of (t, instance, primID) over the oracle's BVH traversal of
    each mesh."""
    n = len(rays)
    best_t = rays["tmax"].astype(np.float32).copy()
    best = np.zeros(n, dtype=api.HIT_DTYPE)
    best["primID"] = -1
    best_i = np.full(n, -1, np.int32)
    for k, (v, ix, nodes, aabbs) in enumerate(trees):
        h, _ = orc.trace_bvh(rays, v, ix, nodes, aabbs, threads=ORACLE_THREADS)
        better = (h["primID"] >= 0) & ((best["primID"] < 0) | (h["t"] < best_t))
        best[better] = h[better]
        best_t[better] = h["t"][better]
        best_i[better] = k
    best["t"][best["primID"] < 0] = rays["tmax"][best["primID"] < 0]
    return best, best_i


def test_dynamic_scene_32_frames(ctx, orc):
    ds = DynamicScene(ctx)
    _log(f"dynamic: {ds.meshes} x {ds.per} triangles, {ds.rays_per_frame} rays/frame, {ds.classes} classes")
    tuned = []
    n = ds.n
    for f in range(32):
        t = ds.frame(f)
        if t is not None:
            tuned.append((f, t))
        assert api.last_error() is None, api.last_error()
        bad = ds.mismatches()
        assert bad == 0, f"frame {f}: {bad} rays differ between the learned and the plain trace"
        if f == 17:
            # One frame against the oracle: 4,096 primary and 4,096 bounce rays, per instance.
            verts = ds.cur.read(np.float32).reshape(-1, 3)
            idx = ds.ibuf.read(np.uint32).reshape(-1, 3)
            trees = []
            for m in range(ds.meshes):
                v = verts[3 * m * ds.per: 3 * (m + 1) * ds.per]
                ix = idx[m * ds.per: (m + 1) * ds.per]
                _, rn, ra = orc.lbvh_build(v, ix)
                trees.append((v, ix, rn, ra))
            sel = np.sort(np.random.default_rng(22).choice(n, 4096, replace=False))
            for rb, hb, ib in ((ds.pb, ds.hp, ds.ip), (ds.bb, ds.hb, ds.ib)):
                rays = rb.read(api.RAY_DTYPE, n)[sel]
                ref, ref_i = _oracle_instances(orc, rays, trees)
                got, got_i = hb.read(api.HIT_DTYPE, n)[sel], ib.read(np.int32, n)[sel]
                np.testing.assert_array_equal(got["primID"], ref["primID"])
                np.testing.assert_array_equal(got_i, ref_i)
                np.testing.assert_array_equal(got.view(np.uint32).reshape(-1, 4), ref.view(np.uint32).reshape(-1, 4))
    assert [f for f, _ in tuned] == [0, 16]
    for f, ((l1, a1), (l2, a2)) in tuned:
        assert np.isfinite(l1) and np.isfinite(l2)
    # The gate was calibrated at each fine-tune (learned hits matched the plain ones above whatever
    # it kept).
    assert ds.gate_keep is not None and 0.0 <= ds.gate_keep <= 1.0
    assert not np.isnan(ctx.lib.gprtProposalGetGate(ds.prop.h))
    _log(f"dynamic: fine-tunes {tuned}; proposal hit rate {ds.proposal_hit_rate():.3f}; gate keeps {ds.gate_keep:.4f}")
    ds.destroy()
