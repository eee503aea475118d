"""Benchmark: LBVH build Mprims/s + closest-hit Mrays/s (plain and learned proposal), one MI355X.

Default workload (the large single scene): 4,194,304 triangles from 64 Gaussian blobs (FP32);
8,388,608 rays = 2048x2048 pinhole primary rays + 4,194,304 cosine-weighted bounce rays from the
primary hits. A step is one full LBVH build (bounds, Morton keys, key sort, hierarchy + boxes)
followed by a closest-hit trace of all rays, once plainly and once with the learned treelet
proposal (top-1 treelet from the bf16 MFMA MLP, traversed first to tighten tmax, then the exact
traversal). Every input is resident in HBM before timing starts. Before the timed steps the
proposal network is trained on the device on rays drawn from the same distribution with other
seeds (jittered primary + bounce), labelled by the plain traversal's hits; training is reported,
not timed in the step. Hits of the two traces are compared for every ray in the timed steps.

`value` is the build throughput (primitives / build time). Build and trace times come from HIP
events the library records around every kernel on its stream during the timed steps; the whole
timed region is also wall-clocked between two stream synchronisations. Roofline entries use
algorithmic bytes per launch (fixed by the workload, not by tuning knobs):
  * LBVH sort pass (u32 Morton key + u32 primID payload): read 8 B + write 8 B per key;
  * trace: 32 B ray + 16 B hit + 64 B per inner node fetched + 48 B per triangle tested, with
    the node / triangle counts measured by the CPU reference traversal on a ray sample;
  * proposal inference: 2 x (102 x 128 + 2 x 128^2 + 128 x classes) FLOP per ray against the bf16
    dense MFMA peak (classes: the treelets bound, 2,897 on this scene; at most 4,096).
The CPU baseline times the oracle (single thread) on a bounded sample and is reported beside it.

Other workloads (one JSON line each): --workload batched (4,096 meshes, log-uniform 64..262,144
triangles: per-mesh launch sequence vs one hipGraph-captured batch launch vs the autotuned
configuration); dynamic (64 meshes rebuilt per frame under an instance accel, plain vs learned
traces); refit (refit-only updates of the blob scene); boxes (LBVH build and trace over 4,194,304
AABB primitives); compute (a user program through gprtComputeLaunch); sort (the radix-sort
microbenchmark: 2^26 u32 keys + u32 payloads, 68 B/key for the whole sort, 16 B/key per pass);
scan (the exclusive-scan microbenchmark: 2^28 u32, 8 B/element); copy (HBM counter calibration).
bench.py --gpus N with N > 1 exits: this build measures one GPU.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--workload scene|batched|dynamic|refit|boxes|compute|sort|scan|copy]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
def mlp_flop_per_ray(classes):
    """Inference FLOP per ray: encode-width input layer, two hidden layers, and the output layer
    over the bound classes (the output kernel multiplies only the 128-class chunks that hold them)."""
    return 2 * (102 * 128 + 2 * 128 * 128 + 128 * classes)


def mlp_out_flop_per_ray(classes):
    return 2 * 128 * classes
N_TRIS = 4_194_304
IMG = 2048
N_BOUNCE = 4_194_304
LBVH_BYTES_PER_PRIM = 368
# HBM-side bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this workload
# (FETCH_SIZE x 2.0 and WRITE_SIZE x 1.0, the factors a plain 1 GiB copy gives): read + write.
# The commands and per-kernel tables are in profiles/hbm_bytes_build_trace.md.
TRAFFIC_BYTES = {
    "trace_closest": 28060.8e6 + 138.6e6,  # the plain trace of the workload's 8,388,608 rays
    "mlp_infer_fused": 275.9e6 + 33.6e6,  # one launch of 8,388,608 rays, 2,897 classes bound
    "lbvh_sort_pass": 30.9e6 + 35.3e6,  # one pass over 4,194,304 keys on 16,384-key tiles (mean of 4)
    # per primitive, summed over the build kernels (init, bounds, keys, histogram scan, 4 sort
    # passes, sub-tile, region, top)
    "build_per_prim": 0.3 + 60.1 + 16.2 + (4.4 + 8.6) + 3 * (8.4 + 8.4) + 263.8 + 1.1,
    "batched_build_per_prim": 0.7 + 60.1 + 16.5 + 0.1 + 4 * (7.61 + 9.18) + 220.1 + 1.4 + 0.04,
    "boxes_build_per_prim": 0.3 + 36.1 + 16.2 + (4.3 + 8.3) + 3 * (8.4 + 8.4) + 262.0 + 1.1,
    "refit_per_prim": 12.0 + 263.4 + 1.2,  # identity check, sub-tile (vertex gather), region + top
    "trace_boxes": 14894.1e6 + 69.4e6,  # one launch of 4,194,304 rays
    "compute_bounds": 268.5e6 + 402.7e6,  # one launch over 16,777,216 spheres
    "sort_onesweep": 556.5e6 + 566.3e6,  # one pass over 2^26 u32 keys + u32 payloads (16,384-key tiles)
    "scan_lookback": 1075.5e6 + 1078.1e6,  # one launch over 2^28 u32
}
# Vector L1 and L2 hit rates of the traversal kernel on this workload (rocprofv3 --pmc TCP_* / TCC_*;
# commands and per-kernel table in profiles/cache_hit_rates_scene.md).
TRACE_CACHE_HIT = {"l1": 0.807, "l2": 0.687}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_report(lib, ctx):
    buf = C.create_string_buffer(1 << 16)
    lib.gprtKernelTimerReport(ctx.h, buf, len(buf))
    out = {}
    for line in buf.value.decode().strip().splitlines():
        name, n, ms = line.split(",")
        out[name] = {"launches": int(n), "total_ms": float(ms)}
    return out


def ms_of(per_kernel, prefix):
    ms = sum(r["total_ms"] for k, r in per_kernel.items() if k.startswith(prefix))
    n = sum(r["launches"] for k, r in per_kernel.items() if k.startswith(prefix))
    return ms, n


def _median_time(fn, setup=None, reps=5, warmup=1):
    """Seconds per call: one warm-up, then the median of `reps` timed calls (steady clock); `setup`
    runs untimed before each call."""
    ts = []
    for i in range(warmup + reps):
        if setup is not None:
            setup()
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        if i >= warmup:
            ts.append(dt)
    return float(np.median(ts))


def _cpu_baseline(value, unit, threads, sample, single_value, single_sample, **extra):
    """cpu_baseline record: the all-cores OpenMP leg as `value` (`cores` = its threads), the
    single-thread oracle leg beside it, and the host the two ran on (read at run time)."""
    from oracle import binding as orc
    info = orc.host_info()
    return {"value": value, "unit": unit, "cores": threads, "kind": "port", "sample": sample,
            "single_thread": {"value": single_value, "unit": unit, "cores": 1, "sample": single_sample},
            "all_cores_vs_single": value / single_value if single_value else None,
            "timing": "steady clock, 1 warm-up, median of 5", **info, **extra}


def _cpu_threads():
    from oracle import binding as orc
    return int(orc.lib().orc_max_threads())



def proposal_fetch_split(lib, ctx, acc, prop, rb, n_rays, hits, n_prims):
    """Inner nodes fetched per ray, plain and proposal-first without the gate, split by what the
    proposal was: the ray's true treelet (the treelet of the triangle its plain trace hits), a
    wrong treelet (or a ray that hits nothing), or the root class (no treelet walk). `best_gate`
    is the mean over rays of the smaller of the two counts: what a per-ray gate that knew the
    outcome could reach. Untimed; the counts come from the traversal kernel itself."""
    from gprt_amd import api
    cls_b = ctx.device_buffer(nbytes=n_rays * 4)
    root_b = ctx.device_buffer(nbytes=4096 * 4)
    pc_b = ctx.device_buffer(nbytes=n_prims * 4)
    v0_b = ctx.device_buffer(nbytes=n_rays * 4)
    v1_b = ctx.device_buffer(nbytes=n_rays * 4)
    tmp_h = ctx.device_buffer(nbytes=n_rays * 16)
    lib.gprtProposalInfer(ctx.h, prop.h, rb.h, n_rays, cls_b.h)
    lib.gprtProposalExportTreelets(ctx.h, prop.h, root_b.h, pc_b.h)
    lib.gprtAccelTraceClosestHitEx(ctx.h, acc.h, None, rb.h, tmp_h.h, n_rays, v0_b.h)
    gate = float(lib.gprtProposalGetGate(prop.h))
    lib.gprtProposalSetGate(ctx.h, prop.h, float("-inf"))
    lib.gprtAccelTraceClosestHitEx(ctx.h, acc.h, prop.h, rb.h, tmp_h.h, n_rays, v1_b.h)
    lib.gprtProposalSetGate(ctx.h, prop.h, gate)
    cls = cls_b.read(np.int32)
    roots = root_b.read(np.int32)
    prim_class = pc_b.read(np.int32)
    v0 = v0_b.read(np.uint32).astype(np.int64)
    v1 = v1_b.read(np.uint32).astype(np.int64)
    for b in (cls_b, root_b, pc_b, v0_b, v1_b, tmp_h):
        b.destroy()
    pid = hits["primID"]
    true_cls = np.where(pid >= 0, prim_class[np.clip(pid, 0, n_prims - 1)], -1)
    valid = (cls >= 0) & (cls < len(roots))
    at_root = valid & (roots[np.clip(cls, 0, len(roots) - 1)] == 0)
    right = valid & ~at_root & (cls == true_cls)
    wrong = ~at_root & ~right
    out = {"rays": int(n_rays), "plain": float(v0.mean()), "proposal_first": float(v1.mean()),
           "best_gate": float(np.minimum(v0, v1).mean())}
    for name, m in (("right_treelet", right), ("wrong_treelet_or_miss", wrong), ("root_class", at_root)):
        k = int(m.sum())
        out[name] = {"share": k / n_rays, "plain": float(v0[m].mean()) if k else 0.0,
                     "proposal_first": float(v1[m].mean()) if k else 0.0,
                     "cheaper_share": float((v1[m] < v0[m]).mean()) if k else 0.0}
    return out

def scene_workload(args):
    from gprt_amd import api, scenes

    lib = api.load()
    t0 = time.time()
    verts, idx = scenes.blob_scene(N_TRIS, seed=2)
    prim = scenes.pinhole_rays(IMG, IMG)
    log(f"scene + primary rays generated in {time.time() - t0:.1f}s")
    ctx = api.Context(0)
    acc = api.TriangleAccel(ctx, verts, idx)
    acc.build()

    def trace_np(rays):
        rb = ctx.device_buffer(rays)
        hb = ctx.device_buffer(nbytes=len(rays) * 16)
        acc.trace(rb, hb, len(rays))
        h = hb.read(api.HIT_DTYPE)
        rb.destroy()
        hb.destroy()
        return h

    ph = trace_np(prim)
    bounce = scenes.bounce_rays(prim, ph, verts, idx, N_BOUNCE, seed=3)
    rays = np.concatenate([prim, bounce])
    n_rays = len(rays)
    log(f"rays ready; primary hit rate {(ph['primID'] >= 0).mean():.3f}")

    # Proposal training data: same distribution, other seeds.
    t0 = time.time()
    tr = scenes.jittered_pinhole_rays(args.train_rays // 2, seed=6)
    th = trace_np(tr)
    tr = np.concatenate([tr, scenes.bounce_rays(tr, th, verts, idx, args.train_rays // 2, seed=16)])
    th = trace_np(tr)
    gen_s = time.time() - t0
    prop = api.Proposal(ctx, max_batch=args.batch, seed=1)
    classes = prop.bind(acc, 12)
    trb, thb = ctx.device_buffer(tr), ctx.device_buffer(th)
    t0 = time.time()
    loss, train_acc = prop.train(trb, thb, len(tr), epochs=args.epochs, lr=args.lr)
    train_s = time.time() - t0
    # Confidence gate: calibrated on a sample of the training rays (not the timed ones) from the
    # node fetches each ray costs plainly and proposal-first.
    t0 = time.time()
    gate_keep = float(lib.gprtProposalCalibrateGate(ctx.h, prop.h, acc.h, trb.h, min(args.gate_rays, len(tr))))
    gate = float(lib.gprtProposalGetGate(prop.h))
    gate_s = time.time() - t0
    trb.destroy()
    thb.destroy()
    log(f"proposal: {classes} treelets; trained on {len(tr)} rays x {args.epochs} epochs in {train_s:.1f}s "
        f"(data {gen_s:.1f}s): loss {loss:.3f}, train top-1 {train_acc:.3f}; gate {gate:.3f} keeps "
        f"{gate_keep:.3f} of {min(args.gate_rays, len(tr))} calibration rays ({gate_s:.1f}s)")

    rb = ctx.device_buffer(rays)
    hb = ctx.device_buffer(nbytes=n_rays * 16)
    hb2 = ctx.device_buffer(nbytes=n_rays * 16)

    def learned_trace():
        """The learned trace's device time (inference + proposal-first traversal, however they
        overlap), between two events on the context's stream."""
        lib.gprtBeginProfile(ctx.h)
        lib.gprtAccelTraceClosestHitEx(ctx.h, acc.h, prop.h, rb.h, hb2.h, n_rays, None)
        return float(lib.gprtEndProfile(ctx.h))

    # Overlap of inference and traversal: a short sweep of chunk counts and traversal workgroups
    # per CU (results do not depend on it) on the ungated trace, where the inference always runs;
    # the fastest is used from then on.
    lib.gprtProposalSetGate(ctx.h, prop.h, float("-inf"))
    overlap_sweep = {}
    for chunks, blocks in [(0, 0)] + [(c, b) for c in args.overlap_chunks for b in args.overlap_blocks]:
        lib.gprtProposalSetOverlap(ctx.h, prop.h, chunks, blocks)
        learned_trace()
        overlap_sweep[f"{chunks}x{blocks}"] = min(learned_trace() for _ in range(2))
    best_overlap = min(overlap_sweep, key=overlap_sweep.get)
    oc, ob = (int(x) for x in best_overlap.split("x"))
    lib.gprtProposalSetOverlap(ctx.h, prop.h, oc, ob)
    lib.gprtProposalSetGate(ctx.h, prop.h, gate)
    log(f"overlap sweep (ungated learned trace ms): {overlap_sweep}; using {best_overlap}")
    learned_wall = []

    def step():
        # The rebuild is deterministic (same keys, same tree), so the proposal's treelet cut stays
        # valid; after a geometry change a stale cut would cost speed, never correctness.
        acc.build()
        acc.trace(rb, hb, n_rays)
        learned_wall.append(learned_trace())

    for _ in range(args.warmup):
        step()
    ctx.sync()
    learned_wall.clear()
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    wall = time.perf_counter() - t0
    lib.gprtKernelTimerEnable(ctx.h, False)
    K = args.steps
    per_kernel = kernel_report(lib, ctx)
    for name, rec in sorted(per_kernel.items(), key=lambda kv: -kv[1]["total_ms"]):
        log(f"  {name:26s} {rec['launches']:5d} launches {rec['total_ms'] / K:9.3f} ms/step "
            f"{rec['total_ms'] / rec['launches']:9.4f} ms/launch")
    build_ms = ms_of(per_kernel, "lbvh_")[0] / K
    cut_ms = (ms_of(per_kernel, "mlp_treelet")[0] + ms_of(per_kernel, "mlp_leaf_class")[0]) / K
    # The plain traversal kernel's time per launch: when the gate admits no ray, the learned trace
    # launches that same kernel (and no inference), so launches are averaged rather than counted
    # per step.
    plain_rec = per_kernel.get("trace_closest", {"launches": 0, "total_ms": 0.0})
    plain_ms = plain_rec["total_ms"] / max(plain_rec["launches"], 1)
    prop_trace_ms = ms_of(per_kernel, "trace_closest_proposal")[0] / K
    # Inference = the fused encoding / hidden / output-layer / argmax kernel (per step, summed over
    # its chunks); the learned trace's time is its device wall time (inference and traversal may
    # overlap).
    infer_ms_tot, _ = ms_of(per_kernel, "mlp_infer")
    infer_ms = infer_ms_tot / K
    learned_ms = float(np.median(learned_wall))
    h1, h2 = hb.read(api.HIT_DTYPE), hb2.read(api.HIT_DTYPE)
    mismatches = int((h1.view(np.uint32).reshape(-1, 4) != h2.view(np.uint32).reshape(-1, 4)).any(1).sum())
    hit_rate = float((h1["primID"] >= 0).mean())

    # The ungated learned trace (every ray walks its proposed treelet first; the inference always
    # runs), after the timed steps: its wall time, its kernels' times and its hits.
    lib.gprtProposalSetGate(ctx.h, prop.h, float("-inf"))
    learned_trace()
    ctx.sync()
    ungated_wall = []
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    ungated_runs = max(1, min(K, 5))
    for _ in range(ungated_runs):
        ungated_wall.append(learned_trace())
    ctx.sync()
    lib.gprtKernelTimerEnable(ctx.h, False)
    ungated_kernels = kernel_report(lib, ctx)
    lib.gprtProposalSetGate(ctx.h, prop.h, gate)
    ungated_ms = float(np.median(ungated_wall))
    ungated_infer_ms = ms_of(ungated_kernels, "mlp_infer")[0] / ungated_runs
    ungated_traverse_ms = ms_of(ungated_kernels, "trace_closest_proposal")[0] / ungated_runs
    h3 = hb2.read(api.HIT_DTYPE)
    mismatches_ungated = int((h1.view(np.uint32).reshape(-1, 4) != h3.view(np.uint32).reshape(-1, 4)).any(1).sum())
    log(f"learned trace: gated {learned_ms:.3f} ms (inference {infer_ms:.3f} ms/step), ungated {ungated_ms:.3f} ms "
        f"(inference {ungated_infer_ms:.3f} + traversal {ungated_traverse_ms:.3f}); plain {plain_ms:.3f} ms")

    fetch_split = proposal_fetch_split(lib, ctx, acc, prop, rb, n_rays, h1, N_TRIS)
    log(f"proposal node fetches per ray: {json.dumps(fetch_split)}")

    sw_ms, sw_n = ms_of(per_kernel, "lbvh_sort_pass")
    sw_avg = sw_ms / max(sw_n, 1)
    roof_sort = {"kernel": "lbvh_sort_pass", "bound": "hbm", "achieved": N_TRIS * 16 / (sw_avg * 1e-3) / 1e9,
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": TRAFFIC_BYTES["lbvh_sort_pass"],
                 "avg_launch_ms": sw_avg,
                 "bytes_per_launch": N_TRIS * 16}
    roof_sort["frac"] = roof_sort["achieved"] / HBM_PEAK_GBS
    roof_build = {"kernel": "lbvh build (all build kernels)", "bound": "hbm",
                  "achieved": N_TRIS * LBVH_BYTES_PER_PRIM / (build_ms * 1e-3) / 1e9,
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "bytes_per_prim": LBVH_BYTES_PER_PRIM,
                  "traffic": TRAFFIC_BYTES["build_per_prim"] * N_TRIS}
    roof_build["frac"] = roof_build["achieved"] / HBM_PEAK_GBS
    out_ms_tot, out_n = ms_of(ungated_kernels, "mlp_infer_fused")
    out_avg = out_ms_tot / max(out_n, 1)
    out_rays = n_rays * ungated_runs / max(out_n, 1)  # rays per launch
    roof_mlp = {"kernel": "mlp_infer_fused (encoding, hidden layers, output layer + argmax)", "bound": "mfma",
                "achieved": out_rays * mlp_flop_per_ray(classes) / (out_avg * 1e-3) / 1e12,
                "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "avg_launch_ms": out_avg,
                "flop_per_ray": mlp_flop_per_ray(classes), "classes": classes,
                "rays_per_launch": out_rays,
                "traffic": TRAFFIC_BYTES["mlp_infer_fused"] * out_rays / 8388608}
    roof_mlp["frac"] = roof_mlp["achieved"] / BF16_PEAK_TFLOPS

    from oracle import binding as orc
    keys, nodes, aabbs = acc.export()
    T = _cpu_threads()
    # Inner nodes visited and triangles tested per ray by the oracle's traversal of this tree,
    # on a seeded sample of the workload's rays (all cores; the counts fix the traversal's
    # algorithmic bytes).
    sel = np.random.default_rng(0).choice(n_rays, args.cpu_rays, replace=False)
    t0 = time.perf_counter()
    ch, counters = orc.trace_bvh(rays[sel], verts, idx, nodes, aabbs, threads=T)
    count_s = time.perf_counter() - t0
    trace_mism = int((ch.view(np.uint32).reshape(-1, 4) != h1[sel].view(np.uint32).reshape(-1, 4)).any(1).sum())
    nodes_per_ray = counters[0] / len(sel)
    tris_per_ray = counters[1] / len(sel)
    # Algorithmic bytes per ray: 32 B ray + 8 B hit + 32 B per inner node visited + 36 B per
    # triangle tested. The layout here fetches 64 B per node (both child boxes) and 48 B per
    # triangle record (the `fetched` figure); HBM traffic is what rocprofv3 measured.
    trace_bytes = 40 + 32 * nodes_per_ray + 36 * tris_per_ray
    fetched_bytes = 48 + 64 * nodes_per_ray + 48 * tris_per_ray
    trace_avg = plain_ms
    ach = n_rays * trace_bytes / (trace_avg * 1e-3) / 1e9
    roof_trace = {"kernel": "trace_closest", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES["trace_closest"],
                  "avg_launch_ms": trace_avg, "rays_per_launch": n_rays,
                  "bytes_per_ray": trace_bytes, "nodes_per_ray": nodes_per_ray, "tris_per_ray": tris_per_ray,
                  "fetched_bytes_per_ray": fetched_bytes,
                  "fetched_gbs": n_rays * fetched_bytes / (trace_avg * 1e-3) / 1e9,
                  "dram_gbs": TRAFFIC_BYTES["trace_closest"] / (trace_avg * 1e-3) / 1e9,
                  "dram_frac": TRAFFIC_BYTES["trace_closest"] / (trace_avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                  "l1_hit_rate": TRACE_CACHE_HIT["l1"], "l2_hit_rate": TRACE_CACHE_HIT["l2"],
                  "counts_from": f"oracle traversal of {len(sel)} sampled rays ({count_s:.2f} s, {T} threads)"}
    log(f"trace: {nodes_per_ray:.1f} nodes + {tris_per_ray:.2f} tris per ray -> {trace_bytes:.0f} B/ray, "
        f"frac {roof_trace['frac']:.3f}; {trace_mism} sampled hits differ from the oracle")
    cpu = None
    if not args.no_cpu:
        m1 = min(args.cpu_build_tris, N_TRIS)
        single_build = _median_time(lambda: orc.lbvh_build(verts[: 3 * m1], idx[:m1]))
        multi_build = _median_time(lambda: orc.par_lbvh_build(verts, idx, threads=T))
        r1 = rays[sel[: args.cpu_rays_single]]
        single_trace = _median_time(lambda: orc.trace_bvh(r1, verts, idx, nodes, aabbs, threads=1))
        rm = rays[sel]
        multi_trace = _median_time(lambda: orc.trace_bvh(rm, verts, idx, nodes, aabbs, threads=T))
        cpu = _cpu_baseline(
            N_TRIS / multi_build / 1e6, "Mprims/s", T,
            f"oracle OpenMP LBVH build of the whole {N_TRIS}-triangle scene ({T} threads); OpenMP traversal "
            f"of {len(rm)} sampled workload rays ({len(rm) / multi_trace / 1e6:.3f} Mrays/s)",
            m1 / single_build / 1e6,
            f"oracle LBVH build of the first {m1} triangles, 1 thread; traversal of {len(r1)} sampled rays "
            f"({len(r1) / single_trace / 1e6:.4f} Mrays/s)",
            trace_mrays_s=len(rm) / multi_trace / 1e6, trace_mrays_s_single=len(r1) / single_trace / 1e6,
            trace_mismatches_vs_gpu=trace_mism)
        log(f"cpu: build {cpu['value']:.2f} Mprims/s ({T} threads), {cpu['single_thread']['value']:.3f} "
            f"(1 thread); trace {cpu['trace_mrays_s']:.3f} / {cpu['trace_mrays_s_single']:.4f} Mrays/s")
    dominant = roof_trace if plain_ms >= out_avg else roof_mlp
    out = {
        "metric": "LBVH build Mprims/s + closest-hit Mrays/s, plain vs learned-proposal, 1 GPU",
        "value": N_TRIS / (build_ms * 1e3),
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (geometry, traversal); bf16 MFMA with f32 accumulate (proposal MLP)",
        "data": "synthetic (seeded blob scene, pinhole + bounce rays; random-init proposal trained on the device)",
        "config": {"workload": "blob scene 4,194,304 tris (64 blobs); 8,388,608 rays (2048^2 primary + 4,194,304 "
                               "bounce); step = LBVH build + plain trace + learned-proposal trace",
                   "triangles": N_TRIS, "rays": n_rays, "treelets": classes, "train_rays": len(tr),
                   "train_epochs": args.epochs},
        "components": {"build_ms": build_ms, "build_mprims_s": N_TRIS / (build_ms * 1e3),
                       "trace_ms_plain": plain_ms, "trace_mrays_s_plain": n_rays / (plain_ms * 1e3),
                       "trace_ms_learned": learned_ms, "trace_mrays_s_learned": n_rays / (learned_ms * 1e3),
                       "learned_infer_ms": infer_ms, "learned_traverse_ms": prop_trace_ms,
                       "learned_serial_sum_ms": infer_ms + prop_trace_ms, "overlap": best_overlap,
                       "overlap_sweep_ms": overlap_sweep,
                       "learned_vs_plain": plain_ms / learned_ms, "treelet_cut_ms": cut_ms,
                       "hit_mismatches_learned_vs_plain": mismatches, "hit_rate": hit_rate,
                       "ungated_trace_ms_learned": ungated_ms, "ungated_infer_ms": ungated_infer_ms,
                       "ungated_traverse_ms": ungated_traverse_ms,
                       "ungated_serial_sum_ms": ungated_infer_ms + ungated_traverse_ms,
                       "ungated_learned_vs_plain": plain_ms / ungated_ms,
                       "hit_mismatches_ungated_vs_plain": mismatches_ungated,
                       "train_s": train_s, "train_loss": loss, "train_top1": train_acc,
                       "gate": gate, "gate_keep_fraction": gate_keep,
                       "proposal_fetches": fetch_split},
        # The roofline of the kernel that takes the most time in the step.
        "roofline": dominant,
        "roofline_build": roof_build,
        "roofline_trace": roof_trace,
        "roofline_sort_pass": roof_sort,
        "roofline_mlp": roof_mlp,
        "cpu_baseline": cpu,
        "kernels": per_kernel,
    }
    print(json.dumps(out))
    for b in (rb, hb, hb2):
        b.destroy()
    prop.destroy()
    acc.destroy()
    ctx.close()


def batched_workload(args):
    from gprt_amd import api, scenes

    lib = api.load()
    ctx = api.Context(0)
    sizes = scenes.batch_sizes(args.meshes, seed=3)
    total = int(sizes.sum())
    t0 = time.time()
    accs, vbs, ibs = [], [], []
    for i, n in enumerate(sizes):
        vb, ib = ctx.device_buffer(nbytes=36 * int(n)), ctx.device_buffer(nbytes=12 * int(n))
        one = np.array([n], dtype=np.int64)
        lib.gprtSynthBlobMeshes(ctx.h, one.ctypes.data_as(C.c_void_p), 1, 1000 + i, 0.002, vb.h, ib.h)
        accs.append(api.TriangleAccel.from_buffers(ctx, vb, ib, 3 * int(n), int(n)))
        vbs.append(vb)
        ibs.append(ib)
    ctx.sync()
    log(f"{len(accs)} meshes, {total} triangles generated on the device in {time.time() - t0:.1f}s")
    ev = (C.c_void_p * len(accs))(*[a.h for a in accs])
    params = api.GPRTBuildParams()

    def timed(fn, reps):
        fn()
        ctx.sync()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ctx.sync()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts)) * 1e3

    def per_mesh():
        for a in accs:
            lib.gprtAccelBuild(ctx.h, a.h, params)

    def batch():
        lib.gprtAccelBuildBatch(ctx.h, ev, len(accs), params)

    per_mesh_ms = timed(per_mesh, args.steps) if not args.no_per_mesh else None
    lib.gprtSetSortConfig(ctx.h, 8, 0, 0)  # the default configuration
    batch_ms = timed(batch, args.steps)
    if args.no_autotune:
        tune, best, tuned_ms = {"best": None, "configs": []}, -1, batch_ms
    else:
        rep = C.create_string_buffer(1 << 14)
        best = lib.gprtAccelBuildBatchAutotune(ctx.h, ev, len(accs), 3, rep, len(rep))
        tune = json.loads(rep.value.decode())
        tuned_ms = timed(batch, args.steps)
    # device-event time of one graph replay (the batch build alone)
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    for _ in range(args.steps):
        batch()
    ctx.sync()
    g_ms, g_n = ms_of(kernel_report(lib, ctx), "lbvh_batch_graph")
    lib.gprtKernelTimerEnable(ctx.h, False)
    graph_ms = g_ms / max(g_n, 1)
    ach = total * LBVH_BYTES_PER_PRIM / (graph_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import binding as orc
        T = _cpu_threads()

        def host_meshes(sel):
            v = np.concatenate([vbs[i].read(np.float32) for i in sel])
            ix = np.concatenate([ibs[i].read(np.uint32) for i in sel])
            return v, ix, sizes[sel]

        # All cores: every 8th mesh, one oracle build per mesh spread over the threads; one thread:
        # every 128th mesh, one after the other.
        vm, im, sm = host_meshes(np.arange(0, len(sizes), 8))
        multi = _median_time(lambda: orc.lbvh_build_many(vm, im, sm, threads=T))
        v1, i1, s1 = host_meshes(np.arange(0, len(sizes), 128))
        single = _median_time(lambda: orc.lbvh_build_many(v1, i1, s1, threads=1))
        cpu = _cpu_baseline(int(sm.sum()) / multi / 1e6, "Mprims/s", T,
                            f"oracle LBVH builds of every 8th mesh ({len(sm)} meshes, {int(sm.sum())} triangles), "
                            f"meshes over {T} threads", int(s1.sum()) / single / 1e6,
                            f"oracle LBVH builds of every 128th mesh ({len(s1)} meshes, {int(s1.sum())} triangles), "
                            f"1 thread")
        del vm, im, v1, i1
        log(f"cpu: {cpu['value']:.2f} Mprims/s ({T} threads), {cpu['single_thread']['value']:.3f} (1 thread)")
    out = {
        "metric": "batched LBVH build Mprims/s (4,096 uneven meshes, one hipGraph launch), 1 GPU",
        "value": total / (graph_ms * 1e3),
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": tuned_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-generated blob meshes, sizes log-uniform 64..262,144)",
        "config": {"workload": f"{len(accs)} meshes, {total} triangles", "meshes": len(accs), "triangles": total},
        "components": {"per_mesh_launches_ms": per_mesh_ms, "batch_graph_default_ms": batch_ms,
                       "batch_graph_tuned_ms": tuned_ms, "batch_graph_device_ms": graph_ms,
                       "speedup_vs_per_mesh": per_mesh_ms / tuned_ms if per_mesh_ms else None,
                       "autotune_best": tune["best"],
                       "autotune_best_index": int(best)},
        "roofline": {"kernel": "lbvh batch graph (all build kernels)", "bound": "hbm", "achieved": ach,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES["batched_build_per_prim"] * total,
                     "bytes_per_prim": LBVH_BYTES_PER_PRIM},
        "autotune": tune,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    for a in accs:
        a.destroy()
    ctx.close()


def dynamic_workload(args):
    """Dynamic scene (gprt_amd.dynamic.DynamicScene): 64 blob meshes of 65,536 triangles (mesh i:
    seed 5000 + i) under one instance accel. Every frame f: all vertices = base + U(-1e-3, 1e-3)
    (seed 7000 + f), the 64 triangle accels are rebuilt in one graph-captured batch and the
    instance accel on top of them, then 1024^2 pinhole primary rays and 1024^2 bounce rays from
    their hits (2,097,152 rays) are traced plainly and with the learned proposal (classes = the
    instance accel's 4,096 top-level leaf groups). Every 16 frames the proposal is fine-tuned on
    that frame's rays with the plain traversal's labels (all 2,097,152 by default) and its
    confidence gate calibrated on all of them. Frame times
    come from the library's per-kernel HIP events: plain frame = perturb + builds + bounce
    generation + plain traces; learned frame = the same with the learned traces in place of the
    plain ones (gated as calibrated at the last fine-tune; their device wall time between two
    events); the ungated learned traces of the last frame are timed apart;
    the fine-tune (with the calibration) is timed on its own (wall clock)."""
    from gprt_amd import api
    from gprt_amd.dynamic import DynamicScene

    lib = api.load()
    ctx = api.Context(0)
    ds = DynamicScene(ctx, max_batch=args.batch)
    n = ds.n
    log(f"dynamic scene: {ds.meshes} x {ds.per} triangles, {ds.rays_per_frame} rays/frame, "
        f"{ds.classes} proposal classes")
    timing = {"on": False, "tune_s": 0.0}

    def on_tune(before):
        if before:
            lib.gprtKernelTimerEnable(ctx.h, False)
            ctx.sync()
            timing["t0"] = time.perf_counter()
        else:
            ctx.sync()
            timing["tune_s"] = time.perf_counter() - timing["t0"]
            lib.gprtKernelTimerEnable(ctx.h, timing["on"])

    tune_rays = args.tune_rays if args.tune_rays > 0 else None
    for f in range(args.warmup):
        ds.frame(f, tune_rays=tune_rays, epochs=args.tune_epochs, lr=args.lr, on_tune=on_tune)
    ctx.sync()
    frames = args.frames
    plain_ms, learned_ms, build_ms, tplain, tlearn, tune_ms = [], [], [], [], [], []
    mism = 0
    hit_rates = []
    t0 = time.perf_counter()
    gate_keeps = []
    for f in range(frames):
        # The frame of DynamicScene.frame, step by step: the plain part under the kernel timer, the
        # learned traces timed whole between two events (with the gate admitting no ray they launch
        # the plain traversal kernel, so kernel names cannot tell them apart).
        lib.gprtKernelTimerReset(ctx.h)
        lib.gprtKernelTimerEnable(ctx.h, True)
        timing["on"] = True
        ds.rebuild(f)
        ds.trace_plain(f)
        tuned = None
        if f % 16 == 0:
            on_tune(True)
            tuned = ds.fine_tune(tune_rays, args.tune_epochs, args.lr)
            on_tune(False)
            gate_keeps.append(ds.gate_keep)
        ctx.sync()
        lib.gprtKernelTimerEnable(ctx.h, False)
        timing["on"] = False
        pk = kernel_report(lib, ctx)
        lib.gprtBeginProfile(ctx.h)
        ds.trace_learned()
        tl = float(lib.gprtEndProfile(ctx.h))
        common = ms_of(pk, "synth_")[0] + ms_of(pk, "lbvh_")[0] + ms_of(pk, "tlas_")[0] + ms_of(pk, "fill_u32")[0] \
            + ms_of(pk, "scan_")[0]
        tp = ms_of(pk, "trace_closest_instances")[0]
        plain_ms.append(common + tp)
        learned_ms.append(common + tl)
        build_ms.append(ms_of(pk, "lbvh_")[0] + ms_of(pk, "tlas_")[0])
        tplain.append(tp)
        tlearn.append(tl)
        if tuned is not None:
            tune_ms.append(timing["tune_s"] * 1e3)
        if f % 16 == 0 or f == frames - 1:
            mism += ds.mismatches()
            hit_rates.append(ds.proposal_hit_rate())
    wall = time.perf_counter() - t0
    # The last frame's learned traces ungated (every ray walks its proposed leaf group first), for
    # comparison; the calibrated gate is put back afterwards.
    gate = float(lib.gprtProposalGetGate(ds.prop.h))
    lib.gprtProposalSetGate(ctx.h, ds.prop.h, float("-inf"))
    ds.trace_learned()
    lib.gprtBeginProfile(ctx.h)
    ds.trace_learned()
    ungated_ms = float(lib.gprtEndProfile(ctx.h))
    mism += ds.mismatches()
    lib.gprtProposalSetGate(ctx.h, ds.prop.h, gate)
    pct = lambda a, q: float(np.percentile(np.asarray(a), q))
    rays = ds.rays_per_frame
    # Node fetches per ray of the last frame's plain traces (untimed), for the traversal's
    # algorithmic bytes: 32 B ray + 8 B hit + 32 B per inner node visited (top and instance
    # levels); leaf tests are not counted, so this is a lower bound.
    vis = ctx.device_buffer(nbytes=4 * n)
    api.trace_detail(ctx, ds.tlas, ds.pb, ds.hp, n, visits_buf=vis)
    v1 = float(vis.read(np.uint32).mean())
    api.trace_detail(ctx, ds.tlas, ds.bb, ds.hb, n, visits_buf=vis)
    v2 = float(vis.read(np.uint32).mean())
    vis.destroy()
    visits = (v1 + v2) / 2
    launch_ms = pct(tplain, 50) / 2  # two traversal launches per frame (primary, bounce)
    roof = _roof_hbm("trace_closest_instances", n * (40 + 32 * visits), launch_ms, bytes_per_ray=40 + 32 * visits,
                     nodes_per_ray=visits, fetched_bytes_per_ray=48 + 64 * visits,
                     note="issue- and L2-bound like the scene traversal; leaf tests not counted")
    cpu = None
    if not args.no_cpu:
        from oracle import binding as orc
        T = _cpu_threads()
        verts = ds.cur.read(np.float32)
        idx = ds.ibuf.read(np.uint32)
        sizes = ds.sizes
        # The last frame's geometry rebuilt on the CPU (one oracle LBVH per mesh) and a sample of
        # its rays traced through the 64 meshes (identity instances); the frame-equivalent rate is
        # the frame's rays over (rebuild + their traversal at the sample's rate).
        multi_build = _median_time(lambda: orc.lbvh_build_many(verts, idx, sizes, threads=T))
        _, nodes, aabbs = orc.lbvh_build_many(verts, idx, sizes, threads=T)
        sample = np.concatenate([ds.pb.read(api.RAY_DTYPE, n)[::64], ds.bb.read(api.RAY_DTYPE, n)[::64]])
        multi_trace = _median_time(lambda: orc.trace_bvh_multi(sample, verts, idx, sizes, nodes, aabbs, threads=T))
        k1 = 8
        single_build = _median_time(lambda: orc.lbvh_build_many(verts[: 9 * k1 * ds.per], idx[: 3 * k1 * ds.per],
                                                                 sizes[:k1], threads=1), reps=3)
        s1 = sample[:: 16]
        single_trace = _median_time(lambda: orc.trace_bvh_multi(s1, verts, idx, sizes, nodes, aabbs, threads=1),
                                    reps=3)
        frame_multi = multi_build + multi_trace / len(sample) * rays
        frame_single = single_build * ds.meshes / k1 + single_trace / len(s1) * rays
        cpu = _cpu_baseline(rays / frame_multi / 1e6, "Mrays/s (whole frame)", T,
                            f"oracle rebuild of the 64 meshes (meshes over {T} threads) + OpenMP traversal of "
                            f"{len(sample)} of the frame's rays through them, scaled to {rays} rays",
                            rays / frame_single / 1e6,
                            f"oracle rebuild of 8 of the 64 meshes and traversal of {len(s1)} rays, 1 thread, "
                            f"scaled to the whole frame (median of 3)",
                            frame_ms=frame_multi * 1e3, frame_ms_single=frame_single * 1e3)
        log(f"cpu: frame {cpu['frame_ms']:.0f} ms ({T} threads), {cpu['frame_ms_single']:.0f} ms (1 thread)")
    out = {
        "metric": "dynamic scene frame time (64 x 65,536 tris rebuilt per frame + 2,097,152 rays), plain vs "
                  "learned proposal, 1 GPU",
        "value": rays / (pct(plain_ms, 50) * 1e3),
        "unit": "Mrays/s (whole frame, plain, p50)",
        "n_gpus": 1, "steps": frames, "warmup": args.warmup, "ms_per_step": wall * 1e3 / frames,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (geometry, traversal); "
        "bf16 MFMA with f32 accumulate (proposal MLP)",
        "data": "synthetic (device-generated blob meshes, seeded per-frame perturbation)",
        "config": {"workload": f"{ds.meshes} meshes x {ds.per} triangles, {rays} rays/frame (1024^2 primary + "
                               f"1024^2 bounce), {frames} frames, fine-tune every 16 frames on "
                               f"{tune_rays or rays} labelled rays x {args.tune_epochs} epochs",
                   "meshes": ds.meshes, "triangles_per_mesh": ds.per, "rays_per_frame": rays, "frames": frames,
                   "proposal_classes": ds.classes},
        "components": {
            "frame_ms_plain_p50": pct(plain_ms, 50), "frame_ms_plain_p99": pct(plain_ms, 99),
            "frame_ms_learned_p50": pct(learned_ms, 50), "frame_ms_learned_p99": pct(learned_ms, 99),
            "build_ms_p50": pct(build_ms, 50), "build_mprims_s_p50": ds.meshes * ds.per / (pct(build_ms, 50) * 1e3),
            "trace_ms_plain_p50": pct(tplain, 50), "trace_mrays_s_plain_p50": rays / (pct(tplain, 50) * 1e3),
            "trace_ms_learned_p50": pct(tlearn, 50), "trace_mrays_s_learned_p50": rays / (pct(tlearn, 50) * 1e3),
            "learned_vs_plain_trace": pct(tplain, 50) / pct(tlearn, 50),
            "gate_keep_by_tune": gate_keeps,
            "ungated_trace_ms_learned": ungated_ms,
            "ungated_learned_vs_plain_trace": pct(tplain, 50) / ungated_ms,
            "fine_tune_wall_ms_mean": float(np.mean(tune_ms)) if tune_ms else None,
            "proposal_hit_rate_by_check": hit_rates, "hit_mismatches_learned_vs_plain": mism},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ds.destroy()
    ctx.close()


def copy_workload(args):
    """Calibration for HBM byte counters: a plain device-to-device copy of 1 GiB (torch's copy
    kernel), timed with HIP events; under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE its counters
    give the factor between counted and moved bytes."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    print(json.dumps({"metric": "device copy GB/s (read + write)", "value": 2 * 4 * n / (ms * 1e-3) / 1e9,
                      "unit": "GB/s", "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n,
                      "ms_per_copy": ms, "steps": args.steps}))


REFIT_BYTES_PER_PRIM = 104  # 48 B triangle read + 8 B parent link + 48 B node box write
BOX_LBVH_BYTES_PER_PRIM = 296  # the 368 B/prim build with a 24-byte box in place of each 48-byte triangle read
N_BOXES = 4_194_304
N_SPHERES = 16_777_216
BOUNDS_BYTES_PER_PRIM = 12 + 4 + 24  # centre + radius read, box written


def _roof_hbm(kernel, bytes_per_launch, avg_ms, **extra):
    ach = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"kernel": kernel, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": avg_ms, "bytes_per_launch": bytes_per_launch,
            **extra}


def refit_workload(args):
    """Refit-only update (gprtAccelUpdate): the 4,194,304-triangle blob scene is built once with
    GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE; a step moves every vertex (base + U(-1e-3, 1e-3), device
    generated, seed 7000 + step) and updates the accel: node boxes recomputed over the kept
    topology. The 2048^2 primary rays are traced after each update and after a full rebuild of the
    same vertices, to show what the kept topology costs the trace."""
    from gprt_amd import api, scenes

    lib = api.load()
    verts, idx = scenes.blob_scene(N_TRIS, seed=2)
    prim = scenes.pinhole_rays(IMG, IMG)
    ctx = api.Context(0)
    base = ctx.device_buffer(verts)
    acc = api.TriangleAccel(ctx, verts, idx)
    acc.build(api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
    rb = ctx.device_buffer(prim)
    hb = ctx.device_buffer(nbytes=16 * len(prim))
    floats = verts.size

    def step(k):
        lib.gprtSynthPerturb(ctx.h, base.h, acc.vbuf.h, floats, 7000 + k, 1e-3)
        acc.update()
        acc.trace(rb, hb, len(prim))

    for k in range(args.warmup):
        step(k)
    ctx.sync()
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k)
    ctx.sync()
    wall = time.perf_counter() - t0
    lib.gprtKernelTimerEnable(ctx.h, False)
    K = args.steps
    pk = kernel_report(lib, ctx)
    refit_ms = (ms_of(pk, "lbvh_build")[0] + ms_of(pk, "lbvh_ident")[0]) / K
    trace_refit_ms = ms_of(pk, "trace_closest")[0] / K
    # The same vertices rebuilt from scratch: build and trace times for comparison.
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    for _ in range(K):
        acc.build(api.GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE)
        acc.trace(rb, hb, len(prim))
    ctx.sync()
    lib.gprtKernelTimerEnable(ctx.h, False)
    pk2 = kernel_report(lib, ctx)
    build_ms = ms_of(pk2, "lbvh_")[0] / K
    trace_rebuilt_ms = ms_of(pk2, "trace_closest")[0] / K
    roof = _roof_hbm("lbvh_ident + build_local + region + top (refit over the kept topology)",
                     N_TRIS * REFIT_BYTES_PER_PRIM,
                     refit_ms, bytes_per_prim=REFIT_BYTES_PER_PRIM)
    roof["traffic"] = TRAFFIC_BYTES["refit_per_prim"] * N_TRIS
    cpu = None
    if not args.no_cpu:
        from oracle import binding as orc
        T = _cpu_threads()
        moved = acc.vbuf.read(np.float32).reshape(-1, 3)
        _, nodes, _ = acc.export()
        single = _median_time(lambda: orc.refit(moved, idx, nodes))
        multi = _median_time(lambda: orc.par_refit(moved, idx, nodes, threads=T))
        cpu = _cpu_baseline(N_TRIS / multi / 1e6, "Mprims/s", T,
                            f"oracle atomic-flag refit of the whole {N_TRIS}-triangle scene over the same topology, "
                            f"{T} threads", N_TRIS / single / 1e6,
                            f"oracle refit (every leaf into every ancestor) of the whole scene, 1 thread")
    out = {
        "metric": "refit-only update Mprims/s (gprtAccelUpdate), 1 GPU", "value": N_TRIS / (refit_ms * 1e3),
        "unit": "Mprims/s", "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded blob scene, device-generated per-step vertex motion)",
        "config": {"workload": "blob scene 4,194,304 tris, every vertex moved by U(+-1e-3) per step; step = "
                               "perturb + refit + trace of 2048^2 primary rays", "triangles": N_TRIS,
                   "rays": len(prim)},
        "components": {"refit_ms": refit_ms, "rebuild_ms": build_ms, "refit_vs_rebuild": build_ms / refit_ms,
                       "trace_ms_after_refit": trace_refit_ms, "trace_ms_after_rebuild": trace_rebuilt_ms},
        "roofline": roof, "cpu_baseline": cpu, "kernels": pk,
    }
    print(json.dumps(out))
    for b in (base, rb, hb):
        b.destroy()
    acc.destroy()
    ctx.close()


def _box_scene(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    h = (rng.random((n, 3), dtype=np.float32) * np.float32(0.002) + np.float32(5e-4)).astype(np.float32)
    return np.concatenate([c - h, c + h], 1).astype(np.float32)


def _box_rays(n, seed):
    from gprt_amd import scenes
    rng = np.random.default_rng(seed)
    o = rng.random((n, 3), dtype=np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    r = np.zeros(n, dtype=scenes.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = o[:, 0], o[:, 1], o[:, 2]
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["tmin"], r["tmax"] = 0.0, 1e30
    return r


def boxes_workload(args):
    """AABB-primitive accel (gprtAABBsSetPositions + gprtAABBAccelCreate): 4,194,304 boxes (centres
    U[0,1]^3, half extents U[5e-4, 2.5e-3] per axis, seed 11); a step is one LBVH build over the
    boxes and a closest-hit trace of 4,194,304 rays (origins U[0,1]^3, Gaussian directions, seed 12)."""
    from gprt_amd import api

    lib = api.load()
    boxes = _box_scene(N_BOXES, 11)
    rays = _box_rays(N_BOXES, 12)
    ctx = api.Context(0)
    acc = api.AABBAccel(ctx, boxes)
    rb = ctx.device_buffer(rays)
    hb = ctx.device_buffer(nbytes=16 * len(rays))
    vis = ctx.device_buffer(nbytes=4 * len(rays))

    def step():
        acc.build()
        acc.trace(rb, hb, len(rays))

    for _ in range(args.warmup):
        step()
    ctx.sync()
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    wall = time.perf_counter() - t0
    lib.gprtKernelTimerEnable(ctx.h, False)
    K = args.steps
    pk = kernel_report(lib, ctx)
    build_ms = ms_of(pk, "lbvh_")[0] / K
    trace_ms = ms_of(pk, "trace_closest")[0] / K
    api.trace_detail(ctx, acc, rb, hb, len(rays), visits_buf=vis)
    visits = float(vis.read(np.uint32).mean())
    hits = hb.read(api.HIT_DTYPE)
    roof = _roof_hbm("LBVH build over boxes (all build kernels)", N_BOXES * BOX_LBVH_BYTES_PER_PRIM, build_ms,
                     bytes_per_prim=BOX_LBVH_BYTES_PER_PRIM)
    roof["traffic"] = TRAFFIC_BYTES["boxes_build_per_prim"] * N_BOXES
    cpu = None
    if not args.no_cpu:
        from oracle import binding as orc
        T = _cpu_threads()
        m = min(args.cpu_build_tris, N_BOXES)
        single = _median_time(lambda: orc.lbvh_build_boxes(boxes[:m]))
        multi = _median_time(lambda: orc.par_lbvh_build_boxes(boxes, threads=T))
        sel = np.random.default_rng(0).choice(len(rays), 256, replace=False)
        sub = boxes[: 1 << 16]
        bf_s = _median_time(lambda: orc.trace_boxes_bruteforce(rays[sel], sub, threads=T), reps=3)
        cpu = _cpu_baseline(
            N_BOXES / multi / 1e6, "Mprims/s", T,
            f"oracle OpenMP LBVH build of all {N_BOXES} boxes ({T} threads); OpenMP brute-force trace of 256 rays "
            f"against the first 65,536 boxes ({256 * 65536 / bf_s / 1e6:.1f} M ray-box tests/s)",
            m / single / 1e6, f"oracle LBVH build of the first {m} boxes, 1 thread")
    out = {
        "metric": "AABB-primitive LBVH build Mprims/s + closest-hit Mrays/s, 1 GPU",
        "value": N_BOXES / (build_ms * 1e3), "unit": "Mprims/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded random boxes and rays)",
        "config": {"workload": "4,194,304 boxes + 4,194,304 random rays; step = LBVH build + closest-hit trace",
                   "boxes": N_BOXES, "rays": len(rays)},
        "components": {"build_ms": build_ms, "trace_ms": trace_ms, "trace_mrays_s": len(rays) / (trace_ms * 1e3),
                       "trace_hbm_bytes_per_launch": TRAFFIC_BYTES["trace_boxes"],
                       "trace_hbm_gbs": TRAFFIC_BYTES["trace_boxes"] / (trace_ms * 1e-3) / 1e9,
                       "node_visits_per_ray": visits, "hit_rate": float((hits["primID"] >= 0).mean())},
        "roofline": roof, "cpu_baseline": cpu, "kernels": pk,
    }
    print(json.dumps(out))
    for b in (rb, hb, vis):
        b.destroy()
    acc.destroy()
    ctx.close()


def compute_workload(args):
    """User compute programs through gprtComputeLaunch: the reference sample's ComputeBounds
    (sphere centre + radius -> box) over 16,777,216 spheres, timed with the library's per-launch
    HIP events, and the host cost of a launch (back-to-back one-group launches, wall clock)."""
    from gprt_amd import api

    lib = api.load()
    code = os.path.join(ROOT, "tests", "programs", "compute_programs.co")
    n = N_SPHERES
    rng = np.random.default_rng(13)
    centres = rng.random((n, 3), dtype=np.float32)
    radii = (rng.random(n, dtype=np.float32) * np.float32(0.01)).astype(np.float32)
    ctx = api.Context(0)
    module = api.Module(ctx, code)
    prog = module.compute("ComputeBounds")
    vb, rbuf = ctx.device_buffer(centres), ctx.device_buffer(radii)
    ab = ctx.device_buffer(nbytes=24 * n)

    class Rec(C.Structure):
        _fields_ = [("vertex", C.c_void_p), ("radius", C.c_void_p), ("aabbs", C.c_void_p)]

    dp = lambda b: lib.gprtBufferGetDevicePointer(b.h, 0)
    rec = Rec(dp(vb), dp(rbuf), dp(ab))
    groups = ((n + 1023) // 1024,)  # at most 65,535 groups per dimension
    for _ in range(args.warmup):
        prog.launch(groups, (1024,), rec, C.c_uint32(n))
    ctx.sync()
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prog.launch(groups, (1024,), rec, C.c_uint32(n))
    ctx.sync()
    wall = time.perf_counter() - t0
    lib.gprtKernelTimerEnable(ctx.h, False)
    K = args.steps
    launches = C.c_uint32(0)
    ms = lib.gprtKernelTimerQuery(ctx.h, b"ComputeBounds", C.byref(launches)) / max(launches.value, 1)
    got = ab.read(np.float32).reshape(-1, 6)
    want_lo = centres - radii[:, None]
    ok = bool(np.array_equal(got[:, :3], want_lo) and np.array_equal(got[:, 3:], centres + radii[:, None]))
    reps = 2000
    one = (1,)
    small = Rec(dp(vb), dp(rbuf), dp(ab))
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        prog.launch(one, (64,), small, C.c_uint32(64))
    host_us = (time.perf_counter() - t0) / reps * 1e6
    ctx.sync()
    roof = _roof_hbm("ComputeBounds (user program via gprtComputeLaunch)", n * BOUNDS_BYTES_PER_PRIM, ms,
                     bytes_per_prim=BOUNDS_BYTES_PER_PRIM)
    roof["traffic"] = TRAFFIC_BYTES["compute_bounds"]
    cpu = None
    if not args.no_cpu:
        from oracle import binding as orc
        T = _cpu_threads()
        m = 1 << 22
        single = _median_time(lambda: orc.par_sphere_bounds(centres[:m], radii[:m], threads=1))
        multi = _median_time(lambda: orc.par_sphere_bounds(centres, radii, threads=T))
        cpu = _cpu_baseline(n / multi / 1e6, "Mprims/s", T,
                            f"oracle OpenMP sphere bounds of all {n} spheres ({T} threads)",
                            m / single / 1e6, f"oracle sphere bounds of the first {m} spheres, 1 thread")
    out = {
        "metric": "user compute program (ComputeBounds) Mprims/s through gprtComputeLaunch, 1 GPU",
        "value": n / (ms * 1e3), "unit": "Mprims/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded sphere centres and radii)",
        "config": {"workload": "16,777,216 spheres -> boxes, 1024-thread groups", "spheres": n},
        "components": {"kernel_ms": ms, "host_us_per_launch": host_us, "boxes_match_numpy": ok},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    for b in (vb, rbuf, ab):
        b.destroy()
    prog.destroy()
    module.destroy()
    ctx.close()


SORT_KEYS = 1 << 26  # u32 keys uniform over 32 bits + u32 payload = index
SORT_BYTES_PER_KEY = 4 + 4 * 2 * (4 + 4)  # histogram read + 4 passes x (read + write) x (key + payload)
SORT_PASS_BYTES_PER_KEY = 2 * (4 + 4)
SCAN_ELEMS = 1 << 28  # u32 uniform in [0, 255]
SCAN_BYTES_PER_ELEM = 4 + 4


SORT_KERNELS = ("fill_u32", "sort_hist", "sort_hist_scan", "sort_onesweep", "sort_final_copy")  # fill: look-back state


def _device_ms(per_kernel, names):
    return sum(per_kernel[k]["total_ms"] for k in names if k in per_kernel)


def sort_workload(args):
    """The radix-sort microbenchmark: gprtBufferSortKV over 2^26 u32 keys uniform over 32 bits
    with a u32 payload (its index). Every step restores the unsorted input with a device copy
    first (outside the sort's kernels), so each sort sees the same random keys; the time is the
    sum of the sort's own kernels from the library's per-launch HIP events."""
    from gprt_amd import api, scenes

    lib = api.load()
    n = SORT_KEYS
    keys = (scenes.rand_u64(4, 0, n) >> np.uint64(32)).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    ctx = api.Context(0)
    k0, v0 = ctx.device_buffer(keys), ctx.device_buffer(vals)
    kb, vb = ctx.device_buffer(nbytes=4 * n), ctx.device_buffer(nbytes=4 * n)
    scratch = ctx.device_buffer(nbytes=16 * n)  # ample; the sort takes what it needs

    def one():
        lib.gprtBufferCopy(ctx.h, k0.h, kb.h, 0, 0, 4, n, 0, 0)
        lib.gprtBufferCopy(ctx.h, v0.h, vb.h, 0, 0, 4, n, 0, 0)
        lib.gprtBufferSortKV(ctx.h, kb.h, vb.h, n, 32, 32, scratch.h)

    for _ in range(args.warmup):
        one()
    ctx.sync()
    assert api.last_error() is None, api.last_error()
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one()
    ctx.sync()
    wall = time.perf_counter() - t0
    lib.gprtKernelTimerEnable(ctx.h, False)
    pk = kernel_report(lib, ctx)
    K = args.steps
    sort_ms = _device_ms(pk, SORT_KERNELS) / K
    passes = pk.get("sort_onesweep", {"launches": 0})["launches"] // max(K, 1)
    pass_ms = pk["sort_onesweep"]["total_ms"] / pk["sort_onesweep"]["launches"]
    # The same sort under each tile size of the sort configuration (8-bit digits), for comparison.
    sweep = {}
    for tile in (2048, 4096, 8192, 16384):
        lib.gprtSetSortConfig(ctx.h, 8, tile, 0)
        one()
        ctx.sync()
        lib.gprtKernelTimerReset(ctx.h)
        lib.gprtKernelTimerEnable(ctx.h, True)
        for _ in range(5):
            one()
        ctx.sync()
        lib.gprtKernelTimerEnable(ctx.h, False)
        sweep[tile] = _device_ms(kernel_report(lib, ctx), SORT_KERNELS) / 5
    lib.gprtSetSortConfig(ctx.h, 8, 0, 0)
    got_k, got_v = kb.read(np.uint32), vb.read(np.uint32)
    order = np.argsort(keys, kind="stable")
    ok = bool(np.array_equal(got_k, keys[order]) and np.array_equal(got_v, order.astype(np.uint32)))
    roof = _roof_hbm("sort_onesweep (one 8-bit pass over key + payload)", n * SORT_PASS_BYTES_PER_KEY, pass_ms,
                     bytes_per_key=SORT_PASS_BYTES_PER_KEY)
    roof["traffic"] = TRAFFIC_BYTES.get("sort_onesweep")
    whole = SORT_BYTES_PER_KEY * n / (sort_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import binding as orc  # the CPU restatement, timed as the baseline only
        T = _cpu_threads()
        L = orc.lib()
        kw, vw = np.empty_like(keys), np.empty_like(vals)

        def restore(m):
            np.copyto(kw[:m], keys[:m])
            np.copyto(vw[:m], vals[:m])

        m1 = 1 << 22
        single = _median_time(lambda: L.orc_stable_sort(orc._p(kw), orc._p(vw), m1, 32, 32),
                              setup=lambda: restore(m1))
        multi = _median_time(lambda: L.orc_par_radix_sort(orc._p(kw), orc._p(vw), n, 32, 32, T),
                             setup=lambda: restore(n))
        cpu = _cpu_baseline(n / multi / 1e6, "Mkeys/s", T,
                            f"oracle OpenMP LSD radix sort (8-bit digits) of all {n} key/payload pairs, {T} threads",
                            m1 / single / 1e6,
                            f"oracle std::stable_sort of the first {m1} key/payload pairs (u32 + u32), 1 thread")
    out = {
        "metric": "radix sort Mkeys/s (2^26 u32 keys + u32 payload, gprtBufferSortKV), 1 GPU",
        "value": n / (sort_ms * 1e3), "unit": "Mkeys/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (SplitMix64 keys, seed 4; payload = index)",
        "config": {"workload": "sort microbenchmark: 67,108,864 u32 keys + u32 payloads", "keys": n},
        "components": {"sort_device_ms": sort_ms, "passes": passes, "pass_ms": pass_ms,
                       "whole_sort_gbs": whole, "whole_sort_frac": whole / HBM_PEAK_GBS,
                       "whole_sort_bytes_per_key": SORT_BYTES_PER_KEY, "sorted_matches_numpy": ok,
                       "sort_ms_by_tile_keys": sweep,
                       "wall_ms_per_step_includes": "two 256 MiB device copies restoring the input"},
        "roofline": roof, "cpu_baseline": cpu, "kernels": pk,
    }
    print(json.dumps(out))
    for b in (k0, v0, kb, vb, scratch):
        b.destroy()
    ctx.close()


def scan_workload(args):
    """The exclusive-scan microbenchmark: gprtBufferExclusiveSum over 2^28 u32 uniform in
    [0, 255] (the total wraps modulo 2^32), timed from the library's per-launch HIP events."""
    from gprt_amd import api, scenes

    lib = api.load()
    n = SCAN_ELEMS
    x = (scenes.rand_u64(5, 0, n) >> np.uint64(56)).astype(np.uint32)
    ctx = api.Context(0)
    inp, outb = ctx.device_buffer(x), ctx.device_buffer(nbytes=4 * n)
    scratch = ctx.device_buffer(nbytes=4)
    for _ in range(args.warmup):
        lib.gprtBufferExclusiveSum(ctx.h, inp.h, outb.h, scratch.h)
    ctx.sync()
    assert api.last_error() is None, api.last_error()
    lib.gprtKernelTimerReset(ctx.h)
    lib.gprtKernelTimerEnable(ctx.h, True)
    t0 = time.perf_counter()
    total = 0
    for _ in range(args.steps):
        total = lib.gprtBufferExclusiveSum(ctx.h, inp.h, outb.h, scratch.h)
    ctx.sync()
    wall = time.perf_counter() - t0
    lib.gprtKernelTimerEnable(ctx.h, False)
    pk = kernel_report(lib, ctx)
    K = args.steps
    lb = pk["scan_lookback"]
    ms = lb["total_ms"] / lb["launches"]
    scan_ms = sum(v["total_ms"] for k, v in pk.items() if k.startswith("scan")) / K
    got = outb.read(np.uint32)
    c = np.cumsum(x, dtype=np.uint64)
    ok = bool(int(total) == int(c[-1]) & 0xFFFFFFFF and got[0] == 0 and
              np.array_equal(got[1:], (c[:-1] & np.uint64(0xFFFFFFFF)).astype(np.uint32)))
    del c
    roof = _roof_hbm("scan_lookback (decoupled look-back exclusive sum)", n * SCAN_BYTES_PER_ELEM, ms,
                     bytes_per_elem=SCAN_BYTES_PER_ELEM)
    roof["traffic"] = TRAFFIC_BYTES.get("scan_lookback")
    cpu = None
    if not args.no_cpu:
        from oracle import binding as orc  # the CPU restatement, timed as the baseline only
        T = _cpu_threads()
        L = orc.lib()
        out = np.empty_like(x)
        single = _median_time(lambda: L.orc_exclusive_sum(orc._p(x), orc._p(out), n))
        multi = _median_time(lambda: L.orc_par_exclusive_sum(orc._p(x), orc._p(out), n, T))
        cpu = _cpu_baseline(n / multi / 1e9, "Gelems/s", T,
                            f"oracle two-pass blocked exclusive scan of all {n} elements, {T} threads",
                            n / single / 1e9, f"oracle std::exclusive_scan of all {n} elements, 1 thread")
    out = {
        "metric": "exclusive scan Gelems/s (2^28 u32, gprtBufferExclusiveSum), 1 GPU",
        "value": n / (scan_ms * 1e-3) / 1e9, "unit": "Gelems/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (SplitMix64 values in [0, 255], seed 5)",
        "config": {"workload": "scan microbenchmark: 268,435,456 u32", "elements": n},
        "components": {"scan_device_ms": scan_ms, "lookback_ms": ms, "total_matches_numpy": ok},
        "roofline": roof, "cpu_baseline": cpu, "kernels": pk,
    }
    print(json.dumps(out))
    for b in (inp, outb, scratch):
        b.destroy()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="scene",
                    choices=["scene", "batched", "dynamic", "copy", "refit", "boxes", "compute", "sort", "scan"])
    ap.add_argument("--frames", type=int, default=256, help="dynamic: timed frames")
    ap.add_argument("--tune-rays", type=int, default=0, help="dynamic: labelled rays per fine-tune (0: the whole frame)")
    ap.add_argument("--tune-epochs", type=int, default=2, help="dynamic: epochs per fine-tune")
    ap.add_argument("--cpu-rays", type=int, default=262144, help="rays of the all-cores CPU traversal sample")
    ap.add_argument("--cpu-rays-single", type=int, default=32768, help="rays of the 1-thread CPU traversal sample")
    ap.add_argument("--cpu-build-tris", type=int, default=1 << 20, help="triangles of the 1-thread CPU build sample")
    ap.add_argument("--train-rays", type=int, default=16 << 20, help="proposal training rays (16,777,216: C4)")
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--gate-rays", type=int, default=262144, help="training rays the proposal gate is calibrated on")
    ap.add_argument("--overlap-chunks", type=int, nargs="*", default=[4, 8, 16],
                    help="scene: chunk counts of the overlapped learned trace to sweep")
    ap.add_argument("--overlap-blocks", type=int, nargs="*", default=[6, 8, 10],
                    help="scene: traversal workgroups per CU of the overlapped learned trace to sweep")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--meshes", type=int, default=4096)
    ap.add_argument("--no-autotune", action="store_true", help="batched: skip the autotuner sweep")
    ap.add_argument("--no-per-mesh", action="store_true", help="batched: skip the per-mesh launch loop")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} requested; this build measures one GPU only "
              f"(the path has no multi-GPU form here). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        batched_workload(args)
    elif args.workload == "copy":
        copy_workload(args)
    elif args.workload == "dynamic":
        dynamic_workload(args)
    elif args.workload == "refit":
        refit_workload(args)
    elif args.workload == "boxes":
        boxes_workload(args)
    elif args.workload == "compute":
        compute_workload(args)
    elif args.workload == "sort":
        sort_workload(args)
    elif args.workload == "scan":
        scan_workload(args)
    else:
        scene_workload(args)


if __name__ == "__main__":
    main()
