"""Benchmark: LBVH build Mprims/s + closest-hit Mrays/s on one MI355X.

Workload (the large single scene): 4,194,304 triangles from 64 Gaussian blobs (FP32), 8,388,608
rays = 2048x2048 pinhole primary rays + 4,194,304 cosine-weighted bounce rays from the primary
hits. A step is one full LBVH build (bounds, Morton keys, key sort, Karras, refit) followed by a
closest-hit trace of all 8,388,608 rays, with every input resident in HBM before timing starts.

`value` is the build throughput (primitives / build time). Build and trace times come from HIP
events the library records around every kernel on its stream during the timed steps; the whole
timed region is also wall-clocked between two stream synchronisations. The roofline entries use
algorithmic bytes per launch (fixed by the workload, not by tuning knobs):
  * sort pass (u32 Morton key + u32 primID payload): read 8 B + write 8 B per key;
  * trace: 32 B ray + 16 B hit + 64 B per inner node fetched + 48 B per triangle tested, with
    the node / triangle counts measured by the CPU reference traversal on a ray sample.
The CPU baseline times the oracle (single thread) on a bounded sample and is reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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
N_TRIS = 4_194_304
IMG = 2048
N_BOUNCE = 4_194_304


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_ms(lib, ctx, name):
    n = C.c_uint32(0)
    ms = lib.gprtKernelTimerQuery(ctx.h, name.encode(), C.byref(n))
    return float(ms), int(n.value)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-rays", type=int, default=16384)
    ap.add_argument("--cpu-build-tris", type=int, default=1 << 20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} requested; this build measures one GPU only "
              f"(the path has no multi-GPU form here). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from gprt_amd import api, scenes

    lib = api.load()
    t0 = time.time()
    verts, idx = scenes.blob_scene(N_TRIS, seed=2)
    prim = scenes.pinhole_rays(IMG, IMG)
    log(f"scene + primary rays generated in {time.time() - t0:.1f}s")

    ctx = api.Context(0)
    acc = api.TriangleAccel(ctx, verts, idx)
    acc.build()
    pb = ctx.device_buffer(prim)
    phb = ctx.device_buffer(nbytes=len(prim) * 16)
    acc.trace(pb, phb, len(prim))
    ph = phb.read(api.HIT_DTYPE)
    t0 = time.time()
    bounce = scenes.bounce_rays(prim, ph, verts, idx, N_BOUNCE, seed=3)
    rays = np.concatenate([prim, bounce])
    log(f"bounce rays generated in {time.time() - t0:.1f}s; primary hit rate {(ph['primID'] >= 0).mean():.3f}")
    pb.destroy()
    phb.destroy()
    n_rays = len(rays)
    rb = ctx.device_buffer(rays)
    hb = ctx.device_buffer(nbytes=n_rays * 16)

    def step():
        acc.build()
        acc.trace(rb, hb, n_rays)

    for _ in range(args.warmup):
        step()
    ctx.sync()

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
    build_names = ["lbvh_"]
    trace_ms_tot, trace_launches = kernel_ms(lib, ctx, "trace_closest")
    build_ms_tot = 0.0
    buf = C.create_string_buffer(1 << 16)
    lib.gprtKernelTimerReport(ctx.h, buf, len(buf))
    per_kernel = {}
    for line in buf.value.decode().strip().splitlines():
        name, n, ms = line.split(",")
        per_kernel[name] = {"launches": int(n), "total_ms": float(ms)}
    for name, rec in per_kernel.items():
        if any(name.startswith(p) for p in build_names):
            build_ms_tot += rec["total_ms"]
    build_ms = build_ms_tot / K
    trace_ms = trace_ms_tot / K
    ms_per_step = wall * 1e3 / K
    build_mprims = N_TRIS / (build_ms * 1e3)
    trace_mrays = n_rays / (trace_ms * 1e3)

    for name, rec in sorted(per_kernel.items(), key=lambda kv: -kv[1]["total_ms"]):
        log(f"  {name:24s} {rec['launches']:5d} launches {rec['total_ms'] / K:9.3f} ms/step "
            f"{rec['total_ms'] / rec['launches']:9.4f} ms/launch")
    hits = hb.read(api.HIT_DTYPE)
    hit_rate = float((hits["primID"] >= 0).mean())

    # Roofline of the dominant build kernel: one onesweep pass over u32 key + u32 payload.
    sw_ms, sw_n = kernel_ms(lib, ctx, "lbvh_sort_pass")
    sw_avg_ms = sw_ms / max(sw_n, 1)
    sw_bytes = N_TRIS * 16
    roof_sort = {"kernel": "lbvh_sort_pass", "bound": "hbm", "achieved": sw_bytes / (sw_avg_ms * 1e-3) / 1e9,
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
                 "avg_launch_ms": sw_avg_ms, "bytes_per_launch": sw_bytes}
    roof_sort["frac"] = roof_sort["achieved"] / HBM_PEAK_GBS

    cpu = None
    trace_bytes = None
    if not args.no_cpu:
        from oracle import binding as orc
        # Node / triangle visit counts on a ray sample (for the trace's algorithmic bytes) and the
        # single-thread CPU traversal rate on the same BVH.
        keys, nodes, aabbs = acc.export()
        sel = np.random.default_rng(0).choice(n_rays, args.cpu_rays, replace=False)
        t0 = time.perf_counter()
        ch, counters = orc.trace_bvh(rays[sel], verts, idx, nodes, aabbs, threads=1)
        cpu_trace_s = time.perf_counter() - t0
        mism = int((ch["primID"] != hits["primID"][sel]).sum())
        nodes_per_ray = counters[0] / len(sel)
        tris_per_ray = counters[1] / len(sel)
        trace_bytes = 32 + 16 + 64 * nodes_per_ray + 48 * tris_per_ray
        # Single-thread oracle LBVH build on a bounded prefix of the scene.
        m = args.cpu_build_tris
        t0 = time.perf_counter()
        orc.lbvh_build(verts[: 3 * m], idx[:m])
        cpu_build_s = time.perf_counter() - t0
        cpu = {"value": m / cpu_build_s / 1e6, "unit": "Mprims/s", "cores": 1, "kind": "port",
               "sample": f"oracle LBVH build of the first {m} triangles of the scene (single thread); "
                         f"oracle BVH traversal of {len(sel)} sampled rays: "
                         f"{len(sel) / cpu_trace_s / 1e6:.4f} Mrays/s single thread",
               "trace_mrays_s": len(sel) / cpu_trace_s / 1e6,
               "host_cpus": os.cpu_count(), "trace_mismatches_vs_gpu": mism}
        log(f"cpu: build {cpu['value']:.3f} Mprims/s, trace {cpu['trace_mrays_s']:.4f} Mrays/s; "
            f"{nodes_per_ray:.1f} nodes + {tris_per_ray:.1f} tris per ray")

    roof_trace = None
    if trace_bytes is not None:
        avg_ms = trace_ms_tot / max(trace_launches, 1)
        ach = n_rays * trace_bytes / (avg_ms * 1e-3) / 1e9
        roof_trace = {"kernel": "trace_closest", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": avg_ms,
                      "bytes_per_ray": trace_bytes}
    # The dominant kernel of a step is the trace (8.4M rays) unless the build outweighs it.
    dominant = roof_trace if (roof_trace and trace_ms > sw_avg_ms) else roof_sort

    out = {
        "metric": "LBVH build Mprims/s + closest-hit Mrays/s, plain vs learned-proposal, 1 GPU",
        "value": build_mprims,
        "unit": "Mprims/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded blob scene + pinhole/bounce rays)",
        "config": {"workload": "blob scene 4,194,304 tris, 64 blobs; 8,388,608 rays (2048^2 primary + "
                               "4,194,304 bounce); step = LBVH build + closest-hit trace",
                   "triangles": N_TRIS, "rays": n_rays},
        "components": {"build_ms": build_ms, "trace_ms": trace_ms, "build_mprims_s": build_mprims,
                       "trace_mrays_s_plain": trace_mrays, "trace_mrays_s_learned": None,
                       "hit_rate": hit_rate, "wall_ms_per_step": ms_per_step},
        "roofline": dominant,
        "roofline_build": roof_sort,
        "cpu_baseline": cpu,
        "kernels": per_kernel,
    }
    print(json.dumps(out))
    rb.destroy()
    hb.destroy()
    acc.destroy()
    ctx.close()


if __name__ == "__main__":
    main()
