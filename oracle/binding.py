"""ctypes binding of the CPU oracle (oracle/liboracle.so). TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline
leg as the timed CPU path. The product (gprt_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

HIT_DTYPE = np.dtype([("t", "<f4"), ("primID", "<i4"), ("u", "<f4"), ("v", "<f4")])


def build(force: bool = False) -> None:
    args = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sz = C.c_size_t
        sigs = {
            "orc_stable_sort": (None, [P, P, sz, C.c_int, C.c_int]),
            "orc_exclusive_sum": (C.c_uint32, [P, P, sz]),
            "orc_partition": (C.c_uint32, [P, P, sz, C.c_int]),
            "orc_select": (C.c_uint32, [P, P, sz, C.c_int]),
            "orc_separate_bits": (C.c_uint32, [C.c_uint32]),
            "orc_morton3": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_scene_bounds": (None, [P, P, sz, P]),
            "orc_morton_keys": (None, [P, P, sz, P, P]),
            "orc_karras": (None, [P, sz, P]),
            "orc_topdown": (None, [P, sz, P]),
            "orc_refit": (None, [P, P, P, sz, P]),
            "orc_refit_bruteforce": (None, [P, P, P, sz, P]),
            "orc_lbvh_build": (None, [P, P, sz, P, P, P]),
            "orc_lbvh_build_pe": (None, [P, P, sz, P, P, P]),
            "orc_intersect": (C.c_int, [P, P, P, P, P, P, P]),
            "orc_plane_t_fp64": (C.c_double, [P, P, P, P]),
            "orc_trace_bruteforce": (None, [P, sz, P, P, sz, P, C.c_int]),
            "orc_trace_bvh": (None, [P, sz, P, P, P, P, sz, P, P, C.c_int]),
            "orc_trace_bvh_ray": (None, [P, P, P, P, P, sz, P]),
            "orc_num_procs": (C.c_int, []),
            "orc_par_radix_sort": (None, [P, P, sz, C.c_int, C.c_int, C.c_int]),
            "orc_par_exclusive_sum": (C.c_uint32, [P, P, sz, C.c_int]),
            "orc_par_lbvh_build": (None, [P, P, sz, P, P, P, C.c_int]),
            "orc_par_lbvh_build_boxes": (None, [P, sz, P, P, P, C.c_int]),
            "orc_par_sphere_bounds": (None, [P, P, sz, P, C.c_int]),
            "orc_par_refit": (None, [P, P, P, sz, P, C.c_int]),
            "orc_par_lbvh_build_many": (None, [P, P, P, sz, P, P, P, C.c_int]),
            "orc_trace_bvh_multi": (None, [P, sz, P, P, P, sz, P, P, P, P, C.c_int]),
            "orc_max_threads": (C.c_int, []),
            "orc_box_bounds": (None, [P, sz, P]),
            "orc_box_keys": (None, [P, sz, P, P]),
            "orc_box_refit": (None, [P, P, sz, P]),
            "orc_lbvh_build_boxes": (None, [P, sz, P, P, P]),
            "orc_box_hit": (C.c_int, [P, P, P]),
            "orc_trace_boxes_bruteforce": (None, [P, sz, P, sz, P, C.c_int]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def stable_sort(keys: np.ndarray, vals: np.ndarray | None = None):
    k = np.ascontiguousarray(keys).copy()
    v = None if vals is None else np.ascontiguousarray(vals).copy()
    lib().orc_stable_sort(_p(k), None if v is None else _p(v), len(k), k.itemsize * 8,
                          0 if v is None else v.itemsize * 8)
    return k, v


def exclusive_sum(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.empty_like(x)
    total = lib().orc_exclusive_sum(_p(x), _p(out), len(x))
    return out, int(total)


def partition(x: np.ndarray, select_positive: bool):
    x = np.ascontiguousarray(x, dtype=np.int32)
    out = np.zeros_like(x)
    total = lib().orc_partition(_p(x), _p(out), len(x), int(select_positive))
    return out, int(total)


def select(x: np.ndarray, select_positive: bool):
    x = np.ascontiguousarray(x, dtype=np.int32)
    out = np.zeros_like(x)
    total = lib().orc_select(_p(x), _p(out), len(x), int(select_positive))
    return out, int(total)


def scene_bounds(verts, idx):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    out = np.zeros(6, dtype=np.float32)
    lib().orc_scene_bounds(_p(verts), _p(idx), len(idx), _p(out))
    return out


def morton_keys(verts, idx, bounds):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    bounds = np.ascontiguousarray(bounds, dtype=np.float32)
    keys = np.zeros(len(idx), dtype=np.uint64)
    lib().orc_morton_keys(_p(verts), _p(idx), len(idx), _p(bounds), _p(keys))
    return keys


def karras(sorted_keys):
    k = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    nodes = np.zeros((2 * len(k) - 1, 4), dtype=np.int32)
    lib().orc_karras(_p(k), len(k), _p(nodes))
    return nodes


def topdown(sorted_keys):
    k = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    nodes = np.zeros((2 * len(k) - 1, 4), dtype=np.int32)
    lib().orc_topdown(_p(k), len(k), _p(nodes))
    return nodes


def refit(verts, idx, nodes, bruteforce: bool = False):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    n = (len(nodes) + 1) // 2
    aabbs = np.zeros((len(nodes), 6), dtype=np.float32)
    fn = lib().orc_refit_bruteforce if bruteforce else lib().orc_refit
    fn(_p(verts), _p(idx), _p(nodes), n, _p(aabbs))
    return aabbs


def lbvh_build(verts, idx):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    n = len(idx)
    keys = np.zeros(n, dtype=np.uint64)
    nodes = np.zeros((2 * n - 1, 4), dtype=np.int32)
    aabbs = np.zeros((2 * n - 1, 6), dtype=np.float32)
    lib().orc_lbvh_build(_p(verts), _p(idx), n, _p(keys), _p(nodes), _p(aabbs))
    return keys, nodes, aabbs


def lbvh_build_points(verts):
    """Point LBVH: keys (morton << 32 | primID, sorted), nodes int32[2n-1, 4], aabbs float32[2n-1, 6]."""
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    n = len(verts)
    keys = np.zeros(n, dtype=np.uint64)
    nodes = np.zeros((2 * n - 1, 4), dtype=np.int32)
    aabbs = np.zeros((2 * n - 1, 6), dtype=np.float32)
    lib().orc_lbvh_build_pe(_p(verts), None, n, _p(keys), _p(nodes), _p(aabbs))
    return keys, nodes, aabbs


def lbvh_build_edges(verts, edges):
    """Edge LBVH over uint2 edges into float3 vertices (same outputs as lbvh_build_points)."""
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    n = len(edges)
    keys = np.zeros(n, dtype=np.uint64)
    nodes = np.zeros((2 * n - 1, 4), dtype=np.int32)
    aabbs = np.zeros((2 * n - 1, 6), dtype=np.float32)
    lib().orc_lbvh_build_pe(_p(verts), _p(edges), n, _p(keys), _p(nodes), _p(aabbs))
    return keys, nodes, aabbs


def _boxes(boxes):
    return np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)


def box_bounds(boxes):
    b = _boxes(boxes)
    out = np.zeros(6, dtype=np.float32)
    lib().orc_box_bounds(_p(b), len(b), _p(out))
    return out


def box_keys(boxes, bounds):
    b = _boxes(boxes)
    bounds = np.ascontiguousarray(bounds, dtype=np.float32)
    keys = np.zeros(len(b), dtype=np.uint64)
    lib().orc_box_keys(_p(b), len(b), _p(bounds), _p(keys))
    return keys


def box_refit(boxes, nodes):
    b = _boxes(boxes)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    aabbs = np.zeros((len(nodes), 6), dtype=np.float32)
    lib().orc_box_refit(_p(b), _p(nodes), len(b), _p(aabbs))
    return aabbs


def lbvh_build_boxes(boxes):
    """LBVH over box primitives {lo.xyz, hi.xyz}: (sorted keys, int4 nodes, node boxes)."""
    b = _boxes(boxes)
    n = len(b)
    keys = np.zeros(n, dtype=np.uint64)
    nodes = np.zeros((2 * n - 1, 4), dtype=np.int32)
    aabbs = np.zeros((2 * n - 1, 6), dtype=np.float32)
    lib().orc_lbvh_build_boxes(_p(b), n, _p(keys), _p(nodes), _p(aabbs))
    return keys, nodes, aabbs


def box_hit(ray8, box6):
    """(hit, t) of the box surface test for one ray {ox, oy, oz, tmin, dx, dy, dz, tmax}."""
    r = np.ascontiguousarray(ray8, dtype=np.float32)
    b = np.ascontiguousarray(box6, dtype=np.float32)
    t = C.c_float()
    hit = lib().orc_box_hit(_p(r), _p(b), C.byref(t))
    return bool(hit), t.value


def trace_boxes_bruteforce(rays, boxes, threads: int = 0):
    r = _rays_f32(rays)
    b = _boxes(boxes)
    hits = np.zeros(len(r), dtype=HIT_DTYPE)
    lib().orc_trace_boxes_bruteforce(_p(r), len(r), _p(b), len(b), _p(hits), threads or os.cpu_count() or 1)
    return hits


def intersect(ray8, v0, v1, v2):
    r = np.ascontiguousarray(ray8, dtype=np.float32)
    a, b, c = (np.ascontiguousarray(v, dtype=np.float32) for v in (v0, v1, v2))
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    ok = lib().orc_intersect(_p(r), _p(a), _p(b), _p(c), C.byref(t), C.byref(u), C.byref(v))
    return bool(ok), t.value, u.value, v.value


def plane_t_fp64(ray8, v0, v1, v2) -> float:
    r = np.ascontiguousarray(ray8, dtype=np.float32)
    a, b, c = (np.ascontiguousarray(v, dtype=np.float32) for v in (v0, v1, v2))
    return float(lib().orc_plane_t_fp64(_p(r), _p(a), _p(b), _p(c)))


def _rays_f32(rays) -> np.ndarray:
    return np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)


def trace_bruteforce(rays, verts, idx, threads: int = 0):
    r = _rays_f32(rays)
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    hits = np.zeros(len(r), dtype=HIT_DTYPE)
    lib().orc_trace_bruteforce(_p(r), len(r), _p(verts), _p(idx), len(idx), _p(hits),
                               threads or os.cpu_count() or 1)
    return hits


def trace_bvh(rays, verts, idx, nodes, aabbs, threads: int = 1):
    r = _rays_f32(rays)
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    aabbs = np.ascontiguousarray(aabbs, dtype=np.float32)
    n = (len(nodes) + 1) // 2
    hits = np.zeros(len(r), dtype=HIT_DTYPE)
    counters = np.zeros(2, dtype=np.uint64)
    lib().orc_trace_bvh(_p(r), len(r), _p(verts), _p(idx), _p(nodes), _p(aabbs), n, _p(hits), _p(counters),
                        threads)
    return hits, counters


# ------------------------------------------------------------------------------------------------
# All-cores forms (baseline_omp.cpp): the multi-threaded CPU baseline. threads=0: all of
# omp_get_max_threads().
# ------------------------------------------------------------------------------------------------
def par_radix_sort(keys: np.ndarray, vals: np.ndarray | None = None, threads: int = 0):
    k = np.ascontiguousarray(keys).copy()
    v = None if vals is None else np.ascontiguousarray(vals).copy()
    lib().orc_par_radix_sort(_p(k), None if v is None else _p(v), len(k), k.itemsize * 8,
                             0 if v is None else v.itemsize * 8, threads)
    return k, v


def par_exclusive_sum(x: np.ndarray, threads: int = 0):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.empty_like(x)
    total = lib().orc_par_exclusive_sum(_p(x), _p(out), len(x), threads)
    return out, int(total)


def par_lbvh_build(verts, idx, threads: int = 0):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    n = len(idx)
    keys = np.zeros(n, dtype=np.uint64)
    nodes = np.zeros((2 * n - 1, 4), dtype=np.int32)
    aabbs = np.zeros((2 * n - 1, 6), dtype=np.float32)
    lib().orc_par_lbvh_build(_p(verts), _p(idx), n, _p(keys), _p(nodes), _p(aabbs), threads)
    return keys, nodes, aabbs


def par_lbvh_build_boxes(boxes, threads: int = 0):
    """All-cores form of lbvh_build_boxes (same keys, nodes and boxes)."""
    b = _boxes(boxes)
    n = len(b)
    keys = np.zeros(n, dtype=np.uint64)
    nodes = np.zeros((max(2 * n - 1, 0), 4), dtype=np.int32)
    aabbs = np.zeros((max(2 * n - 1, 0), 6), dtype=np.float32)
    if n:
        lib().orc_par_lbvh_build_boxes(_p(b), n, _p(keys), _p(nodes), _p(aabbs), threads)
    return keys, nodes, aabbs


def par_sphere_bounds(centres, radii, threads: int = 0):
    """Boxes {centre - r, centre + r} of spheres, all cores."""
    c = np.ascontiguousarray(centres, dtype=np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(radii, dtype=np.float32)
    out = np.zeros((len(c), 6), dtype=np.float32)
    lib().orc_par_sphere_bounds(_p(c), _p(r), len(c), _p(out), threads)
    return out


def par_refit(verts, idx, nodes, threads: int = 0):
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    n = (len(nodes) + 1) // 2
    aabbs = np.zeros((len(nodes), 6), dtype=np.float32)
    lib().orc_par_refit(_p(verts), _p(idx), _p(nodes), n, _p(aabbs), threads)
    return aabbs


def lbvh_build_many(verts, idx, sizes, threads: int = 0):
    """Concatenated meshes (3 vertices per triangle, indices local to each mesh): one oracle build
    per mesh, meshes spread over the threads. Outputs are concatenated (mesh m's nodes at
    2 * first - m)."""
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    total, m = int(sizes.sum()), len(sizes)
    keys = np.zeros(total, dtype=np.uint64)
    nodes = np.zeros((2 * total - m, 4), dtype=np.int32)
    aabbs = np.zeros((2 * total - m, 6), dtype=np.float32)
    lib().orc_par_lbvh_build_many(_p(verts), _p(idx), _p(sizes), m, _p(keys), _p(nodes), _p(aabbs), threads)
    return keys, nodes, aabbs


def trace_bvh_multi(rays, verts, idx, sizes, nodes, aabbs, threads: int = 0):
    """Closest hit over the meshes of lbvh_build_many (identity instances): hits and the mesh of
    each hit (-1 on a miss), the minimum of (t, mesh, primID)."""
    r = _rays_f32(rays)
    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    hits = np.zeros(len(r), dtype=HIT_DTYPE)
    inst = np.zeros(len(r), dtype=np.int32)
    lib().orc_trace_bvh_multi(_p(r), len(r), _p(np.ascontiguousarray(verts, dtype=np.float32)),
                              _p(np.ascontiguousarray(idx, dtype=np.uint32)), _p(sizes), len(sizes),
                              _p(np.ascontiguousarray(nodes, dtype=np.int32)),
                              _p(np.ascontiguousarray(aabbs, dtype=np.float32)), _p(hits), _p(inst), threads)
    return hits, inst


def host_info() -> dict:
    """The CPU the baseline runs on, read at run time."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": int(lib().orc_num_procs()), "omp_max_threads": int(lib().orc_max_threads()),
            "os_cpu_count": os.cpu_count(), "cpu_model": model}
