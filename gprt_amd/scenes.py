"""Seeded synthetic scenes and rays for the acceleration-structure path (numpy, host side).

Counter-based SplitMix64 streams make every value a pure function of (seed, stream, index), so
the CPU oracle and the GPU trace identical inputs and a subset of a workload can be regenerated
without generating the rest. Floats are drawn as (u >> 40) * 2^-24 (24 random bits, exact in
FP32). Geometry and directions are computed in FP64 and rounded once to FP32.

Workloads:
  uniform_scene  16,384 triangles, centroids U[0,1]^3, vertex offsets U[-0.01, 0.01]
  blob_scene     64 Gaussian blobs (centres U[0,1]^3, sigma log-uniform [0.02, 0.1]),
                 centroids by FP64 Box-Muller, vertex offsets U[-0.002, 0.002]
  pinhole_rays   origin (0.5, 0.5, -1), +z, 60 degree vertical field of view, pixel centres
  bounce_rays    cosine-weighted hemisphere rays (Malley's method) from primary hit points,
                 geometric normal facing the incoming ray, tmin = 1e-4; misses are skipped and
                 hits are cycled in order
"""
from __future__ import annotations

import math

import numpy as np

from .api import RAY_DTYPE

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
TMAX = np.float32(1e30)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def rand_u64(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        key = _mix(np.array([(seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03 + 1) & (2**64 - 1)],
                            dtype=np.uint64))[0]
        ctr = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        return _mix(key + ctr * _GOLD)


def rand01(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """Uniform [0, 1) doubles with 24 random bits."""
    return (rand_u64(seed, stream, n, start) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


def _tris_from_centroids(cent: np.ndarray, off: np.ndarray):
    n = len(cent)
    verts = (cent[:, None, :] + off.reshape(n, 3, 3)).astype(np.float32).reshape(3 * n, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return verts, idx


def uniform_scene(n: int = 16384, seed: int = 1):
    """Triangles with centroids uniform in the unit cube, vertex offsets uniform in +-0.01."""
    cent = rand01(seed, 0, 3 * n).reshape(n, 3)
    off = rand01(seed, 1, 9 * n) * 0.02 - 0.01
    return _tris_from_centroids(cent, off)


def blob_scene(n: int = 4194304, seed: int = 2, blobs: int = 64, offset: float = 0.002):
    """Triangles drawn from Gaussian blobs (equal share per blob), vertex offsets +-offset."""
    centers = rand01(seed, 10, 3 * blobs).reshape(blobs, 3)
    sig = np.exp(math.log(0.02) + rand01(seed, 11, blobs) * (math.log(0.1) - math.log(0.02)))
    b = (np.arange(n, dtype=np.int64) * blobs) // n
    u1 = 1.0 - rand01(seed, 12, 3 * n)
    u2 = rand01(seed, 13, 3 * n)
    g = (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)).reshape(n, 3)
    cent = centers[b] + sig[b][:, None] * g
    off = rand01(seed, 14, 9 * n) * (2.0 * offset) - offset
    return _tris_from_centroids(cent, off)


def pinhole_rays(width: int, height: int, origin=(0.5, 0.5, -1.0), fov_deg: float = 60.0) -> np.ndarray:
    tan_half = math.tan(math.radians(fov_deg) / 2.0)
    aspect = width / height
    ys, xs = np.meshgrid(np.arange(height, dtype=np.float64), np.arange(width, dtype=np.float64), indexing="ij")
    px = ((xs + 0.5) / width * 2.0 - 1.0) * tan_half * aspect
    py = ((ys + 0.5) / height * 2.0 - 1.0) * tan_half
    d = np.stack([px.ravel(), py.ravel(), np.ones(width * height)], axis=1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(width * height, dtype=RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = origin
    r["tmin"] = 0.0
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["tmax"] = TMAX
    return r


def bounce_rays(prim_rays: np.ndarray, hits: np.ndarray, verts: np.ndarray, idx: np.ndarray, count: int,
                seed: int = 3, tmin: float = 1e-4) -> np.ndarray:
    """Cosine-weighted bounce rays from the hit points of `prim_rays` (hits: HIT_DTYPE)."""
    hit_ids = np.nonzero(hits["primID"] >= 0)[0]
    out = np.zeros(count, dtype=RAY_DTYPE)
    if len(hit_ids) == 0 or count == 0:
        return out[:0]
    src = hit_ids[np.arange(count) % len(hit_ids)]
    r, h = prim_rays[src], hits[src]
    o = np.stack([r["ox"], r["oy"], r["oz"]], 1).astype(np.float64)
    d = np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(np.float64)
    p = o + h["t"].astype(np.float64)[:, None] * d
    tri = idx[h["primID"]].astype(np.int64)
    v0, v1, v2 = (verts[tri[:, k]].astype(np.float64) for k in range(3))
    nrm = np.cross(v1 - v0, v2 - v0)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-300)
    flip = np.einsum("ij,ij->i", nrm, d) > 0
    nrm[flip] *= -1.0
    u1 = rand01(seed, 20, count)
    u2 = rand01(seed, 21, count)
    rad = np.sqrt(u1)
    phi = 2.0 * math.pi * u2
    lx, ly, lz = rad * np.cos(phi), rad * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - u1))
    # Orthonormal basis around the normal (Duff et al., JCGT 2017).
    nx, ny, nz = nrm[:, 0], nrm[:, 1], nrm[:, 2]
    sgn = np.where(nz >= 0.0, 1.0, -1.0)
    a = -1.0 / (sgn + nz)
    b = nx * ny * a
    t1 = np.stack([1.0 + sgn * nx * nx * a, sgn * b, -sgn * nx], 1)
    t2 = np.stack([b, sgn + ny * ny * a, -ny], 1)
    dirs = lx[:, None] * t1 + ly[:, None] * t2 + lz[:, None] * nrm
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    out["ox"], out["oy"], out["oz"] = p[:, 0], p[:, 1], p[:, 2]
    out["tmin"] = tmin
    out["dx"], out["dy"], out["dz"] = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    out["tmax"] = TMAX
    return out


def bounce_rays_multi(prim_rays: np.ndarray, hits: np.ndarray, inst: np.ndarray, meshes, count: int,
                      seed: int = 3, tmin: float = 1e-4, transforms=None) -> np.ndarray:
    """bounce_rays over instances: hit i lies on triangle hits[i].primID of meshes[inst[i]], whose
    vertices are carried to world space by transforms[inst[i]] (3x4 [A | b], FP64, fixed order)."""
    world = []
    for k, (v, idx) in enumerate(meshes):
        if transforms is None:
            world.append(v)
            continue
        m = np.asarray(transforms[k], dtype=np.float32).astype(np.float64)
        vd = v.astype(np.float64)
        w = np.empty_like(vd)
        for c in range(3):
            w[:, c] = ((m[c, 0] * vd[:, 0] + m[c, 1] * vd[:, 1]) + m[c, 2] * vd[:, 2]) + m[c, 3]
        world.append(w)
    offs_v = np.cumsum([0] + [len(v) for v in world])
    offs_t = np.cumsum([0] + [len(i) for _, i in meshes])
    verts = np.concatenate(world)
    idx = np.concatenate([i.astype(np.int64) + offs_v[k] for k, (_, i) in enumerate(meshes)])
    h = hits.copy()
    hit = h["primID"] >= 0
    h["primID"][hit] = h["primID"][hit] + offs_t[inst[hit]]
    return bounce_rays(prim_rays, h, verts, idx, count, seed, tmin)


def object_rays(rays: np.ndarray, transform) -> np.ndarray:
    """Rays in the object space of an instance with 3x4 transform [A | b] (world = A x + b): the
    inverse in FP64 (adjugate / determinant), rounded to FP32, applied in FP32 in a fixed order."""
    A = np.asarray(transform, dtype=np.float32).astype(np.float64)
    a, b = A[:, :3], A[:, 3]
    det = (a[0, 0] * (a[1, 1] * a[2, 2] - a[1, 2] * a[2, 1]) - a[0, 1] * (a[1, 0] * a[2, 2] - a[1, 2] * a[2, 0])
           + a[0, 2] * (a[1, 0] * a[2, 1] - a[1, 1] * a[2, 0]))
    idt = 1.0 / det
    inv = np.array([
        [(a[1, 1] * a[2, 2] - a[1, 2] * a[2, 1]) * idt, (a[0, 2] * a[2, 1] - a[0, 1] * a[2, 2]) * idt,
         (a[0, 1] * a[1, 2] - a[0, 2] * a[1, 1]) * idt],
        [(a[1, 2] * a[2, 0] - a[1, 0] * a[2, 2]) * idt, (a[0, 0] * a[2, 2] - a[0, 2] * a[2, 0]) * idt,
         (a[0, 2] * a[1, 0] - a[0, 0] * a[1, 2]) * idt],
        [(a[1, 0] * a[2, 1] - a[1, 1] * a[2, 0]) * idt, (a[0, 1] * a[2, 0] - a[0, 0] * a[2, 1]) * idt,
         (a[0, 0] * a[1, 1] - a[0, 1] * a[1, 0]) * idt]])
    binv = -(inv[:, 0] * b[0] + inv[:, 1] * b[1] + inv[:, 2] * b[2])
    m = np.concatenate([inv, binv[:, None]], 1).astype(np.float32)
    out = rays.copy()
    o = [rays["ox"], rays["oy"], rays["oz"]]
    d = [rays["dx"], rays["dy"], rays["dz"]]
    for c, (on, dn) in enumerate([("ox", "dx"), ("oy", "dy"), ("oz", "dz")]):
        out[on] = ((m[c, 0] * o[0] + m[c, 1] * o[1]) + m[c, 2] * o[2]) + m[c, 3]
        out[dn] = (m[c, 0] * d[0] + m[c, 1] * d[1]) + m[c, 2] * d[2]
    return out


def random_rays(n: int, seed: int, lo=-0.5, hi=1.5) -> np.ndarray:
    """Rays with origins uniform in [lo, hi]^3 and directions uniform on the sphere."""
    o = rand01(seed, 30, 3 * n).reshape(n, 3) * (hi - lo) + lo
    z = rand01(seed, 31, n) * 2.0 - 1.0
    ph = rand01(seed, 32, n) * 2.0 * math.pi
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([s * np.cos(ph), s * np.sin(ph), z], 1)
    r = np.zeros(n, dtype=RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = o[:, 0], o[:, 1], o[:, 2]
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["tmin"] = 0.0
    r["tmax"] = TMAX
    return r


def jittered_pinhole_rays(n: int, seed: int, width: int = 2048, height: int = 2048, origin=(0.5, 0.5, -1.0),
                          fov_deg: float = 60.0) -> np.ndarray:
    """Pinhole rays through uniformly random image-plane points (a seeded variant of pinhole_rays,
    used for training data drawn from the same distribution as the primary rays)."""
    tan_half = math.tan(math.radians(fov_deg) / 2.0)
    aspect = width / height
    px = (rand01(seed, 40, n) * 2.0 - 1.0) * tan_half * aspect
    py = (rand01(seed, 41, n) * 2.0 - 1.0) * tan_half
    d = np.stack([px, py, np.ones(n)], axis=1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, dtype=RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = origin
    r["tmin"] = 0.0
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["tmax"] = TMAX
    return r


def _rand01_keyed(keys: np.ndarray, ctr: np.ndarray) -> np.ndarray:
    """rand01 with a per-element stream key (as rand_u64's `key`) and 0-based counter."""
    with np.errstate(over="ignore"):
        return (_mix(keys + (ctr.astype(np.uint64) + np.uint64(1)) * _GOLD) >> np.uint64(40)).astype(np.float64) \
            * (2.0 ** -24)


def _stream_key(seeds: np.ndarray, stream: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        k = seeds.astype(np.uint64) * _GOLD + np.uint64((stream * 0xD1B54A32D192ED03 + 1) & (2**64 - 1))
        return _mix(k)


def batch_sizes(count: int = 4096, seed: int = 3, lo: int = 64, hi: int = 262144) -> np.ndarray:
    """Triangle counts log-uniform in [lo, hi]: round(lo * (hi/lo)^u)."""
    u = rand01(seed, 50, count)
    return np.rint(lo * (hi / lo) ** u).astype(np.int64)


def _blob_params(seeds: np.ndarray, blobs: int):
    M = len(seeds)
    centers = _rand01_keyed(np.repeat(_stream_key(seeds, 10), 3 * blobs), np.tile(np.arange(3 * blobs), M))
    sig = np.exp(math.log(0.02) + _rand01_keyed(np.repeat(_stream_key(seeds, 11), blobs),
                                                np.tile(np.arange(blobs), M)) * (math.log(0.1) - math.log(0.02)))
    return centers.reshape(M, blobs, 3), sig.reshape(M, blobs)


def _blob_range(sizes, offs, seeds, centers, sig, blobs, offset, t0, t1, out):
    """Vertices of global triangles [t0, t1) into out[3*t0 : 3*t1]."""
    m = np.searchsorted(offs, np.arange(t0, t1), side="right") - 1
    loc = np.arange(t0, t1) - offs[m]
    n = sizes[m]
    b = (loc * blobs) // n
    c3 = (3 * loc)[:, None] + np.arange(3)
    u1 = 1.0 - _rand01_keyed(_stream_key(seeds, 12)[m][:, None], c3)
    u2 = _rand01_keyed(_stream_key(seeds, 13)[m][:, None], c3)
    g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
    cent = centers[m, b] + sig[m, b][:, None] * g
    c9 = (9 * loc)[:, None] + np.arange(9)
    off = _rand01_keyed(_stream_key(seeds, 14)[m][:, None], c9) * (2.0 * offset) - offset
    out[3 * t0:3 * t1] = (cent[:, None, :] + off.reshape(-1, 3, 3)).astype(np.float32).reshape(-1, 3)


_SHARED = {}


def _blob_worker(args):
    t0, t1 = args
    g = _SHARED
    _blob_range(g["sizes"], g["offs"], g["seeds"], g["centers"], g["sig"], g["blobs"], g["offset"], t0, t1, g["out"])
    return t1 - t0


def blob_meshes(sizes, seed0: int = 1000, blobs: int = 64, offset: float = 0.002, workers: int = 1):
    """Concatenated blob meshes: mesh i equals blob_scene(sizes[i], seed0 + i) exactly.
    Returns (verts float32 [3T, 3], idx uint32 [T, 3] with per-mesh local indices, prim offsets).
    `workers` > 1 splits the triangle range over forked processes writing into shared memory."""
    from multiprocessing import get_context, shared_memory

    sizes = np.asarray(sizes, dtype=np.int64)
    M = len(sizes)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    T = int(offs[-1])
    seeds = seed0 + np.arange(M, dtype=np.int64)
    centers, sig = _blob_params(seeds, blobs)
    chunk = 1 << 20
    ranges = [(t0, min(T, t0 + chunk)) for t0 in range(0, T, chunk)]
    if workers <= 1 or len(ranges) == 1:
        verts = np.empty((3 * T, 3), dtype=np.float32)
        for t0, t1 in ranges:
            _blob_range(sizes, offs, seeds, centers, sig, blobs, offset, t0, t1, verts)
    else:
        shm = shared_memory.SharedMemory(create=True, size=max(1, 3 * T * 3 * 4))
        try:
            out = np.ndarray((3 * T, 3), dtype=np.float32, buffer=shm.buf)
            _SHARED.update(sizes=sizes, offs=offs, seeds=seeds, centers=centers, sig=sig, blobs=blobs,
                           offset=offset, out=out)
            with get_context("fork").Pool(workers) as pool:
                for _ in pool.imap_unordered(_blob_worker, ranges):
                    pass
            verts = out.copy()
            del out
        finally:
            _SHARED.clear()
            shm.close()
            shm.unlink()
    idx = (np.arange(3 * T, dtype=np.int64) - 3 * np.repeat(offs[:-1], 3 * sizes)).astype(np.uint32).reshape(T, 3)
    return verts, idx, offs
