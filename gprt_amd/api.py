"""ctypes bindings of libgprt_hip.so (declared in include/gprt_hip.h and include/gprt_hip_ext.h).

The names, argument order and meaning are GPRT's host API (gprt/gprt_host.h) for the
acceleration-structure path; see the header for the reference line each one replaces. The
library is the product: this module only loads it and declares signatures. It raises at import
time of `load()` if the shared library has not been built — there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgprt_hip.so")

GPRT_ERROR_RAISE = 0
GPRT_ERROR_RECORD = 1

GPRT_AABBS = 1
GPRT_TRIANGLES = 2
GPRT_BUILD_MODE_FAST_BUILD_NO_UPDATE = 1
GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE = 2
GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE = 3
GPRT_BUILD_MODE_FAST_TRACE_AND_UPDATE = 4

RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("primID", "<i4"), ("u", "<f4"), ("v", "<f4")])


class GPRTBuildParams(C.Structure):
    """gprt/gprt_host.h:1423-1469 (same members, order and defaults)."""
    _fields_ = [("buildMode", C.c_int), ("allowCompaction", C.c_bool), ("minimizeMemory", C.c_bool),
                ("hasMotionBlur", C.c_bool), ("test0", C.c_bool), ("test1", C.c_bool)]

    def __init__(self, buildMode=GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE):
        super().__init__(buildMode, False, False, False, False, False)


GPRT_LBVH_POINTS = 1
GPRT_LBVH_EDGES = 2
GPRT_LBVH_TRIANGLES = 4


class GPRTLBVH(C.Structure):
    """gprt/development/gprt_lbvh.h:64-85: the LBVH and its result buffers."""
    _fields_ = [("type", C.c_int), ("numPrims", C.c_uint32), ("mortonCodes", C.c_void_p), ("ids", C.c_void_p),
                ("nodes", C.c_void_p), ("aabbs", C.c_void_p), ("impl", C.c_void_p)]


class GPRTInstance(C.Structure):
    """gprt::Instance (gprt/gprt_shared.h:92-101): 3x4 row-major transform [A | b], then
    instanceCustomIndex:24 | mask:8, geometryIndex:24 | flags:8, accelAddress (64 bytes)."""
    _fields_ = [("transform", (C.c_float * 4) * 3), ("instanceCustomIndex", C.c_uint32, 24), ("mask", C.c_uint32, 8),
                ("geometryIndex", C.c_uint32, 24), ("flags", C.c_uint32, 8), ("accelAddress", C.c_uint64)]


# The same record as a numpy dtype (little-endian bitfields: low 24 bits, then the top 8).
INSTANCE_DTYPE = np.dtype([("transform", "<f4", (3, 4)), ("word0", "<u4"), ("word1", "<u4"),
                           ("accelAddress", "<u8")])


def instance_records(accels, transforms=None, masks=None) -> np.ndarray:
    """gprt::Instance records for `accels` (gprtAccelGetInstance), optionally with 3x4 transforms
    and 8-bit masks; instanceCustomIndex = position in the list."""
    lib = load()
    rec = np.zeros(len(accels), dtype=INSTANCE_DTYPE)
    for i, a in enumerate(accels):
        inst = lib.gprtAccelGetInstance(a.h)
        rec["accelAddress"][i] = inst.accelAddress
        rec["transform"][i] = np.eye(3, 4, dtype=np.float32) if transforms is None else transforms[i]
        m = 0xFF if masks is None else int(masks[i]) & 0xFF
        rec["word0"][i] = (i & 0xFFFFFF) | (m << 24)
    return rec


_P = C.c_void_p
_SIGS = {
    # name: (restype, argtypes)
    "gprtVersion": (C.c_char_p, []),
    "gprtSetErrorMode": (None, [C.c_int]),
    "gprtSetSyncTimeout": (None, [_P, C.c_double]),
    "gprtGetLastError": (C.c_char_p, []),
    "gprtClearLastError": (None, []),
    "gprtContextCreate": (_P, [C.POINTER(C.c_int32), C.c_int]),
    "gprtContextDestroy": (None, [_P]),
    "gprtHostBufferCreate": (_P, [_P, C.c_size_t, C.c_size_t, _P, C.c_size_t]),
    "gprtDeviceBufferCreate": (_P, [_P, C.c_size_t, C.c_size_t, _P, C.c_size_t]),
    "gprtSharedBufferCreate": (_P, [_P, C.c_size_t, C.c_size_t, _P, C.c_size_t]),
    "gprtBufferClear": (None, [_P]),
    "gprtBufferDestroy": (None, [_P]),
    "gprtBufferGetSize": (C.c_size_t, [_P, C.c_int]),
    "gprtBufferGetHostPointer": (_P, [_P, C.c_int]),
    "gprtBufferMap": (None, [_P, C.c_int]),
    "gprtBufferUnmap": (None, [_P, C.c_int]),
    "gprtBufferGetDevicePointer": (_P, [_P, C.c_int]),
    "gprtBufferResize": (None, [_P, _P, C.c_size_t, C.c_size_t, C.c_bool, C.c_int]),
    "gprtBufferCopy": (None, [_P, _P, _P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int, C.c_int]),
    "gprtBufferExclusiveSum": (C.c_uint32, [_P, _P, _P, _P]),
    "gprtBufferExclusiveSumNoAlloc": (C.c_uint32, [_P, _P, _P]),
    "gprtBufferPartition": (C.c_uint32, [_P, _P, C.c_bool, _P, _P]),
    "gprtBufferSelect": (C.c_uint32, [_P, _P, C.c_bool, _P, _P]),
    "gprtBufferSort": (None, [_P, _P, _P]),
    "gprtBufferSortPayload": (None, [_P, _P, _P, _P]),
    "gprtBufferSortKV": (None, [_P, _P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "gprtBufferSort32": (None, [_P, _P, _P]),
    "gprtBufferSort64": (None, [_P, _P, _P]),
    "gprtBufferSortPayload32": (None, [_P, _P, _P, _P]),
    "gprtBufferSortPayload64": (None, [_P, _P, _P, _P]),
    "gprtSetSortConfig": (None, [_P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "gprtGetSortConfig": (None, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "gprtAccelBuildBatchAutotune": (C.c_uint32, [_P, C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_char_p,
                                                 C.c_size_t]),
    "gprtGeomTypeCreate": (_P, [_P, C.c_int, C.c_size_t]),
    "gprtAABBsSetPositions": (None, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "gprtAABBAccelCreate": (_P, [_P, _P, C.c_uint]),
    "gprtGeomTypeDestroy": (None, [_P]),
    "gprtGeomCreate": (_P, [_P, _P]),
    "gprtGeomDestroy": (None, [_P]),
    "gprtTrianglesSetVertices": (None, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "gprtTrianglesSetIndices": (None, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "gprtTriangleAccelCreate": (_P, [_P, _P, C.c_uint]),
    "gprtAccelDestroy": (None, [_P]),
    "gprtAccelBuild": (None, [_P, _P, GPRTBuildParams]),
    "gprtAccelBuildBatch": (None, [_P, C.POINTER(C.c_void_p), C.c_uint32, GPRTBuildParams]),
    "gprtAccelUpdate": (None, [_P, _P]),
    "gprtAccelGetSize": (C.c_size_t, [_P, C.c_int]),
    "gprtAccelGetDeviceAddress": (_P, [_P, C.c_int]),
    "gprtAccelGetPrimCount": (C.c_uint32, [_P]),
    "gprtAccelGetBounds": (None, [_P, _P, _P]),
    "gprtPointsLBVHCreate": (GPRTLBVH, [_P, _P, C.c_uint32]),
    "gprtEdgesLBVHCreate": (GPRTLBVH, [_P, _P, _P, C.c_uint32]),
    "gprtTriangleLBVHCreate": (GPRTLBVH, [_P, _P, _P, C.c_uint32]),
    "gprtLBVHBuild": (None, [_P, C.POINTER(GPRTLBVH)]),
    "gprtLBVHDestroy": (None, [C.POINTER(GPRTLBVH)]),
    "gprtAccelTraceClosestHit": (None, [_P, _P, _P, _P, C.c_uint32]),
    "gprtAccelExportLBVH": (None, [_P, _P, _P, _P, _P]),
    "gprtProposalCreate": (_P, [_P, C.c_uint32, C.c_uint64]),
    "gprtProposalDestroy": (None, [_P, _P]),
    "gprtProposalBindAccel": (C.c_uint32, [_P, _P, _P, C.c_uint32]),
    "gprtProposalParamCount": (C.c_size_t, []),
    "gprtProposalSetParams": (None, [_P, _P, _P, C.c_size_t]),
    "gprtProposalGetParams": (None, [_P, _P, _P, C.c_size_t]),
    "gprtAccelSerialize": (C.c_size_t, [_P, _P, _P, C.c_size_t]),
    "gprtAccelDeserialize": (C.c_bool, [_P, _P, _P, C.c_size_t]),
    "gprtProposalSerialize": (C.c_size_t, [_P, _P, _P, C.c_size_t, C.c_bool]),
    "gprtProposalDeserialize": (C.c_bool, [_P, _P, _P, C.c_size_t]),
    "gprtProposalTrain": (C.c_float, [_P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_float, C.POINTER(C.c_float)]),
    "gprtProposalInfer": (None, [_P, _P, _P, C.c_uint32, _P]),
    "gprtProposalLogits": (None, [_P, _P, _P, C.c_uint32, _P]),
    "gprtProposalGradients": (None, [_P, _P, _P, _P, C.c_uint32, _P, C.c_size_t]),
    "gprtProposalExportTreelets": (None, [_P, _P, _P, _P]),
    "gprtProposalSetGate": (None, [_P, _P, C.c_float]),
    "gprtProposalSetOverlap": (None, [_P, _P, C.c_uint32, C.c_uint32]),
    "gprtProposalGetGate": (C.c_float, [_P]),
    "gprtProposalCalibrateGate": (C.c_float, [_P, _P, _P, _P, C.c_uint32]),
    "gprtAccelTraceClosestHitEx": (None, [_P, _P, _P, _P, _P, C.c_uint32, _P]),
    "gprtSynthBlobMeshes": (None, [_P, _P, C.c_uint32, C.c_uint64, C.c_float, _P, _P]),
    "gprtSynthPerturb": (None, [_P, _P, _P, C.c_uint64, C.c_uint64, C.c_float]),
    "gprtSynthBounceRays": (None, [_P, _P, _P, _P, _P, C.c_uint32, _P, C.c_uint32, C.c_uint64, C.c_float]),
    "gprtInstanceAccelCreate": (_P, [_P, C.c_uint32, _P]),
    "gprtAccelGetInstance": (GPRTInstance, [_P]),
    "gprtAccelTraceClosestHitDetail": (None, [_P, _P, _P, _P, _P, C.c_uint32, _P, _P, _P]),
    "gprtProposalTrainLabels": (C.c_float, [_P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_float,
                                            C.POINTER(C.c_float)]),
    "gprtModuleCreate": (_P, [_P, _P, C.c_size_t]),
    "gprtModuleDestroy": (None, [_P]),
    "gprtComputeCreate": (_P, [_P, _P, C.c_char_p]),
    "gprtComputeDestroy": (None, [_P]),
    "gprtComputeLaunchArgs": (C.c_uint64, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _P, C.c_size_t]),
    "gprtComputeSynchronize": (C.c_uint64, [_P]),
    "gprtDeviceSynchronize": (C.c_uint64, [_P]),
    "gprtBeginProfile": (None, [_P]),
    "gprtEndProfile": (C.c_float, [_P]),
    "gprtKernelTimerEnable": (None, [_P, C.c_bool]),
    "gprtKernelTimerReset": (None, [_P]),
    "gprtKernelTimerQuery": (C.c_double, [_P, C.c_char_p, C.POINTER(C.c_uint32)]),
    "gprtKernelTimerReport": (C.c_size_t, [_P, C.c_char_p, C.c_size_t]),
}

_lib = None


def load(path: str | None = None) -> C.CDLL:
    """Load libgprt_hip.so and declare every signature. Raises if it is missing or incomplete."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libgprt_hip.so not built ({p}); run __graft_entry__.build() first")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is missing
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def symbols() -> list[str]:
    return list(_SIGS)


# ------------------------------------------------------------------------------------------------
# Thin helpers (numpy <-> GPRT buffers)
# ------------------------------------------------------------------------------------------------
class Context:
    """A GPRT context on one HIP device (gprtContextCreate / gprtContextDestroy)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        dev = (C.c_int32 * 1)(device)
        self.h = self.lib.gprtContextCreate(dev, 1)
        if not self.h:
            raise RuntimeError("gprtContextCreate failed: %s" % last_error())

    def close(self):
        if self.h:
            self.lib.gprtContextDestroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_buffer(self, data=None, nbytes: int | None = None) -> "Buffer":
        if data is not None:
            arr = np.ascontiguousarray(data)
            h = self.lib.gprtDeviceBufferCreate(self.h, arr.itemsize, arr.size,
                                                arr.ctypes.data_as(C.c_void_p), 16)
            size = arr.nbytes
        else:
            size = int(nbytes or 0)
            h = self.lib.gprtDeviceBufferCreate(self.h, 1, size, None, 16)
        if not h:
            raise RuntimeError("gprtDeviceBufferCreate failed: %s" % last_error())
        return Buffer(self, h)

    def sync(self):
        self.lib.gprtComputeSynchronize(self.h)


class Buffer:
    def __init__(self, ctx: Context, h):
        self.ctx, self.h, self.lib = ctx, h, ctx.lib

    @property
    def nbytes(self) -> int:
        return int(self.lib.gprtBufferGetSize(self.h, 0))

    def read(self, dtype, count: int | None = None) -> np.ndarray:
        """Map (waits for the stream), copy out, unmap."""
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        self.lib.gprtBufferMap(self.h, 0)
        ptr = self.lib.gprtBufferGetHostPointer(self.h, 0)
        out = np.empty(n, dtype=dt)
        if n:
            C.memmove(out.ctypes.data, ptr, n * dt.itemsize)
        self.lib.gprtBufferUnmap(self.h, 0)
        return out

    def write(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.lib.gprtBufferMap(self.h, 0)
        ptr = self.lib.gprtBufferGetHostPointer(self.h, 0)
        C.memmove(ptr, arr.ctypes.data, arr.nbytes)
        self.lib.gprtBufferUnmap(self.h, 0)

    def destroy(self):
        if self.h:
            self.lib.gprtBufferDestroy(self.h)
            self.h = None


def last_error() -> str | None:
    e = load().gprtGetLastError()
    return e.decode() if e else None


class TriangleAccel:
    """gprtGeomTypeCreate + gprtGeomCreate + gprtTriangles* + gprtTriangleAccelCreate."""

    def __init__(self, ctx: Context, verts: np.ndarray = None, idx: np.ndarray = None, *, buffers=None,
                 offsets=(0, 0), owns_buffers: bool = True):
        self.ctx, self.lib = ctx, ctx.lib
        if buffers is None:
            verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
            idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
            self.nv, self.nt = len(verts), len(idx)
            self.vbuf = ctx.device_buffer(verts)
            self.ibuf = ctx.device_buffer(idx)
        else:
            self.vbuf, self.ibuf, self.nv, self.nt = buffers
        self.owns_buffers = owns_buffers
        self.gtype = self.lib.gprtGeomTypeCreate(ctx.h, GPRT_TRIANGLES, 0)
        self.geom = self.lib.gprtGeomCreate(ctx.h, self.gtype)
        self.lib.gprtTrianglesSetVertices(self.geom, self.vbuf.h, self.nv, 12, offsets[0])
        self.lib.gprtTrianglesSetIndices(self.geom, self.ibuf.h, self.nt, 12, offsets[1])
        self.h = self.lib.gprtTriangleAccelCreate(ctx.h, self.geom, 0)

    @classmethod
    def from_buffers(cls, ctx: Context, vbuf: Buffer, ibuf: Buffer, nverts: int, ntris: int) -> "TriangleAccel":
        """An accel over geometry already in device buffers (float3 vertices, uint3 indices)."""
        return cls(ctx, buffers=(vbuf, ibuf, nverts, ntris))

    def build(self, mode=GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE):
        self.lib.gprtAccelBuild(self.ctx.h, self.h, GPRTBuildParams(mode))

    def update(self):
        self.lib.gprtAccelUpdate(self.ctx.h, self.h)

    def export(self):
        """Reference-layout LBVH: (keys u64[N], nodes int32[2N-1,4], aabbs float32[2N-1,6])."""
        nb = self.ctx.device_buffer(nbytes=16)
        ab = self.ctx.device_buffer(nbytes=16)
        kb = self.ctx.device_buffer(nbytes=16)
        self.lib.gprtAccelExportLBVH(self.ctx.h, self.h, nb.h, ab.h, kb.h)
        n = self.nt
        keys = kb.read(np.uint64, n)
        nodes = nb.read(np.int32, 4 * (2 * n - 1)).reshape(-1, 4)
        aabbs = ab.read(np.float32, 6 * (2 * n - 1)).reshape(-1, 6)
        for b in (nb, ab, kb):
            b.destroy()
        return keys, nodes, aabbs

    def trace(self, rays_buf: Buffer, hits_buf: Buffer, count: int):
        self.lib.gprtAccelTraceClosestHit(self.ctx.h, self.h, rays_buf.h, hits_buf.h, count)

    def bounds(self) -> np.ndarray:
        """gprtAccelGetBounds: the root box {min.xyz, max.xyz}."""
        return accel_bounds(self.ctx, self)

    def serialize(self) -> bytes:
        """gprtAccelSerialize: the built accel as bytes."""
        n = int(self.lib.gprtAccelSerialize(self.ctx.h, self.h, None, 0))
        buf = C.create_string_buffer(max(n, 1))
        if int(self.lib.gprtAccelSerialize(self.ctx.h, self.h, buf, n)) != n or n == 0:
            raise RuntimeError("gprtAccelSerialize failed: %s" % last_error())
        return buf.raw[:n]

    def deserialize(self, blob: bytes) -> bool:
        return bool(self.lib.gprtAccelDeserialize(self.ctx.h, self.h, blob, len(blob)))

    def destroy(self):
        if self.h:
            self.lib.gprtAccelDestroy(self.h)
            self.lib.gprtGeomDestroy(self.geom)
            self.lib.gprtGeomTypeDestroy(self.gtype)
            if self.owns_buffers:
                self.vbuf.destroy()
                self.ibuf.destroy()
            self.h = None


class AABBAccel(TriangleAccel):
    """gprtGeomTypeCreate(GPRT_AABBS) + gprtGeomCreate + gprtAABBsSetPositions + gprtAABBAccelCreate
    over boxes {minX, minY, minZ, maxX, maxY, maxZ} (float32 [n, 6])."""

    def __init__(self, ctx: Context, boxes: np.ndarray):
        self.ctx, self.lib = ctx, ctx.lib
        boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
        self.nt = self.nv = len(boxes)
        self.vbuf = ctx.device_buffer(boxes)
        self.ibuf = None
        self.owns_buffers = True
        self.gtype = self.lib.gprtGeomTypeCreate(ctx.h, GPRT_AABBS, 0)
        self.geom = self.lib.gprtGeomCreate(ctx.h, self.gtype)
        self.lib.gprtAABBsSetPositions(self.geom, self.vbuf.h, self.nt, 24, 0)
        self.h = self.lib.gprtAABBAccelCreate(ctx.h, self.geom, 0)

    def destroy(self):
        if self.h:
            self.lib.gprtAccelDestroy(self.h)
            self.lib.gprtGeomDestroy(self.geom)
            self.lib.gprtGeomTypeDestroy(self.gtype)
            self.vbuf.destroy()
            self.h = None


def build_batch(ctx: Context, accels, mode=GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE):
    """gprtAccelBuildBatch over a list of TriangleAccel objects."""
    arr = (C.c_void_p * len(accels))(*[a.h for a in accels])
    ctx.lib.gprtAccelBuildBatch(ctx.h, arr, len(accels), GPRTBuildParams(mode))


def synth_blob_meshes(ctx: Context, sizes, seed0: int, offset: float = 0.002):
    """gprtSynthBlobMeshes: blob meshes generated on the device into one vertex and one index
    buffer (mesh i: seed seed0 + i, indices local to the mesh). Returns (vbuf, ibuf)."""
    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    total = int(sizes.sum())
    vb, ib = ctx.device_buffer(nbytes=36 * total), ctx.device_buffer(nbytes=12 * total)
    ctx.lib.gprtSynthBlobMeshes(ctx.h, sizes.ctypes.data_as(C.c_void_p), len(sizes), seed0, offset, vb.h, ib.h)
    return vb, ib


def mesh_accels(ctx: Context, vbuf: Buffer, ibuf: Buffer, sizes):
    """One TriangleAccel per mesh of concatenated geometry (3 vertices per triangle, as written
    by synth_blob_meshes), referencing the shared buffers at the mesh's offsets."""
    accs, t0 = [], 0
    if 36 * int(np.sum(sizes)) > 0xFFFFFFFF:
        raise ValueError("geometry offsets are 32-bit (gprtTrianglesSetVertices): use one buffer per mesh beyond 4 GiB")
    for n in sizes:
        n = int(n)
        accs.append(TriangleAccel(ctx, buffers=(vbuf, ibuf, 3 * n, n), offsets=(36 * t0, 12 * t0),
                                  owns_buffers=False))
        t0 += n
    return accs


def _h(b):
    return b.h if b is not None else None


class LBVH:
    """gprtPointsLBVHCreate / gprtEdgesLBVHCreate / gprtTriangleLBVHCreate + gprtLBVHBuild: the
    development LBVH over float3 vertices and (for edges / triangles) uint2 / uint3 indices."""

    def __init__(self, ctx: Context, verts: np.ndarray, prims: np.ndarray | None = None, kind: str = "triangles"):
        self.ctx, self.lib = ctx, ctx.lib
        verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        self.vbuf = ctx.device_buffer(verts)
        self.pbuf = None
        if kind == "points":
            self.h = self.lib.gprtPointsLBVHCreate(ctx.h, self.vbuf.h, len(verts) if prims is None else int(prims))
        else:
            width = 2 if kind == "edges" else 3
            p = np.ascontiguousarray(prims, dtype=np.uint32).reshape(-1, width)
            self.pbuf = ctx.device_buffer(p)
            fn = self.lib.gprtEdgesLBVHCreate if kind == "edges" else self.lib.gprtTriangleLBVHCreate
            self.h = fn(ctx.h, self.vbuf.h, self.pbuf.h, len(p))
        if not self.h.impl:
            raise RuntimeError("LBVH create failed: %s" % last_error())
        self.n = int(self.h.numPrims)

    def build(self):
        self.lib.gprtLBVHBuild(self.ctx.h, C.byref(self.h))

    def read(self):
        """(mortonCodes u32[n], ids u32[n], nodes int32[2n-1, 4], aabbs float32[2n-1, 6])."""
        n = self.n
        out = []
        for h, dt, cnt in ((self.h.mortonCodes, np.uint32, n), (self.h.ids, np.uint32, n),
                           (self.h.nodes, np.int32, 4 * (2 * n - 1)), (self.h.aabbs, np.float32, 6 * (2 * n - 1))):
            out.append(Buffer(self.ctx, h).read(dt, cnt))
        return out[0], out[1], out[2].reshape(-1, 4), out[3].reshape(-1, 6)

    def destroy(self):
        if self.h.impl:
            self.lib.gprtLBVHDestroy(C.byref(self.h))
        self.vbuf.destroy()
        if self.pbuf is not None:
            self.pbuf.destroy()


def accel_bounds(ctx: Context, accel) -> np.ndarray:
    """gprtAccelGetBounds of any built accel: float32[6] {min.xyz, max.xyz}."""
    out = np.zeros(6, dtype=np.float32)
    ctx.lib.gprtAccelGetBounds(ctx.h, accel.h, out.ctypes.data_as(C.c_void_p))
    return out


class InstanceAccel:
    """gprtInstanceAccelCreate over a device buffer of gprt::Instance records."""

    def __init__(self, ctx: Context, records: np.ndarray):
        self.ctx, self.lib = ctx, ctx.lib
        self.records = np.ascontiguousarray(records, dtype=INSTANCE_DTYPE)
        self.rbuf = ctx.device_buffer(self.records)
        self.h = self.lib.gprtInstanceAccelCreate(ctx.h, len(self.records), self.rbuf.h)
        if not self.h:
            raise RuntimeError("gprtInstanceAccelCreate failed: %s" % last_error())

    def set_records(self, records: np.ndarray):
        self.records = np.ascontiguousarray(records, dtype=INSTANCE_DTYPE)
        self.rbuf.write(self.records)

    def build(self, mode=GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE):
        self.lib.gprtAccelBuild(self.ctx.h, self.h, GPRTBuildParams(mode))

    def trace(self, rays_buf: Buffer, hits_buf: Buffer, count: int):
        self.lib.gprtAccelTraceClosestHit(self.ctx.h, self.h, rays_buf.h, hits_buf.h, count)

    def destroy(self):
        if self.h:
            self.lib.gprtAccelDestroy(self.h)
            self.rbuf.destroy()
            self.h = None


def trace_detail(ctx: Context, accel, rays_buf: Buffer, hits_buf: Buffer, count: int, inst_buf: Buffer = None,
                 leaf_buf: Buffer = None, visits_buf: Buffer = None, proposal=None):
    """gprtAccelTraceClosestHitDetail (any accel; optional instance / top-leaf / visit outputs)."""
    ctx.lib.gprtAccelTraceClosestHitDetail(ctx.h, accel.h, _h(proposal), rays_buf.h, hits_buf.h, count,
                                           _h(inst_buf), _h(leaf_buf), _h(visits_buf))


def bounce_rays(ctx: Context, accel, rays_buf: Buffer, hits_buf: Buffer, inst_buf: Buffer | None, count: int,
                out_buf: Buffer, out_count: int, seed: int, tmin: float = 1e-4):
    """gprtSynthBounceRays: cosine-weighted bounce rays from the hits of `count` traced rays."""
    ctx.lib.gprtSynthBounceRays(ctx.h, accel.h, rays_buf.h, hits_buf.h, _h(inst_buf), count, out_buf.h, out_count,
                                seed, tmin)


class Proposal:
    """gprtProposal*: the learned treelet proposal bound to one accel."""

    def __init__(self, ctx: Context, max_batch: int = 8192, seed: int = 1):
        self.ctx, self.lib = ctx, ctx.lib
        self.h = self.lib.gprtProposalCreate(ctx.h, max_batch, seed)
        if not self.h:
            raise RuntimeError("gprtProposalCreate failed: %s" % last_error())
        self.n_params = int(self.lib.gprtProposalParamCount())

    def bind(self, accel: "TriangleAccel", cut_depth: int = 12) -> int:
        return int(self.lib.gprtProposalBindAccel(self.ctx.h, self.h, accel.h, cut_depth))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        self.lib.gprtProposalGetParams(self.ctx.h, self.h, out.ctypes.data_as(C.c_void_p), self.n_params)
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.lib.gprtProposalSetParams(self.ctx.h, self.h, p.ctypes.data_as(C.c_void_p), len(p))

    def train(self, rays_buf: Buffer, hits_buf: Buffer, count: int, epochs: int = 1, lr: float = 1e-3):
        acc = C.c_float(0.0)
        loss = self.lib.gprtProposalTrain(self.ctx.h, self.h, rays_buf.h, hits_buf.h, count, epochs, lr,
                                          C.byref(acc))
        return float(loss), float(acc.value)

    def train_labels(self, rays_buf: Buffer, labels_buf: Buffer, count: int, epochs: int = 1, lr: float = 1e-3):
        acc = C.c_float(0.0)
        loss = self.lib.gprtProposalTrainLabels(self.ctx.h, self.h, rays_buf.h, labels_buf.h, count, epochs, lr,
                                                C.byref(acc))
        return float(loss), float(acc.value)

    def serialize(self, with_optimizer_state: bool = True) -> bytes:
        """gprtProposalSerialize: the versioned weight checkpoint."""
        n = int(self.lib.gprtProposalSerialize(self.ctx.h, self.h, None, 0, with_optimizer_state))
        buf = C.create_string_buffer(n)
        got = int(self.lib.gprtProposalSerialize(self.ctx.h, self.h, buf, n, with_optimizer_state))
        if got != n:
            raise RuntimeError("gprtProposalSerialize failed: %s" % last_error())
        return buf.raw

    def deserialize(self, blob: bytes) -> bool:
        return bool(self.lib.gprtProposalDeserialize(self.ctx.h, self.h, blob, len(blob)))

    def destroy(self):
        if self.h:
            self.lib.gprtProposalDestroy(self.ctx.h, self.h)
            self.h = None


# ------------------------------------------------------------------------------------------------
# User compute programs
# ------------------------------------------------------------------------------------------------
PUSH_CONSTANTS_LIMIT = 256


def pack_args(*args) -> bytes:
    """Packs ctypes values as a kernel's parameter list: each at its natural alignment (the HIP
    kernel-argument layout, and the layout gprtPackArgs of the C++ header writes)."""
    out = bytearray()
    for a in args:
        al = C.alignment(a)
        out += b"\0" * (-len(out) % al)
        out += bytes(a)
    return bytes(out)


class Module:
    """gprtModuleCreate over a gfx950 code object file (hipcc --genco)."""

    def __init__(self, ctx: Context, path_or_bytes):
        self.ctx, self.lib = ctx, ctx.lib
        if isinstance(path_or_bytes, (bytes, bytearray)):
            self.code = bytes(path_or_bytes)
        else:
            with open(path_or_bytes, "rb") as f:
                self.code = f.read()
        self.h = self.lib.gprtModuleCreate(ctx.h, self.code, len(self.code))
        if not self.h:
            raise RuntimeError("gprtModuleCreate failed: %s" % last_error())

    def compute(self, name: str) -> "Compute":
        return Compute(self, name)

    def destroy(self):
        if self.h:
            self.lib.gprtModuleDestroy(self.h)
            self.h = None


class Compute:
    """gprtComputeCreate / gprtComputeLaunch: one kernel of a module."""

    def __init__(self, module: Module, name: str):
        self.ctx, self.lib = module.ctx, module.lib
        self.h = self.lib.gprtComputeCreate(self.ctx.h, module.h, name.encode())
        if not self.h:
            raise RuntimeError("gprtComputeCreate failed: %s" % last_error())

    def launch(self, groups, threads, *args):
        g = (C.c_uint32 * 3)(*(list(groups) + [1] * (3 - len(groups))))
        t = (C.c_uint32 * 3)(*(list(threads) + [1] * (3 - len(threads))))
        blob = pack_args(*args)
        return self.lib.gprtComputeLaunchArgs(self.h, g, t, blob, len(blob))

    def destroy(self):
        if self.h:
            self.lib.gprtComputeDestroy(self.h)
            self.h = None
