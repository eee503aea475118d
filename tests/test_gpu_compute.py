"""User compute programs (gprtModuleCreate / gprtComputeCreate / gprtComputeLaunch) over GPRT
buffers: a gfx950 code object built from tests/programs/compute_programs.hip is loaded at run
time and its kernels are launched on the context's stream.

The first test follows the reference's compute-bounds sample
(samples/s5-computePrograms/s5-0-computeAABBs/hostCode.cpp:54-95): a compute program writes
sphere boxes into the positions buffer of an AABB geometry, the AABB accel is built from them, a
second program writes the rays, and the trace is checked against the oracle's brute force over
the same boxes. The others check the argument layout, the grid and group shape, the full
256-byte argument block, ordering on the stream, the kernel timer and the error paths."""
import ctypes as C
import os

import numpy as np
import pytest

from gprt_amd import api, scenes

pytestmark = pytest.mark.gpu

CODE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "programs", "compute_programs.co")


class AABBGeomData(C.Structure):  # samples/s5-computePrograms/s5-0-computeAABBs/sharedCode.h:27-34
    _fields_ = [("vertex", C.c_void_p), ("radius", C.c_void_p), ("aabbs", C.c_void_p)]


def _dptr(buf):
    return C.c_void_p(buf.lib.gprtBufferGetDevicePointer(buf.h, 0))


@pytest.fixture(scope="module")
def module(ctx):
    m = api.Module(ctx, CODE)
    yield m
    m.destroy()


def test_compute_bounds_then_aabb_accel_trace(ctx, orc, module):
    n = 20000
    rng = np.random.default_rng(3)
    centres = rng.random((n, 3), dtype=np.float32)
    radii = (rng.random(n, dtype=np.float32) * np.float32(0.01) + np.float32(1e-3)).astype(np.float32)
    vb, rb = ctx.device_buffer(centres), ctx.device_buffer(radii)
    acc = api.AABBAccel(ctx, np.zeros((n, 6), np.float32))  # positions buffer the program fills
    bounds = module.compute("ComputeBounds")
    bounds.launch(((n + 63) // 64,), (64,), AABBGeomData(_dptr(vb), _dptr(rb), _dptr(acc.vbuf)), C.c_uint32(n))
    acc.build()  # ordered after the launch on the context's stream
    assert api.last_error() is None, api.last_error()
    boxes = np.concatenate([centres - radii[:, None], centres + radii[:, None]], 1).astype(np.float32)
    np.testing.assert_array_equal(acc.vbuf.read(np.float32).reshape(-1, 6).view(np.uint32), boxes.view(np.uint32))
    k, nodes, aabbs = acc.export()
    rk, rn, ra = orc.lbvh_build_boxes(boxes)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(nodes, rn)
    np.testing.assert_array_equal(aabbs.view(np.uint32), ra.view(np.uint32))

    w, h = 96, 64
    rays_buf = ctx.device_buffer(nbytes=32 * w * h)
    ortho = module.compute("OrthoRays")
    ortho.launch(((w + 15) // 16, (h + 15) // 16), (16, 16), C.c_void_p(_dptr(rays_buf).value), C.c_uint32(w),
                 C.c_uint32(h), C.c_float(-1.0))
    hits_buf = ctx.device_buffer(nbytes=16 * w * h)
    acc.trace(rays_buf, hits_buf, w * h)
    rays = rays_buf.read(scenes.RAY_DTYPE)
    y, x = np.divmod(np.arange(w * h), w)
    np.testing.assert_array_equal(rays["ox"], ((x.astype(np.float32) + np.float32(0.5)) / np.float32(w)))
    np.testing.assert_array_equal(rays["oy"], ((y.astype(np.float32) + np.float32(0.5)) / np.float32(h)))
    assert np.all(rays["oz"] == -1) and np.all(rays["dz"] == 1) and np.all(rays["dx"] == 0)
    hits = hits_buf.read(api.HIT_DTYPE)
    ref = orc.trace_boxes_bruteforce(rays, boxes)
    np.testing.assert_array_equal(hits["primID"], ref["primID"])
    np.testing.assert_array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))
    assert (hits["primID"] >= 0).mean() > 0.5
    for b in (vb, rb, rays_buf, hits_buf):
        b.destroy()
    bounds.destroy()
    ortho.destroy()
    acc.destroy()


def test_mixed_arguments_sit_at_natural_alignment(ctx, module):
    n = 1000
    out = ctx.device_buffer(nbytes=8 * n)
    prog = module.compute("Mixed")
    prog.launch(((n + 255) // 256,), (256,), C.c_uint32(n), _dptr(out), C.c_uint16(7), C.c_double(1.75),
                C.c_uint8(5))
    got = out.read(np.uint64)
    want = (np.arange(n, dtype=np.float64) * 1.75).astype(np.uint64) + np.uint64(7) + (np.uint64(5) << np.uint64(40))
    np.testing.assert_array_equal(got, want)
    prog.destroy()
    out.destroy()


def test_grid_and_group_shape(ctx, module):
    groups, threads = (3, 2, 4), (8, 4, 2)
    slots = int(np.prod(groups) * np.prod(threads))
    out = ctx.device_buffer(nbytes=8 * slots)
    prog = module.compute("Grid3D")
    prog.launch(groups, threads, _dptr(out))
    got = out.read(np.uint32).reshape(-1, 2)
    gz, gy, gx, tz, ty, tx = np.meshgrid(*[np.arange(v) for v in (groups[2], groups[1], groups[0], threads[2],
                                                                    threads[1], threads[0])], indexing="ij")
    want_g = (gx | (gy << 16) | (gz << 24)).ravel().astype(np.uint32)
    want_t = (tx | (ty << 10) | (tz << 20)).ravel().astype(np.uint32)
    np.testing.assert_array_equal(got[:, 0], want_g)
    np.testing.assert_array_equal(got[:, 1], want_t)
    prog.destroy()
    out.destroy()


def test_full_argument_block(ctx, module):
    class Full(C.Structure):
        _fields_ = [("out", C.c_void_p), ("words", C.c_uint32 * 62)]

    assert C.sizeof(Full) == api.PUSH_CONSTANTS_LIMIT
    out = ctx.device_buffer(nbytes=4 * 62)
    words = np.arange(62, dtype=np.uint32) * np.uint32(2654435761)
    prog = module.compute("Echo")
    prog.launch((1,), (64,), Full(_dptr(out).value, (C.c_uint32 * 62)(*words.tolist())))
    np.testing.assert_array_equal(out.read(np.uint32), words)
    prog.destroy()
    out.destroy()


def test_launches_are_ordered_and_timed(ctx, module):
    """Launches run in order with the library's own work on the context's stream: a program
    overwrites the first half of random keys, the sort runs, a second launch overwrites the
    first quarter of the sorted keys."""
    n = 1 << 16
    rng = np.random.default_rng(9)
    start = rng.integers(0, 1 << 40, n, dtype=np.uint64)
    keys = ctx.device_buffer(start)
    prog = module.compute("Mixed")
    lib = ctx.lib
    lib.gprtKernelTimerEnable(ctx.h, True)
    lib.gprtKernelTimerReset(ctx.h)
    # out[i] = floor(i * scale) + add + (flag << 40), i < count
    prog.launch(((n // 2 + 255) // 256,), (256,), C.c_uint32(n // 2), _dptr(keys), C.c_uint16(3), C.c_double(7.25),
                C.c_uint8(2))
    lib.gprtBufferSort(ctx.h, keys.h, None)
    prog.launch(((n // 4 + 255) // 256,), (256,), C.c_uint32(n // 4), _dptr(keys), C.c_uint16(0), C.c_double(1.0),
                C.c_uint8(0))
    got = keys.read(np.uint64)
    first = (np.arange(n // 2, dtype=np.float64) * 7.25).astype(np.uint64) + np.uint64(3) + (np.uint64(2) << np.uint64(40))
    want = np.sort(np.concatenate([first, start[n // 2:]]))
    want[:n // 4] = np.arange(n // 4, dtype=np.uint64)
    np.testing.assert_array_equal(got, want)
    launches = C.c_uint32(0)
    ms = lib.gprtKernelTimerQuery(ctx.h, b"Mixed", C.byref(launches))
    assert launches.value == 2 and ms > 0
    lib.gprtKernelTimerEnable(ctx.h, False)
    prog.destroy()
    keys.destroy()


def test_errors(ctx, module):
    lib = ctx.lib
    lib.gprtClearLastError()
    assert not lib.gprtComputeCreate(ctx.h, module.h, b"NoSuchProgram")
    assert "NoSuchProgram" in api.last_error()
    lib.gprtClearLastError()
    assert not lib.gprtModuleCreate(ctx.h, b"\0" * 128, 128)
    assert "neither a code object" in api.last_error()
    lib.gprtClearLastError()
    # sized for every rejected launch below (65536 x 1 and 1 x 2048 slots), should one run anyway
    out = ctx.device_buffer(nbytes=8 * 65536)
    prog = module.compute("Grid3D")
    prog.launch((65536,), (1,), _dptr(out))
    assert "WORKGROUP_LIMIT" in api.last_error()
    lib.gprtClearLastError()
    prog.launch((1,), (32, 32, 2), _dptr(out))
    assert "THREADGROUP_LIMIT" in api.last_error()
    lib.gprtClearLastError()
    g = (C.c_uint32 * 3)(1, 1, 1)
    lib.gprtComputeLaunchArgs(prog.h, g, g, b"\0" * 300, 300)
    assert "PUSH_CONSTANTS_LIMIT" in api.last_error()
    lib.gprtClearLastError()
    # a zero-sized dispatch runs nothing and is not an error
    out.write(np.full(2 * 65536, 0xAB, np.uint32))
    prog.launch((0, 1, 1), (1,), _dptr(out))
    assert api.last_error() is None
    assert np.all(out.read(np.uint32) == 0xAB)
    # a handle outlives its module's destruction and then reports it
    m2 = api.Module(ctx, CODE)
    p2 = m2.compute("Grid3D")
    m2.destroy()
    p2.launch((1,), (1,), _dptr(out))
    assert "was destroyed" in api.last_error()
    lib.gprtClearLastError()
    p2.destroy()
    prog.destroy()
    out.destroy()


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Camera(C.Structure):  # samples/s5-computePrograms/s5-0-computeAABBs/sharedCode.h (camera)
    _fields_ = [("pos", Float3), ("dir_00", Float3), ("dir_du", Float3), ("dir_dv", Float3)]


class RayGenData(C.Structure):
    _fields_ = [("world", C.c_void_p), ("frameBuffer", C.c_void_p), ("rays", C.c_void_p)]


def _bits(h):
    return h.view(np.uint32).reshape(-1, 4)


def test_program_traces_aabb_accel_like_the_library(ctx, orc, module):
    """The reference sample's whole flow inside programs: ComputeBounds writes the sphere boxes,
    the AABB accel is built from them, and the sample's ray-generation program traces its camera
    rays against the accel it names by gprtAccelGetDeviceAddress (include/gprt_hip_device.h).
    Hits equal gprtAccelTraceClosestHit on the same rays bit for bit, and the oracle's brute
    force over the boxes."""
    n = 5000
    rng = np.random.default_rng(7)
    centres = rng.random((n, 3), dtype=np.float32)
    radii = (rng.random(n, dtype=np.float32) * np.float32(0.02) + np.float32(2e-3)).astype(np.float32)
    vb, rb = ctx.device_buffer(centres), ctx.device_buffer(radii)
    acc = api.AABBAccel(ctx, np.zeros((n, 6), np.float32))
    bounds = module.compute("ComputeBounds")
    bounds.launch(((n + 63) // 64,), (64,), AABBGeomData(_dptr(vb), _dptr(rb), _dptr(acc.vbuf)), C.c_uint32(n))
    acc.build()
    world = ctx.lib.gprtAccelGetDeviceAddress(acc.h, 0)
    w, h = 80, 60
    fb, rays_buf = ctx.device_buffer(nbytes=16 * w * h), ctx.device_buffer(nbytes=32 * w * h)
    cam = Camera(Float3(0.5, 0.5, -1.5), Float3(-0.4, -0.3, 1.0), Float3(0.8, 0.0, 0.0), Float3(0.0, 0.6, 0.0))
    raygen = module.compute("SimpleRayGen")
    print(f"device trace: {w}x{h} camera rays against {n} boxes", flush=True)
    raygen.launch(((w + 15) // 16, (h + 15) // 16), (16, 16), RayGenData(world, _dptr(fb), _dptr(rays_buf)), cam,
                  C.c_uint32(w), C.c_uint32(h))
    assert api.last_error() is None, api.last_error()
    got = fb.read(api.HIT_DTYPE)
    lib_hits = ctx.device_buffer(nbytes=16 * w * h)
    acc.trace(rays_buf, lib_hits, w * h)
    want = lib_hits.read(api.HIT_DTYPE)
    np.testing.assert_array_equal(_bits(got), _bits(want))
    boxes = np.concatenate([centres - radii[:, None], centres + radii[:, None]], 1).astype(np.float32)
    ref = orc.trace_boxes_bruteforce(rays_buf.read(scenes.RAY_DTYPE), boxes)
    np.testing.assert_array_equal(got["primID"], ref["primID"])
    np.testing.assert_array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32))
    assert (got["primID"] >= 0).mean() > 0.2
    for b in (vb, rb, fb, rays_buf, lib_hits):
        b.destroy()
    bounds.destroy()
    raygen.destroy()
    acc.destroy()


def test_program_traces_triangle_accel_like_the_library(ctx, orc, module):
    """A ray batch traced inside a program against a triangle accel equals the library's trace
    bit for bit (t, primID, u, v); an accel not built yet, a destroyed accel and a null or
    misaligned address give misses."""
    verts, idx = scenes.blob_scene(200_000, seed=12)
    acc = api.TriangleAccel(ctx, verts, idx)
    rays = np.concatenate([scenes.pinhole_rays(128, 128), scenes.random_rays(30_000, 5)])
    n = len(rays)
    rb = ctx.device_buffer(rays)
    hb, hb2 = ctx.device_buffer(nbytes=16 * n), ctx.device_buffer(nbytes=16 * n)
    trace = module.compute("TraceRays")
    addr = ctx.lib.gprtAccelGetDeviceAddress(acc.h, 0)

    def program_trace(address):
        trace.launch(((n + 255) // 256,), (256,), C.c_void_p(address), _dptr(rb), _dptr(hb), C.c_uint32(n))
        return hb.read(api.HIT_DTYPE)

    def all_miss(hits):
        return (hits["primID"] == -1).all() and np.array_equal(hits["t"], rays["tmax"]) and (hits["u"] == 0).all()

    assert all_miss(program_trace(addr))  # not built yet
    acc.build()
    print(f"device trace: {n} rays against {len(idx)} triangles", flush=True)
    got = program_trace(addr)
    acc.trace(rb, hb2, n)
    want = hb2.read(api.HIT_DTYPE)
    assert api.last_error() is None, api.last_error()
    np.testing.assert_array_equal(_bits(got), _bits(want))
    assert (got["primID"] >= 0).mean() > 0.02
    sel = np.random.default_rng(1).choice(n, 256, replace=False)
    ref = orc.trace_bruteforce(rays[sel], verts, idx)
    np.testing.assert_array_equal(_bits(got[sel]), _bits(ref))
    assert all_miss(program_trace(0)) and all_miss(program_trace(addr + 16))
    acc.destroy()
    assert all_miss(program_trace(addr))  # the slot outlives the accel, cleared
    for b in (rb, hb, hb2):
        b.destroy()
    trace.destroy()
