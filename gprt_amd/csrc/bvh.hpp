// HBM layout of the triangle LBVH and the device-side geometric tests.
//
// Node numbering follows the reference's development LBVH (gprt/development/gprt_lbvh.h:41-51):
// inner nodes [0, N-2] with the root at 0, leaf i (the i-th primitive in sorted key order) at
// N-1+i. For traversal each inner node is stored as one 64-byte record holding both children's
// boxes and indices, so a single node fetch (4 x 16 B, half an L2 line) decides both children:
//
//   float[0..5]  child 0 box {lo.xyz, hi.xyz}
//   float[6..11] child 1 box {lo.xyz, hi.xyz}
//   int[12], int[13]  child 0 / child 1: inner index >= 0, or ~leafPosition (< 0) for a leaf
//   int[14..15]  padding
//
// Triangles are copied into leaf order (48 B each): {v0.xyz, primID}, {v1.xyz, 0}, {v2.xyz, 0},
// so a leaf test reads one contiguous record instead of gathering index + 3 vertices.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gprt_hip_device.h"

namespace gprt {

// The node and leaf-record layouts, the ray setup and the box / triangle tests are those of the
// public device header (device programs trace with the same code).
using BVHNode = gprt_device::BVHNode;
using TriRecord = gprt_device::LeafRecord;
using gprt_device::box_hit;
using gprt_device::boxprim_hit;
using gprt_device::clamp_dir;
using gprt_device::comp;
using gprt_device::fmax_ord;
using gprt_device::fmin_ord;
using gprt_device::ford;
using gprt_device::kBoxSlack;
using gprt_device::ray_setup;
using gprt_device::RayPre;
using gprt_device::tri_hit;

// Rays and hits are read and written once per trace: streaming (non-temporal) accesses keep them
// from displacing the tree, which is reused by every ray, from the caches.
typedef float stream_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ GPRTRay load_ray_stream(const GPRTRay *p) {
  const stream_f4 a = __builtin_nontemporal_load((const stream_f4 *)p);
  const stream_f4 b = __builtin_nontemporal_load((const stream_f4 *)p + 1);
  return GPRTRay{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}
__device__ __forceinline__ void store_hit_stream(GPRTHit *p, const GPRTHit &h) {
  const stream_f4 v = {h.t, __int_as_float(h.primID), h.u, h.v};
  __builtin_nontemporal_store(v, (stream_f4 *)p);
}

// Header words (float): [0..5] scene box as ordered uints during the build, then floats;
// [8..13] root box; [15] reserved.
constexpr int kHeaderFloats = 16;

// ------------------------------------------------------------------------------------------------
// Morton code: 10 bits per axis interleaved x | y<<1 | z<<2 (gprt/development/gprt_lbvh.hlsl:25-38).
// ------------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__host__ __device__ __forceinline__ uint32_t morton30(uint32_t x, uint32_t y, uint32_t z) {
  return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
}

}  // namespace gprt
