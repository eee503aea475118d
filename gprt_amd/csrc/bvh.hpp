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

namespace gprt {

struct __attribute__((aligned(16))) BVHNode {
  float4 a, b, c;
  int4 d;
};
static_assert(sizeof(BVHNode) == 64, "BVHNode must be 64 bytes");

struct __attribute__((aligned(16))) TriRecord {
  float4 v0, v1, v2;
};
static_assert(sizeof(TriRecord) == 48, "TriRecord must be 48 bytes");

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

// Exact min/max under the total order of float bit patterns (-0 < +0), so results do not depend
// on the order in which children or primitives are combined.
__host__ __device__ __forceinline__ uint32_t ford(float f) {
  union {
    float f;
    uint32_t u;
  } x;
  x.f = f;
  return (x.u & 0x80000000u) ? ~x.u : (x.u | 0x80000000u);
}
__host__ __device__ __forceinline__ float fmin_ord(float a, float b) { return ford(b) < ford(a) ? b : a; }
__host__ __device__ __forceinline__ float fmax_ord(float a, float b) { return ford(b) > ford(a) ? b : a; }

// ------------------------------------------------------------------------------------------------
// Ray setup shared by the box and triangle tests.
// ------------------------------------------------------------------------------------------------
struct RayPre {
  float ox, oy, oz, tmin;
  float ix, iy, iz;   // reciprocal direction, |d| clamped away from 0 so no 0*inf NaN arises
  int kx, ky, kz;     // watertight axis permutation
  float sx, sy, sz;   // watertight shear constants
  float dx, dy, dz;
};

__device__ __forceinline__ float clamp_dir(float d) {
  return fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d;
}

__device__ __forceinline__ float comp(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ RayPre ray_setup(float ox, float oy, float oz, float tmin, float dx, float dy, float dz) {
  RayPre r;
  r.ox = ox;
  r.oy = oy;
  r.oz = oz;
  r.tmin = tmin;
  r.dx = dx;
  r.dy = dy;
  r.dz = dz;
  r.ix = 1.0f / clamp_dir(dx);
  r.iy = 1.0f / clamp_dir(dy);
  r.iz = 1.0f / clamp_dir(dz);
  const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
  int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  const float dkz = comp(dx, dy, dz, kz);
  if (dkz < 0.0f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.sx = comp(dx, dy, dz, kx) / dkz;
  r.sy = comp(dx, dy, dz, ky) / dkz;
  r.sz = 1.0f / dkz;
  return r;
}

// Conservative slab test. The far distance is widened by 1+2*gamma(3) (Ize 2013) and the near
// distance is compared against tbest widened by a relative slack that covers the rounding of
// the triangle test, so a primitive the triangle test accepts is never culled by its box.
constexpr float kFarScale = 1.0f + 2.0f * (3.0f * 5.9604645e-08f) / (1.0f - 3.0f * 5.9604645e-08f);
constexpr float kNearSlack = 1.0f / 16384.0f;

__device__ __forceinline__ bool box_hit(const RayPre &r, float lox, float loy, float loz, float hix, float hiy,
                                        float hiz, float tbest, float &tnear) {
  const float t0x = (lox - r.ox) * r.ix, t1x = (hix - r.ox) * r.ix;
  const float t0y = (loy - r.oy) * r.iy, t1y = (hiy - r.oy) * r.iy;
  const float t0z = (loz - r.oz) * r.iz, t1z = (hiz - r.oz) * r.iz;
  const float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
  const float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  const float slack = kNearSlack * (fabsf(tn) + fabsf(tf));
  tnear = tn;
  const float lo = fmaxf(tn - slack, r.tmin);
  const float hi = fminf(tf * kFarScale + slack, tbest);
  return lo <= hi;
}

// Box primitive (AABB geometry) as a closed surface: the entry distance when the ray enters the
// box after tmin, else the exit distance when it starts inside; exact FP32 slab test in a fixed
// operation order, with the total-order min/max (so the sign of a zero is deterministic).
__device__ __forceinline__ bool boxprim_hit(const RayPre &r, const float4 &lo, const float4 &hi, float &t) {
  const float t0x = (lo.x - r.ox) * r.ix, t1x = (hi.x - r.ox) * r.ix;
  const float t0y = (lo.y - r.oy) * r.iy, t1y = (hi.y - r.oy) * r.iy;
  const float t0z = (lo.z - r.oz) * r.iz, t1z = (hi.z - r.oz) * r.iz;
  const float tn = fmax_ord(fmax_ord(fmin_ord(t0x, t1x), fmin_ord(t0y, t1y)), fmin_ord(t0z, t1z));
  const float tf = fmin_ord(fmin_ord(fmax_ord(t0x, t1x), fmax_ord(t0y, t1y)), fmax_ord(t0z, t1z));
  if (!(tn <= tf)) return false;
  if (tn > r.tmin) {
    t = tn;
    return true;
  }
  if (tf > r.tmin) {
    t = tf;
    return true;
  }
  return false;
}

// Watertight ray/triangle intersection (Woop, Benthin, Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013), FP32 with a fixed operation order and an FP64 re-evaluation
// of the 2D edge functions when one of them is exactly zero. Returns true and (t, u, v) when the
// ray meets the triangle with tmin < t; the caller applies the closest-hit rule.
__device__ __forceinline__ bool tri_hit(const RayPre &r, const float4 &p0, const float4 &p1, const float4 &p2,
                                        float &t, float &u, float &v) {
  const float Ax0 = p0.x - r.ox, Ay0 = p0.y - r.oy, Az0 = p0.z - r.oz;
  const float Bx0 = p1.x - r.ox, By0 = p1.y - r.oy, Bz0 = p1.z - r.oz;
  const float Cx0 = p2.x - r.ox, Cy0 = p2.y - r.oy, Cz0 = p2.z - r.oz;
  const float Akx = comp(Ax0, Ay0, Az0, r.kx), Aky = comp(Ax0, Ay0, Az0, r.ky), Akz = comp(Ax0, Ay0, Az0, r.kz);
  const float Bkx = comp(Bx0, By0, Bz0, r.kx), Bky = comp(Bx0, By0, Bz0, r.ky), Bkz = comp(Bx0, By0, Bz0, r.kz);
  const float Ckx = comp(Cx0, Cy0, Cz0, r.kx), Cky = comp(Cx0, Cy0, Cz0, r.ky), Ckz = comp(Cx0, Cy0, Cz0, r.kz);
  const float Ax = Akx - r.sx * Akz, Ay = Aky - r.sy * Akz;
  const float Bx = Bkx - r.sx * Bkz, By = Bky - r.sy * Bkz;
  const float Cx = Ckx - r.sx * Ckz, Cy = Cky - r.sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  t = T / det;
  if (!(t > r.tmin)) return false;
  u = V / det;
  v = W / det;
  return true;
}

}  // namespace gprt
