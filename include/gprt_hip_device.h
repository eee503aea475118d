/*
 * gprt_hip_device.h — closest-hit traversal callable from device code.
 *
 * In the reference, device programs trace rays against an accel they name by address
 * (TraceRay on a SurfaceAccelerationStructure built from gprtAccelGetDeviceAddress,
 * gprt/gprt_builtins.slang:105-124; gprt/gprt_host.h:1000-1002; used by
 * samples/s5-computePrograms/s5-0-computeAABBs/deviceCode.slang:45). Here a program compiled for
 * gfx950 (`hipcc --genco --offload-arch=gfx950`, loaded with gprtModuleCreate and launched with
 * gprtComputeLaunch) includes this header and calls
 *
 *     GPRTHit gprtTraceClosestHit(const void *accel, GPRTRay ray);
 *
 * with `accel` = gprtAccelGetDeviceAddress of a built triangle or AABB accel, passed in as a
 * kernel argument. The result is the one gprtAccelTraceClosestHit returns for the same ray, bit
 * for bit: the lexicographic minimum of (t, primID) over every primitive the test accepts with
 * tmin < t < tmax, t = tmax and primID = -1 on a miss. The address is a 64-byte descriptor slot
 * that the library keeps for the context's lifetime, so the address of a destroyed or not (yet)
 * built accel stays readable and returns a miss (its magic word is cleared), as does a null or
 * misaligned address; instance accels are not traced here (their address is not a descriptor).
 *
 * The ray and box tests below are the ones the library's trace kernels use (the library includes
 * this header), with floating-point contraction switched off inside them, so results do not
 * depend on the caller's -ffp-contract; do not build callers with fast-math options that change
 * division or comparisons (-ffast-math, -fno-honor-nans).
 *
 * The walk keeps its stack in private memory (64 entries; a binary radix tree over distinct
 * 64-bit keys is at most 64 levels deep). The node and leaf-record layouts are those of the
 * built accel in HBM; they are described here so that device code can read them.
 */
#ifndef GPRT_HIP_DEVICE_H
#define GPRT_HIP_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gprt_hip.h"

namespace gprt_device {

/* Inner node, 64 B: child 0 box {lo.xyz, hi.xyz} (a.xyzw b.xy), child 1 box (b.zw c.xyzw), child
 * references d.x / d.y (an inner node index >= 0, or ~leafPosition < 0). Inner nodes [0, n-2],
 * the root at 0 (the numbering of gprt/development/gprt_lbvh.h:41-51). */
struct __attribute__((aligned(16))) BVHNode {
  float4 a, b, c;
  int4 d;
};
static_assert(sizeof(BVHNode) == 64, "BVHNode must be 64 bytes");

/* Leaf record in leaf order, 48 B. Triangles: {v0.xyz, primID} {v1.xyz, -} {v2.xyz, -}; boxes:
 * {lo.xyz, primID} {hi.xyz, -} {-} with the corners as given. */
struct __attribute__((aligned(16))) LeafRecord {
  float4 v0, v1, v2;
};
static_assert(sizeof(LeafRecord) == 48, "LeafRecord must be 48 bytes");

/* The 64-byte descriptor at gprtAccelGetDeviceAddress of a triangle or AABB accel. */
struct AccelDescriptor {
  const void *nodes;  /* BVHNode[n - 1] */
  const void *tris;   /* LeafRecord[n] */
  const float *header;
  const char *verts;
  const char *idx;
  uint32_t n, vstride, istride, nverts;
  uint32_t magic; /* kTriangleAccelMagic / kBoxAccelMagic while the accel is built and alive */
  uint32_t slot;
};
static_assert(sizeof(AccelDescriptor) == 64, "AccelDescriptor must be 64 bytes");
constexpr uint32_t kTriangleAccelMagic = 0x53414C42u;
constexpr uint32_t kBoxAccelMagic = 0x58424C42u;

/* Ray prepared for the box and triangle tests. */
struct RayPre {
  float ox, oy, oz, tmin;
  float ix, iy, iz; /* reciprocal direction, |d| clamped away from 0 so no 0*inf NaN arises */
  int kx, ky, kz;   /* watertight axis permutation */
  float sx, sy, sz; /* watertight shear constants */
  float dx, dy, dz;
};

__device__ __forceinline__ float clamp_dir(float d) { return fabsf(d) < 1e-30f ? copysignf(1e-30f, d) : d; }

__device__ __forceinline__ float comp(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ RayPre ray_setup(float ox, float oy, float oz, float tmin, float dx, float dy, float dz) {
#pragma clang fp contract(off)
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

/* Exact min / max under the total order of float bit patterns (-0 < +0). */
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

/* Conservative slab test of a node's child box. The box is accepted unless the later of its entry
 * distance and tmin exceeds the earlier of its exit distance and tbest by more than
 * kBoxSlack * (|tnear| + |tfar|). That margin (2^-13 relative) is far above the rounding of the
 * slab distances (Ize 2013 bounds it by 2*gamma(3) of the far distance) and of the triangle test,
 * so a primitive the triangle test accepts is never culled by its box; a NaN difference (from
 * overflowing distances) accepts. */
constexpr float kBoxSlack = 1.0f / 8192.0f;

__device__ __forceinline__ bool box_hit(const RayPre &r, float lox, float loy, float loz, float hix, float hiy,
                                        float hiz, float tbest, float &tnear) {
#pragma clang fp contract(off)
  const float t0x = (lox - r.ox) * r.ix, t1x = (hix - r.ox) * r.ix;
  const float t0y = (loy - r.oy) * r.iy, t1y = (hiy - r.oy) * r.iy;
  const float t0z = (loz - r.oz) * r.iz, t1z = (hiz - r.oz) * r.iz;
  const float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
  const float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  tnear = tn;
  const float slack = kBoxSlack * (fabsf(tn) + fabsf(tf));
  return !(fmaxf(tn, r.tmin) - fminf(tf, tbest) > slack);
}

/* Box primitive (AABB geometry) as a closed surface: the entry distance when the ray enters the
 * box after tmin, else the exit distance when it starts inside; exact FP32 slab test in a fixed
 * operation order with the total-order min / max (so the sign of a zero is deterministic). */
__device__ __forceinline__ bool boxprim_hit(const RayPre &r, const float4 &lo, const float4 &hi, float &t) {
#pragma clang fp contract(off)
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

/* Watertight ray/triangle intersection (Woop, Benthin, Wald, "Watertight Ray/Triangle
 * Intersection", JCGT 2(1), 2013), FP32 with a fixed operation order and an FP64 re-evaluation
 * of the 2D edge functions when one of them is exactly zero. Returns true and (t, u, v) when the
 * ray meets the triangle with tmin < t; the caller applies the closest-hit rule. */
__device__ __forceinline__ bool tri_hit(const RayPre &r, const float4 &p0, const float4 &p1, const float4 &p2,
                                        float &t, float &u, float &v) {
#pragma clang fp contract(off)
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

/* Closest hit of one ray in a triangle (kBoxes = false) or AABB accel: near-first depth-first
 * walk, both children of a node decided by one node fetch, leaves tested at once; the hit rule
 * makes the result independent of the walk order. */
template <bool kBoxes>
__device__ inline GPRTHit trace_closest(const BVHNode *__restrict__ nodes, const LeafRecord *__restrict__ recs,
                                        uint32_t n, const GPRTRay &ray) {
  const RayPre r = ray_setup(ray.ox, ray.oy, ray.oz, ray.tmin, ray.dx, ray.dy, ray.dz);
  float bt = ray.tmax, bu = 0.f, bv = 0.f;
  int bp = -1;
  auto leaf = [&](int pos) {
    if ((uint32_t)pos >= n) return;
    const LeafRecord rec = recs[pos];
    float t, u = 0.f, v = 0.f;
    const bool hit = kBoxes ? boxprim_hit(r, rec.v0, rec.v1, t) : tri_hit(r, rec.v0, rec.v1, rec.v2, t, u, v);
    if (!hit) return;
    const int prim = (int)__float_as_uint(rec.v0.w);
    if (t < bt || (t == bt && (unsigned)prim < (unsigned)bp)) {
      bt = t;
      bp = prim;
      bu = u;
      bv = v;
    }
  };
  if (n == 1) {
    leaf(0);
  } else if (n > 1) {
    int stack[64];
    int sp = 0, node = 0;
    /* Each inner node is fetched at most once; the step bound only guards a malformed tree. */
    for (uint32_t step = 0; step < n; ++step) {
      if ((uint32_t)node >= n - 1) break;
      const BVHNode nd = nodes[node];
      float tn0, tn1;
      bool h0 = box_hit(r, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, bt, tn0);
      bool h1 = box_hit(r, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, bt, tn1);
      const int c0 = nd.d.x, c1 = nd.d.y;
      if (h0 && c0 < 0) {
        leaf(~c0);
        h0 = false;
      }
      if (h1 && c1 < 0) {
        leaf(~c1);
        h1 = false;
      }
      if (h0 && h1) {
        const bool firstNear = tn0 <= tn1;
        if (sp < 64) stack[sp++] = firstNear ? c1 : c0;
        node = firstNear ? c0 : c1;
      } else if (h0) {
        node = c0;
      } else if (h1) {
        node = c1;
      } else {
        if (sp == 0) break;
        node = stack[--sp];
      }
    }
  }
  GPRTHit o;
  o.t = bp >= 0 ? bt : ray.tmax;
  o.primID = bp;
  o.u = bu;
  o.v = bv;
  return o;
}

}  // namespace gprt_device

/* Closest hit of `ray` against the triangle or AABB accel whose gprtAccelGetDeviceAddress is
 * `accel`; a miss (t = tmax, primID = -1, u = v = 0) when `accel` names no live, built triangle
 * or AABB accel. */
__device__ inline GPRTHit gprtTraceClosestHit(const void *accel, GPRTRay ray) {
  GPRTHit miss;
  miss.t = ray.tmax;
  miss.primID = -1;
  miss.u = miss.v = 0.f;
  if (!accel || ((uintptr_t)accel & 63u)) return miss;
  const gprt_device::AccelDescriptor d = *(const gprt_device::AccelDescriptor *)accel;
  if (!d.nodes && d.n != 1) return miss;
  if (d.magic == gprt_device::kTriangleAccelMagic)
    return gprt_device::trace_closest<false>((const gprt_device::BVHNode *)d.nodes,
                                             (const gprt_device::LeafRecord *)d.tris, d.n, ray);
  if (d.magic == gprt_device::kBoxAccelMagic)
    return gprt_device::trace_closest<true>((const gprt_device::BVHNode *)d.nodes,
                                            (const gprt_device::LeafRecord *)d.tris, d.n, ray);
  return miss;
}

#endif /* GPRT_HIP_DEVICE_H */
