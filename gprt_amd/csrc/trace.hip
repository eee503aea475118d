// Closest-hit traversal of the triangle LBVH for gfx950.
//
// The reference has no software traversal: closest hit comes from the Vulkan driver via
// TraceRay / vkCmdTraceRaysKHR (gprt/gprt_builtins.slang:105-124, gprt/gprt.cpp:9404); its
// conventions are kept: a ray {origin, tmin, direction, tmax} (RayDesc), an open interval
// tmin < t (gprt/gprt_fallbacks.slang:108,139) and a closest hit reporting PrimitiveIndex, t and
// barycentrics (samples/s2-hitPrograms/s2-0-triangles/deviceCode.slang:16-19).
//
// Kernel shape: persistent workgroups of 128 lanes; each wave takes 64 rays at a time from a
// global counter (load balance across rays of very different cost). One node fetch decides both
// children (64-byte node with both child boxes); leaf children are intersected at once; the near
// inner child is descended and the far one pushed on a per-lane stack whose first kLdsStack
// entries live in LDS and the rest in a per-lane global spill area. A binary radix tree over
// distinct 64-bit keys is at most 64 levels deep, so 64 entries always suffice.
#include "bvh.hpp"
#include "device_common.hpp"
#include "internal.hpp"

namespace gprt {

namespace {

constexpr int kTraceBlock = 128;
constexpr int kLdsStack = 24;
constexpr int kMaxStack = 64;

__device__ __forceinline__ void intersect_leaf(const RayPre &r, const TriRecord *__restrict__ tris, int leaf,
                                               uint32_t nprims, float &tbest, int &best, float &bu, float &bv) {
  if ((uint32_t)leaf >= nprims) return;
  const TriRecord tr = tris[leaf];
  float t, u, v;
  if (tri_hit(r, tr.v0, tr.v1, tr.v2, t, u, v)) {
    const int prim = (int)__float_as_uint(tr.v0.w);
    if (t < tbest || (t == tbest && (unsigned)prim < (unsigned)best)) {
      tbest = t;
      best = prim;
      bu = u;
      bv = v;
    }
  }
}

__global__ void __launch_bounds__(kTraceBlock)
    k_trace(const BVHNode *__restrict__ nodes, const TriRecord *__restrict__ tris, uint32_t nprims,
            const GPRTRay *__restrict__ rays, GPRTHit *__restrict__ hits, uint32_t count,
            uint32_t *__restrict__ work, uint32_t *__restrict__ spill) {
  __shared__ uint32_t s_stack[kLdsStack][kTraceBlock];
  const int tid = threadIdx.x;
  const uint32_t gtid = blockIdx.x * kTraceBlock + tid;
  uint32_t *mySpill = spill + (size_t)gtid * (kMaxStack - kLdsStack);
  const int l = tid & 63;

  while (true) {
    uint32_t base = 0;
    if (l == 0) base = atomicAdd(work, 64u);
    base = __shfl(base, 0, 64);
    if (base >= count) break;
    const uint32_t ri = base + (uint32_t)l;
    if (ri < count) {
    const GPRTRay ray = rays[ri];
    const RayPre r = ray_setup(ray.ox, ray.oy, ray.oz, ray.tmin, ray.dx, ray.dy, ray.dz);
    float tbest = ray.tmax, bu = 0.f, bv = 0.f;
    int best = -1;

    if (nprims == 1) {
      intersect_leaf(r, tris, 0, nprims, tbest, best, bu, bv);
    } else {
      int node = 0;
      int sp = 0;
      // Each inner node is fetched at most once per ray; the step bound and the index check
      // only guard against a malformed tree.
      for (uint32_t step = 0; step < nprims; ++step) {
        if ((uint32_t)node >= nprims - 1) break;
        const BVHNode nd = nodes[node];
        float tn0, tn1;
        bool h0 = box_hit(r, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, tbest, tn0);
        bool h1 = box_hit(r, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, tbest, tn1);
        const int c0 = nd.d.x, c1 = nd.d.y;
        if (h0 && c0 < 0) {
          intersect_leaf(r, tris, ~c0, nprims, tbest, best, bu, bv);
          h0 = false;
        }
        if (h1 && c1 < 0) {
          intersect_leaf(r, tris, ~c1, nprims, tbest, best, bu, bv);
          h1 = false;
        }
        if (h0 && h1) {
          const bool firstNear = tn0 <= tn1;
          const int nearC = firstNear ? c0 : c1;
          const int farC = firstNear ? c1 : c0;
          if (sp < kLdsStack)
            s_stack[sp][tid] = (uint32_t)farC;
          else if (sp < kMaxStack)
            mySpill[sp - kLdsStack] = (uint32_t)farC;
          ++sp;
          node = nearC;
        } else if (h0) {
          node = c0;
        } else if (h1) {
          node = c1;
        } else {
          if (sp == 0) break;
          --sp;
          if (sp >= kMaxStack) break;  // unreachable for trees of depth <= 64
          node = (int)(sp < kLdsStack ? s_stack[sp][tid] : mySpill[sp - kLdsStack]);
        }
      }
    }
    GPRTHit h;
    h.t = best >= 0 ? tbest : ray.tmax;
    h.primID = best;
    h.u = bu;
    h.v = bv;
    hits[ri] = h;
    }
  }
}

}  // namespace

bool trace_closest(Context *ctx, hipStream_t s, Accel *a, const void *rays, void *hits, uint32_t count) {
  if (count == 0) return true;
  const unsigned blocks = (unsigned)ctx->cuCount * 8;
  const size_t need = (size_t)blocks * kTraceBlock * (kMaxStack - kLdsStack) * 4 + 256;
  if (need > ctx->traceBytes) {
    if (ctx->traceScratch) {
      hipStreamSynchronize(s);
      hipFree(ctx->traceScratch);
      ctx->traceScratch = nullptr;
      ctx->traceBytes = 0;
    }
    if (!hip_ok(hipMalloc(&ctx->traceScratch, need), "hipMalloc(trace scratch)")) return false;
    ctx->traceBytes = need;
  }
  uint32_t *work = (uint32_t *)ctx->traceScratch;
  void *spill = (char *)ctx->traceScratch + 256;
  fill_u32(ctx, s, work, 0u, 1);
  launch(ctx, s, "trace_closest", k_trace, dim3(blocks), dim3(kTraceBlock), 0, (const BVHNode *)a->nodes,
         (const TriRecord *)a->tris, a->n, (const GPRTRay *)rays, (GPRTHit *)hits, count, work, (uint32_t *)spill);
  return true;
}

}  // namespace gprt
