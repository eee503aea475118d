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
// children (64-byte node with both child boxes); leaf children are queued and tested in batches
// across the wave (the triangle test then runs for many lanes at once); the near inner child is
// descended and the far one pushed on a per-lane stack whose first kLdsStack
// entries live in LDS and the rest in a per-lane global spill area. A binary radix tree over
// distinct 64-bit keys is at most 64 levels deep, so 64 entries always suffice.
//
// With a proposal, each ray first traverses the subtree of its proposed treelet (a real hit found
// there becomes the starting (t, primID)), then the ordinary traversal from the root runs with
// that tighter bound; the lexicographic (t, primID) rule makes the result identical.
#include <algorithm>

#include "bvh.hpp"
#include "device_common.hpp"
#include "internal.hpp"
#include "traverse.hpp"

namespace gprt {

namespace {

// kPlain: no proposal, no per-ray fetch counts and fewer than kSmallTreePrims primitives (the
// common trace, with fewer instructions per node step); otherwise the general form.
template <bool kBoxes, bool kPlain>
__global__ void __launch_bounds__(kTraceBlock)
    k_trace(const BVHNode *__restrict__ nodes, const TriRecord *__restrict__ tris, uint32_t nprims,
            const GPRTRay *__restrict__ rays, GPRTHit *__restrict__ hits, uint32_t count,
            uint32_t *__restrict__ work, uint32_t *__restrict__ spill, const int32_t *__restrict__ proposal,
            const int32_t *__restrict__ classRoot, uint32_t numClasses, uint32_t *__restrict__ visitsOut) {
  __shared__ uint32_t s_stack[kLdsStack][kTraceBlock];
  const int tid = threadIdx.x;
  const uint32_t gtid = blockIdx.x * kTraceBlock + tid;
  const LaneStack st{s_stack, tid, spill + (size_t)gtid * (kMaxStack - kLdsStack)};
  const int l = tid & 63;

  while (true) {
    uint32_t base = 0;
    if (l == 0) base = atomicAdd(work, 64u);
    base = __shfl(base, 0, 64);
    if (base >= count) break;
    const uint32_t ri = base + (uint32_t)l;
    if (ri < count) {
      const GPRTRay ray = load_ray_stream(rays + ri);
      const RayPre r = ray_setup(ray.ox, ray.oy, ray.oz, ray.tmin, ray.dx, ray.dy, ray.dz);
      HitState h;
      hit_init(h, ray.tmax);
      if constexpr (kPlain) {
        traverse<kMaxStack, kBoxes, true>(r, nodes, tris, nprims, nprims == 1 ? ~0 : 0, h, st, 0, 0u, -1);
      } else {
        uint32_t visits = 0;
        int skip = kNoSkip;
        if (proposal) {
          const int32_t c = proposal[ri];
          if (c >= 0 && (uint32_t)c < numClasses) {
            const int32_t root = classRoot[c];
            // A treelet rooted at the tree's root (cut depth 0) would only repeat the full pass.
            if (root != 0) {
              visits += traverse<kMaxStack, kBoxes>(r, nodes, tris, nprims, root, h, st, 0, 0u, -1);
              skip = root;  // searched completely: the full pass does not enter it again
            }
          }
        }
        visits +=
            traverse<kMaxStack, kBoxes>(r, nodes, tris, nprims, nprims == 1 ? ~0 : 0, h, st, 0, 0u, -1, skip);
        if (visitsOut) visitsOut[ri] = visits;
      }
      GPRTHit o;
      o.t = h.prim >= 0 ? h.t : ray.tmax;
      o.primID = h.prim;
      o.u = h.u;
      o.v = h.v;
      store_hit_stream(hits + ri, o);
    }
  }
}

}  // namespace

uint32_t *ensure_trace_scratch(Context *ctx, hipStream_t s, int spillWordsPerLane) {
  const unsigned blocks = (unsigned)ctx->cuCount * kBlocksPerCU;
  const size_t need = (size_t)blocks * kTraceBlock * (size_t)spillWordsPerLane * 4 + 256;
  if (need > ctx->traceBytes) {
    if (ctx->traceScratch) {
      hipStreamSynchronize(s);
      hipFree(ctx->traceScratch);
      ctx->traceScratch = nullptr;
      ctx->traceBytes = 0;
    }
    if (!hip_ok(hipMalloc(&ctx->traceScratch, need), "hipMalloc(trace scratch)")) return nullptr;
    ctx->traceBytes = need;
  }
  return (uint32_t *)ctx->traceScratch;
}

bool trace_closest_ex(Context *ctx, hipStream_t s, Accel *a, const void *rays, void *hits, uint32_t count,
                      const int32_t *proposal, const int32_t *classRoot, uint32_t numClasses, uint32_t *visits,
                      unsigned blocksPerCU) {
  if (count == 0) return true;
  uint32_t *work = ensure_trace_scratch(ctx, s, kMaxStack - kLdsStack);
  if (!work) return false;
  const unsigned blocks = (unsigned)ctx->cuCount * (blocksPerCU ? std::min<unsigned>(blocksPerCU, kBlocksPerCU)
                                                               : (unsigned)kBlocksPerCU);
  fill_u32(ctx, s, work, 0u, 1);
  const char *name = proposal ? "trace_closest_proposal" : "trace_closest";
  const bool plain = !proposal && !visits && a->n < kSmallTreePrims;
  auto go = [&](auto kernel) {
    launch(ctx, s, name, kernel, dim3(blocks), dim3(kTraceBlock), 0, (const BVHNode *)a->nodes,
           (const TriRecord *)a->tris, a->n, (const GPRTRay *)rays, (GPRTHit *)hits, count, work, work + 64, proposal,
           classRoot, numClasses, visits);
  };
  if (a->kind == AccelKind::AABBs)
    plain ? go(k_trace<true, true>) : go(k_trace<true, false>);
  else
    plain ? go(k_trace<false, true>) : go(k_trace<false, false>);
  return true;
}

bool trace_closest(Context *ctx, hipStream_t s, Accel *a, const void *rays, void *hits, uint32_t count) {
  return trace_closest_ex(ctx, s, a, rays, hits, count, nullptr, nullptr, 0, nullptr);
}

}  // namespace gprt
