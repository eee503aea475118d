// Device-side closest-hit traversal shared by the single-accel and multi-instance trace kernels.
#pragma once

#include "bvh.hpp"

namespace gprt {

constexpr int kTraceBlock = 128;
constexpr int kLdsStack = 20;
constexpr int kBlocksPerCU = 14;  // 28 waves per CU: the register-limited occupancy at ~60 VGPRs
constexpr int kMaxStack = 64;

struct HitState {
  float t, u, v;
  int prim;
};

__device__ __forceinline__ void intersect_leaf(const RayPre &r, const TriRecord *__restrict__ tris, int leaf,
                                               uint32_t nprims, HitState &h) {
  if ((uint32_t)leaf >= nprims) return;
  const TriRecord tr = tris[leaf];
  float t, u, v;
  if (tri_hit(r, tr.v0, tr.v1, tr.v2, t, u, v)) {
    const int prim = (int)__float_as_uint(tr.v0.w);
    if (t < h.t || (t == h.t && (unsigned)prim < (unsigned)h.prim)) {
      h.t = t;
      h.prim = prim;
      h.u = u;
      h.v = v;
    }
  }
}

// Depth-first traversal of the subtree rooted at `start` (an inner node index, or ~leaf).
__device__ __forceinline__ uint32_t traverse(const RayPre &r, const BVHNode *__restrict__ nodes,
                                             const TriRecord *__restrict__ tris, uint32_t nprims, int start,
                                             HitState &h, uint32_t (*s_stack)[kTraceBlock], int tid,
                                             uint32_t *__restrict__ mySpill) {
  if (start < 0) {
    intersect_leaf(r, tris, ~start, nprims, h);
    return 0;
  }
  int node = start;
  int sp = 0;
  uint32_t visits = 0;
  // Each inner node is fetched at most once per traversal; the step bound and the index check
  // only guard against a malformed tree.
  for (uint32_t step = 0; step < nprims; ++step) {
    if ((uint32_t)node >= nprims - 1) break;
    const BVHNode nd = nodes[node];
    ++visits;
    float tn0, tn1;
    bool h0 = box_hit(r, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, h.t, tn0);
    bool h1 = box_hit(r, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, h.t, tn1);
    const int c0 = nd.d.x, c1 = nd.d.y;
    if (h0 && c0 < 0) {
      intersect_leaf(r, tris, ~c0, nprims, h);
      h0 = false;
    }
    if (h1 && c1 < 0) {
      intersect_leaf(r, tris, ~c1, nprims, h);
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
  return visits;
}

}  // namespace gprt
