// Device-side closest-hit traversal shared by the triangle-accel and instance-accel trace kernels.
//
// One node fetch decides both children (64-byte node with both child boxes); leaf children are
// handed to a leaf callback (at once by walk, in batches across the wave by walk_batched); the
// near inner child is descended and the far one pushed on
// a per-lane stack whose first kLdsStack entries live in LDS and the rest in a per-lane global
// spill area. The stack is shared by nested traversals (instance accels: the bottom-level walk
// starts above the top-level entries), so a traversal takes the stack slot it starts at.
#pragma once

#include <climits>

#include "bvh.hpp"

namespace gprt {

constexpr int kTraceBlock = 128;
constexpr int kLdsStack = 20;
// 14 workgroups (28 waves) per CU: 16, the most the LDS stacks allow, measured no faster.
constexpr int kBlocksPerCU = 14;
// A binary radix tree over distinct 64-bit keys is at most 64 levels deep, so a single-level walk
// needs at most 64 entries and a two-level walk at most 128.
constexpr int kMaxStack = 64;
constexpr int kMaxStack2 = 128;

// Closest hit so far. Hits are ordered by (t, inst, prim) with unsigned comparisons, so the
// result does not depend on the order in which subtrees or instances are visited.
struct HitState {
  float t, u, v;
  int prim;      // -1: none yet
  uint32_t inst; // instance index (0 for a triangle accel); 0xFFFFFFFF: none yet
  int leaf;      // instance accels: index of the top-level leaf (sub-root) holding the hit
};

__device__ __forceinline__ void hit_init(HitState &h, float tmax) {
  h.t = tmax;
  h.u = h.v = 0.f;
  h.prim = -1;
  h.inst = 0xFFFFFFFFu;
  h.leaf = -1;
}

__device__ __forceinline__ bool hit_closer(float t, uint32_t inst, int prim, const HitState &h) {
  return t < h.t || (t == h.t && (inst < h.inst || (inst == h.inst && (unsigned)prim < (unsigned)h.prim)));
}

struct LaneStack {
  uint32_t (*lds)[kTraceBlock];
  int tid;
  uint32_t *spill;  // per-lane global entries [kLdsStack, maxStack)
};

// Reads stack entry `sp` (the first kLdsStack in LDS, the rest in the lane's spill area). The spill
// read is a relaxed wave-scope atomic load, an ordinary global load in the code, so that the
// compiler does not fold the two reads into one flat load behind an address select: that costs
// about seven more vector instructions per pop and waits for every outstanding memory operation.
__device__ __forceinline__ uint32_t stack_read(const LaneStack &st, int sp) {
  if (sp < kLdsStack) return st.lds[sp][st.tid];
  return __hip_atomic_load(&st.spill[sp - kLdsStack], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Tests one triangle record ({v0, primID} {v1} {v2}) of instance `inst`; `tag` is recorded with a hit.
__device__ __forceinline__ void intersect_record(const RayPre &r, const TriRecord &tr, uint32_t inst, int tag,
                                                 HitState &h) {
  float t, u, v;
  if (tri_hit(r, tr.v0, tr.v1, tr.v2, t, u, v)) {
    const int prim = (int)__float_as_uint(tr.v0.w);
    if (hit_closer(t, inst, prim, h)) {
      h.t = t;
      h.prim = prim;
      h.u = u;
      h.v = v;
      h.inst = inst;
      h.leaf = tag;
    } else if (t == h.t && inst == h.inst && prim == h.prim && h.leaf < 0) {
      h.leaf = tag;  // the same hit, first found by a walk that did not know its top-level leaf
    }
  }
}

// Tests one box-primitive record ({lo, primID} {hi} {0}); box hits report u = v = 0.
__device__ __forceinline__ void intersect_box_record(const RayPre &r, const TriRecord &tr, uint32_t inst, int tag,
                                                     HitState &h) {
  float t;
  if (boxprim_hit(r, tr.v0, tr.v1, t)) {
    const int prim = (int)__float_as_uint(tr.v0.w);
    if (hit_closer(t, inst, prim, h)) {
      h.t = t;
      h.prim = prim;
      h.u = 0.f;
      h.v = 0.f;
      h.inst = inst;
      h.leaf = tag;
    } else if (t == h.t && inst == h.inst && prim == h.prim && h.leaf < 0) {
      h.leaf = tag;
    }
  }
}

// kBoxes: the leaves are box primitives rather than triangles.
// kSmall: the tree has fewer than 2^26 primitives, so node and record byte offsets fit in 32 bits
// (an address is then one shift from the array's base instead of 64-bit arithmetic).
constexpr uint32_t kSmallTreePrims = 1u << 26;

template <class T>
__device__ __forceinline__ const T &at32(const T *base, uint32_t i) {
  return *(const T *)((const char *)base + i * (uint32_t)sizeof(T));
}

template <bool kBoxes = false, bool kSmall = false>
__device__ __forceinline__ void intersect_leaf(const RayPre &r, const TriRecord *__restrict__ tris, int leaf,
                                               uint32_t nprims, uint32_t inst, int tag, HitState &h) {
  if ((uint32_t)leaf >= nprims) return;
  const TriRecord &rec = kSmall ? at32(tris, (uint32_t)leaf) : tris[(uint32_t)leaf];
  if (kBoxes)
    intersect_box_record(r, rec, inst, tag, h);
  else
    intersect_record(r, rec, inst, tag, h);
}

// Depth-first walk of the subtree rooted at `start` (an inner node index, or ~leaf position) of a
// tree with `nprims` leaves, using stack slots [sp0, MaxStack). leaf(position, spTop) is called
// for every leaf whose box the ray meets; spTop is the first free stack slot. A child equal to
// `skip` (a subtree already searched completely; kNoSkip for none) is not entered. Returns the
// number of inner nodes fetched.
constexpr int kNoSkip = INT_MIN;

template <int MaxStack, class LeafFn>
__device__ __forceinline__ uint32_t walk(const RayPre &r, const BVHNode *__restrict__ nodes, uint32_t nprims,
                                         int start, const HitState &h, const LaneStack &st, int sp0, int skip,
                                         LeafFn &&leaf) {
  if (nprims == 0) return 0;  // an empty tree has no nodes
  if (start < 0) {
    leaf(~start, sp0);
    return 0;
  }
  int node = start;
  int sp = sp0;
  uint32_t visits = 0;
  // Each inner node is fetched at most once per walk; the step bound and the index check only
  // guard against a malformed tree.
  for (uint32_t step = 0; step < nprims; ++step) {
    if ((uint32_t)node >= nprims - 1) break;
    const BVHNode nd = nodes[node];
    ++visits;
    float tn0, tn1;
    bool h0 = box_hit(r, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, h.t, tn0);
    bool h1 = box_hit(r, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, h.t, tn1);
    const int c0 = nd.d.x, c1 = nd.d.y;
    h0 = h0 && c0 != skip;
    h1 = h1 && c1 != skip;
    if (h0 && c0 < 0) {
      leaf(~c0, sp);
      h0 = false;
    }
    if (h1 && c1 < 0) {
      // The first leaf may have tightened h.t; the box test above used the older bound, which
      // only costs a redundant leaf test.
      leaf(~c1, sp);
      h1 = false;
    }
    if (h0 && h1) {
      const bool firstNear = tn0 <= tn1;
      const int nearC = firstNear ? c0 : c1;
      const int farC = firstNear ? c1 : c0;
      if (sp < kLdsStack)
        st.lds[sp][st.tid] = (uint32_t)farC;
      else if (sp < MaxStack)
        st.spill[sp - kLdsStack] = (uint32_t)farC;
      ++sp;
      node = nearC;
    } else if (h0) {
      node = c0;
    } else if (h1) {
      node = c1;
    } else {
      if (sp == sp0) break;
      --sp;
      if (sp >= MaxStack) break;  // unreachable within the depth bounds above
      node = (int)stack_read(st, sp);
    }
  }
  return visits;
}

// Depth-first walk like walk(), with leaf tests postponed and batched across the wave: a lane
// queues the leaf children it meets (up to kLeafQueue, in registers) and keeps walking; the wave
// runs its leaf tests when at least half of its active lanes have leaves queued, or when a lane can
// go no further (no node left, or no room for another node's two leaves). The triangle test is
// long, and run per node it executes whenever any one lane of the wave needs it; batched, it
// executes for many lanes at once. Queued leaves are tested later, so box culling sees a
// bound that is less tight in between (a few more node fetches); the closest-hit rule makes the
// result independent of the order of the tests. Returns the number of inner nodes fetched.
constexpr int kLeafQueue = 4;

// kPlain: no subtree is skipped and fetches are not counted (the walk returns 0), and the tree has
// fewer than kSmallTreePrims primitives: the form of a plain trace, with fewer instructions per node.
template <int MaxStack, class LeafFn, bool kPlain = false>
__device__ __forceinline__ uint32_t walk_batched(const RayPre &r, const BVHNode *__restrict__ nodes, uint32_t nprims,
                                                 int start, const HitState &h, const LaneStack &st, int sp0, int skip,
                                                 LeafFn &&leaf) {
  constexpr int kDone = INT_MIN;
  if (nprims == 0) return 0;  // an empty tree has no nodes
  // Queued leaves as child codes (~position, negative; q0 is the next tested): decoded when tested.
  int q0 = -1, q1 = -1, q2 = -1, q3 = -1, nq = 0;
  auto enqueue = [&](int code) __attribute__((always_inline)) {
    q3 = q2;
    q2 = q1;
    q1 = q0;
    q0 = code;
    ++nq;
  };
  int node = start;
  if (start < 0) {
    enqueue(start);
    node = kDone;
  }
  int sp = sp0;
  uint32_t visits = 0;
  // Every iteration either fetches a node (each inner node at most once per walk) or tests all
  // queued leaves, so the bound only guards against a malformed tree.
  const uint32_t limit = 4u * nprims + 8u;
  for (uint32_t it = 0; it < limit; ++it) {
    if (node != kDone && nq <= kLeafQueue - 2) {
      if ((uint32_t)node >= nprims - 1) {
        node = kDone;  // only a malformed tree
      } else {
        // node >= 0 here: no sign extension
        const BVHNode nd = kPlain ? at32(nodes, (uint32_t)node) : nodes[(uint32_t)node];
        if (!kPlain) ++visits;
        float tn0, tn1;
        bool h0 = box_hit(r, nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, h.t, tn0);
        bool h1 = box_hit(r, nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, h.t, tn1);
        const int c0 = nd.d.x, c1 = nd.d.y;
        if (!kPlain) {
          h0 = h0 && c0 != skip;
          h1 = h1 && c1 != skip;
        }
        if (h0 && c0 < 0) {
          enqueue(c0);
          h0 = false;
        }
        if (h1 && c1 < 0) {
          enqueue(c1);
          h1 = false;
        }
        if (h0 && h1) {
          const bool firstNear = tn0 <= tn1;
          const int farC = firstNear ? c1 : c0;
          if (sp < kLdsStack)
            st.lds[sp][st.tid] = (uint32_t)farC;
          else if (sp < MaxStack)
            st.spill[sp - kLdsStack] = (uint32_t)farC;
          ++sp;
          node = firstNear ? c0 : c1;
        } else if (h0) {
          node = c0;
        } else if (h1) {
          node = c1;
        } else if (sp == sp0 || sp > MaxStack) {
          node = kDone;
        } else {
          --sp;
          node = (int)stack_read(st, sp);
        }
      }
    }
    const bool stuck = nq > 0 && (node == kDone || nq > kLeafQueue - 2);
    const uint64_t active = __ballot(1);
    const uint64_t waiting = __ballot(nq > 0);
    if (__ballot(stuck) || 2 * __popcll(waiting) >= __popcll(active)) {
      // One call site for the triangle test: lanes drain their queues together.
      while (nq > 0) {
        const int code = q0;
        q0 = q1;
        q1 = q2;
        q2 = q3;
        --nq;
        leaf(~code, sp);
      }
    }
    if (node == kDone && nq == 0) break;
  }
  return visits;
}

// Primitive-accel walk: leaves are triangles (or, with kBoxes, box primitives) of instance `inst`;
// `tag` is recorded with a hit.
template <int MaxStack, bool kBoxes = false, bool kPlain = false>
__device__ __forceinline__ uint32_t traverse(const RayPre &r, const BVHNode *__restrict__ nodes,
                                             const TriRecord *__restrict__ tris, uint32_t nprims, int start,
                                             HitState &h, const LaneStack &st, int sp0, uint32_t inst, int tag,
                                             int skip = kNoSkip) {
  auto leafTest = [&](int leaf, int) __attribute__((always_inline)) {
    intersect_leaf<kBoxes, kPlain>(r, tris, leaf, nprims, inst, tag, h);
  };
  return walk_batched<MaxStack, decltype(leafTest) &, kPlain>(r, nodes, nprims, start, h, st, sp0, skip, leafTest);
}

}  // namespace gprt
