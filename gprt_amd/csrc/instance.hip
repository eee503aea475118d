// Instance accels (two-level scenes): gprtInstanceAccelCreate over a buffer of gprt::Instance
// records (gprt/gprt_host.h:1418, gprt/gprt_shared.h:92-101), gprtAccelGetInstance (:1570), build,
// closest-hit trace, and bounce-ray generation over such scenes.
//
// The reference hands instances to the Vulkan driver's TLAS. Here the top level opens every
// instance down to its triangles: an LBVH over one world-space box per triangle of every enabled
// instance, built by the same builder as triangle accels. Instances that overlap in space (every
// blob mesh spans the unit cube) then cost a ray what one merged tree over all their triangles
// costs; leaves of one mesh only (whole instances, subtrees of 64 triangles or runs of 4
// consecutive triangles were all tried) left the top level several times less selective. A ray
// walks the top tree and intersects a leaf's triangle in its instance's object space; closest hits
// are ordered by (t, instance index, primID), so the result equals the per-instance closest hits
// combined in that order, whatever the visiting order. The instances' own trees serve the learned
// proposal: its classes are fixed groups of consecutive (Morton-ordered) triangles of an instance,
// each searched through the smallest node of the instance's tree that covers it.
//
// Instance records name triangle accels by gprt::Instance::accelAddress, which is a slot of the
// context's descriptor pool (gprtAccelGetDeviceAddress); the build and the trace check every
// address and descriptor on the device before reading through them, so a stale or foreign address
// disables its instance instead of faulting.
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh.hpp"
#include "device_common.hpp"
#include "internal.hpp"
#include "traverse.hpp"

namespace gprt {

// gprt::Instance (64 B): float3x4 transform (rows [A | b], world = A * object + b),
// instanceCustomIndex:24 | mask:8, geometryIndex:24 | flags:8, accelAddress.
struct InstanceRecord {
  float xf[12];
  uint32_t w0, w1;
  uint64_t addr;
};
static_assert(sizeof(InstanceRecord) == 64, "gprt::Instance is 64 bytes");
static_assert(sizeof(InstanceRecord) == sizeof(GPRTInstance), "GPRTInstance mirrors gprt::Instance");
static_assert(sizeof(BlasDesc) == 64, "descriptor slots are 64 bytes");

constexpr uint32_t kInstEnabled = 1u, kInstIdentity = 2u;

// Per-instance table written by the build (160 B).
struct TlasInst {
  const BVHNode *nodes;
  const TriRecord *tris;
  const char *verts;
  const char *idx;
  uint32_t n, flags, vstride, istride, nverts, slot;
  uint32_t first;      // first top-level leaf of this instance
  uint32_t classSpan;  // triangles per proposal class of this instance
  float inv[12];  // object = inv[:, :3] * world + inv[:, 3]
  float fwd[12];  // the record's transform
};
static_assert(sizeof(TlasInst) == 160, "TlasInst layout");

struct Tlas {
  Buffer *instances = nullptr;
  uint32_t numInstances = 0, leaves = 0, leafCap = 0;
  // Proposal classes: 2^classDepth per instance, class c of instance c >> classDepth holding its
  // triangles [g * span, (g + 1) * span) in the instance's leaf order, g = c & (2^classDepth - 1);
  // stable across rebuilds, unlike any numbering of the top tree.
  uint32_t classDepth = 0, classes = 0;
  int2 *classNodes = nullptr;
  int32_t *leafClass = nullptr;   // [leaves] class of each top-level leaf, or -1
  uint32_t *instFirst = nullptr;  // [numInstances + 1] first top-level leaf of each instance
  std::vector<uint32_t> hostFirst;
  std::vector<unsigned char> hostRecords;
  uint64_t builtEpoch = ~0ull;  // the context's descriptor epoch when the table was last built
  uint64_t builtContentEpoch = ~0ull;  // the context's triangle-accel content epoch at that build
  void *block = nullptr;
  size_t blockBytes = 0;
  TlasInst *insts = nullptr;
  int2 *subroots = nullptr;  // {instance or -1, node (>= 0 inner, ~leaf position)}
  // The top tree is an ordinary triangle-accel build over one "triangle" per sub-root whose
  // vertices are the box corners lo, hi and their midpoint (its bounding box is the sub-root box).
  Buffer vbuf, ibuf;
  GeomType gtype{};
  Geom geom{};
  Accel top;
};

namespace {

// The instance's triangle accel is still alive and still laid out where the instance table says
// (a rebuild that moved it, or its destruction, makes this false until the instance accel is
// rebuilt).
__device__ __forceinline__ bool blas_live(const BlasDesc *__restrict__ pool, const TlasInst &ti) {
  if (ti.slot >= kBlasPoolSlots) return false;
  const BlasDesc d = pool[ti.slot];
  return d.magic == kBlasMagic && d.nodes == ti.nodes && d.tris == ti.tris && d.n == ti.n && d.verts == ti.verts &&
         d.idx == ti.idx;
}

// One thread per instance: resolve the record's address to a live triangle accel of this context,
// invert the transform (FP64, one rounding to FP32) and fill the instance table.
__global__ void __launch_bounds__(256) k_tlas_inst(const InstanceRecord *__restrict__ recs, uint32_t ninst,
                                                   uint32_t classDepth, const uint32_t *__restrict__ instFirst,
                                                   const BlasDesc *__restrict__ pool, TlasInst *__restrict__ insts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  const InstanceRecord rec = recs[i];
  const uint64_t base = (uint64_t)pool;
  bool valid = rec.addr >= base && rec.addr < base + (uint64_t)kBlasPoolSlots * sizeof(BlasDesc) &&
               ((rec.addr - base) % sizeof(BlasDesc)) == 0;
  const uint32_t slot = valid ? (uint32_t)((rec.addr - base) / sizeof(BlasDesc)) : 0xFFFFFFFFu;
  BlasDesc d{};
  if (valid) {
    d = pool[slot];
    valid = d.magic == kBlasMagic && d.n > 0;
  }
  double A[3][3], b[3];
  bool identity = true, finite = true;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) {
      A[r][c] = rec.xf[4 * r + c];
      identity = identity && rec.xf[4 * r + c] == (r == c ? 1.0f : 0.0f);
      finite = finite && isfinite(rec.xf[4 * r + c]);
    }
    b[r] = rec.xf[4 * r + 3];
    identity = identity && rec.xf[4 * r + 3] == 0.0f;
    finite = finite && isfinite(rec.xf[4 * r + 3]);
  }
  double inv[3][3];
  const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                     A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                     A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  valid = valid && finite && det != 0.0 && isfinite(det);
  if (valid) {
    const double id = 1.0 / det;
    inv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id;
    inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
    inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
    inv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * id;
    inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
    inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
    inv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id;
    inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
    inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  }
  const uint32_t mask = rec.w0 >> 24;
  TlasInst t;
  t.nodes = (const BVHNode *)d.nodes;
  t.tris = (const TriRecord *)d.tris;
  t.verts = d.verts;
  t.idx = d.idx;
  t.n = valid ? d.n : 0u;
  t.flags = (valid && mask != 0 ? kInstEnabled : 0u) | (identity ? kInstIdentity : 0u);
  t.vstride = d.vstride;
  t.istride = d.istride;
  t.nverts = d.nverts;
  t.slot = slot;
  // The host laid the leaves out from its own view of the instances; an instance the device finds
  // different (never for a consistent context) keeps its leaves disabled.
  const uint32_t hostCount = instFirst[i + 1] - instFirst[i];
  if (hostCount != t.n) {
    t.n = 0;
    t.flags &= ~kInstEnabled;
  }
  t.first = instFirst[i];
  t.classSpan = t.n ? (t.n + (1u << classDepth) - 1u) >> classDepth : 0u;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) t.inv[4 * r + c] = valid ? (float)inv[r][c] : 0.f;
    t.inv[4 * r + 3] = valid ? (float)(-(inv[r][0] * b[0] + inv[r][1] * b[1] + inv[r][2] * b[2])) : 0.f;
    for (int c = 0; c < 4; ++c) t.fwd[4 * r + c] = rec.xf[4 * r + c];
  }
  insts[i] = t;
}

// World box of an object-space box under [A | b]: centre / half-extent form in FP64, widened by
// 2^-18 of its magnitude so that rounding of the object-space ray cannot cull a hit.
__device__ __forceinline__ void world_box(const float *fwd, float box[6]) {
  double lo[3], hi[3];
  for (int r = 0; r < 3; ++r) {
    double c = fwd[4 * r + 3], e = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double ck = 0.5 * ((double)box[k] + (double)box[3 + k]);
      const double ek = 0.5 * ((double)box[3 + k] - (double)box[k]);
      c += (double)fwd[4 * r + k] * ck;
      e += fabs((double)fwd[4 * r + k]) * ek;
    }
    const double pad = (fabs(c) + e) * 3.814697265625e-06 + 1e-30;
    lo[r] = c - e - pad;
    hi[r] = c + e + pad;
  }
  for (int r = 0; r < 3; ++r) {
    box[r] = (float)lo[r];
    box[3 + r] = (float)hi[r];
    if ((double)box[r] > lo[r]) box[r] = nextafterf(box[r], -INFINITY);
    if ((double)box[3 + r] < hi[r]) box[3 + r] = nextafterf(box[3 + r], INFINITY);
  }
}

// One thread per top-level leaf s: the instance holding it (binary search over the instances' first
// leaves) and its triangle p; the leaf's box is the world box of that triangle record. Leaves of a
// disabled instance become a point (rays practically never enter them).
__global__ void __launch_bounds__(256) k_tlas_cut(const TlasInst *__restrict__ insts, uint32_t ninst,
                                                  const uint32_t *__restrict__ instFirst, uint32_t leaves,
                                                  uint32_t classDepth, int2 *__restrict__ subroots,
                                                  int32_t *__restrict__ leafClass, float *__restrict__ topVerts) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= leaves) return;
  uint32_t lo = 0, hi = ninst;  // largest i with instFirst[i] <= s
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (instFirst[mid] <= s) lo = mid; else hi = mid;
  }
  const uint32_t i = lo;
  const TlasInst t = insts[i];
  const uint32_t p = s - instFirst[i];
  const bool enabled = (t.flags & kInstEnabled) && p < t.n;
  float box[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (p < t.n) {
    const TriRecord tr = t.tris[p];
    box[0] = fmin_ord(tr.v0.x, fmin_ord(tr.v1.x, tr.v2.x));
    box[1] = fmin_ord(tr.v0.y, fmin_ord(tr.v1.y, tr.v2.y));
    box[2] = fmin_ord(tr.v0.z, fmin_ord(tr.v1.z, tr.v2.z));
    box[3] = fmax_ord(tr.v0.x, fmax_ord(tr.v1.x, tr.v2.x));
    box[4] = fmax_ord(tr.v0.y, fmax_ord(tr.v1.y, tr.v2.y));
    box[5] = fmax_ord(tr.v0.z, fmax_ord(tr.v1.z, tr.v2.z));
    if (!(t.flags & kInstIdentity)) world_box(t.fwd, box);
  }
  if (!enabled) {
    for (int c = 0; c < 3; ++c) box[c] = box[3 + c] = box[c] * 0.5f + box[3 + c] * 0.5f;
  }
  subroots[s] = make_int2(enabled ? (int)i : -1, (int)p);
  leafClass[s] = enabled ? (int32_t)((i << classDepth) + p / t.classSpan) : -1;
  float *v = topVerts + (size_t)9 * s;
  for (int c = 0; c < 3; ++c) {
    v[c] = box[c];
    v[3 + c] = box[3 + c];
    v[6 + c] = box[c] * 0.5f + box[3 + c] * 0.5f;
  }
}

// Node of each proposal class: class c of instance i = c >> Dc covers `classSpan` consecutive leaves
// of the instance's tree; its node is the smallest node whose leaf range holds them, found by
// descending on the split positions (a left child's range ends at its index, or at ~index for a
// leaf).
__global__ void __launch_bounds__(256) k_tlas_classes(const TlasInst *__restrict__ insts, uint32_t ninst,
                                                      uint32_t classDepth, int2 *__restrict__ classNodes) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (ninst << classDepth)) return;
  const uint32_t i = c >> classDepth, g = c & ((1u << classDepth) - 1u);
  const TlasInst t = insts[i];
  const uint32_t span = t.classSpan;
  const uint32_t a = g * span;
  bool enabled = (t.flags & kInstEnabled) && a < t.n;
  int node = 0;
  if (enabled) {
    const uint32_t b = min(a + span, t.n) - 1;
    node = t.n == 1 ? ~0 : 0;
    for (int level = 0; level < 72 && node >= 0; ++level) {
      if ((uint32_t)node >= t.n - 1) {
        enabled = false;  // malformed tree: never reached for a built accel
        break;
      }
      const int4 d = t.nodes[node].d;
      const uint32_t split = d.x >= 0 ? (uint32_t)d.x : (uint32_t)~d.x;
      if (b <= split)
        node = d.x;
      else if (a > split)
        node = d.y;
      else
        break;
    }
  }
  classNodes[c] = make_int2(enabled ? (int)i : -1, node);
}

__global__ void __launch_bounds__(256) k_tlas_index(uint32_t *__restrict__ idx, uint32_t words) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) idx[i] = i;
}

// Object-space ray of instance `ti` (fixed operation order, FP32).
__device__ __forceinline__ RayPre object_ray(const TlasInst &ti, const GPRTRay &ray) {
  const float *m = ti.inv;
  const float ox = ((m[0] * ray.ox + m[1] * ray.oy) + m[2] * ray.oz) + m[3];
  const float oy = ((m[4] * ray.ox + m[5] * ray.oy) + m[6] * ray.oz) + m[7];
  const float oz = ((m[8] * ray.ox + m[9] * ray.oy) + m[10] * ray.oz) + m[11];
  const float dx = (m[0] * ray.dx + m[1] * ray.dy) + m[2] * ray.dz;
  const float dy = (m[4] * ray.dx + m[5] * ray.dy) + m[6] * ray.dz;
  const float dz = (m[8] * ray.dx + m[9] * ray.dy) + m[10] * ray.dz;
  return ray_setup(ox, oy, oz, ray.tmin, dx, dy, dz);
}

// kCheck: the instance table may be stale (a triangle accel was destroyed or moved since this
// instance accel was built), so every use of an instance first checks its descriptor.
// The top-level build writes no leaf records of its own: after it, top-level leaf position `pos`
// (key order, sorted primitive id k) gets a copy of its triangle record in the instance's own
// space: {v0, primID} {v1, instance | kLeafIdentity, or -1 when disabled} {v2, k}; a disabled leaf
// gets zero geometry, which no walk reads. Only the parent nodes' boxes steer the top-level walk,
// so a leaf test then reads one record instead of three dependent loads (leaf -> instance ->
// triangle). An instance accel thus traces the triangles its instances held when it was built.
constexpr uint32_t kLeafIdentity = 0x40000000u;
constexpr uint32_t kLeafInstMask = 0x3FFFFFFFu;

__global__ void __launch_bounds__(256) k_tlas_tag(TriRecord *__restrict__ topTris, uint32_t leaves,
                                                  const uint32_t *__restrict__ sortedIds,
                                                  const int2 *__restrict__ subroots,
                                                  const TlasInst *__restrict__ insts) {
  const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= leaves) return;
  const uint32_t k = sortedIds[pos];
  const int2 sr = k < leaves ? subroots[k] : make_int2(-1, 0);
  TriRecord r;
  r.v0 = r.v1 = r.v2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int code = -1;
  if (sr.x >= 0) {
    const TlasInst &ti = insts[sr.x];
    r = ti.tris[sr.y];  // enabled leaves satisfy sr.y < ti.n
    code = (int)((uint32_t)sr.x | ((ti.flags & kInstIdentity) ? kLeafIdentity : 0u));
  }
  r.v1.w = __int_as_float(code);
  r.v2.w = __uint_as_float(k);
  topTris[pos] = r;
}

// kPlain: no proposal, no per-ray fetch counts and fewer than kSmallTreePrims top-level leaves
// (the common trace: no skip compares or fetch counter, 32-bit node and record offsets).
template <bool kCheck, bool kPlain>
__global__ void __launch_bounds__(kTraceBlock)
    k_trace_tlas(const BVHNode *__restrict__ topNodes, const TriRecord *__restrict__ topTris, uint32_t leaves,
                 const TlasInst *__restrict__ insts,
                 const BlasDesc *__restrict__ pool, const GPRTRay *__restrict__ rays, GPRTHit *__restrict__ hits,
                 uint32_t count, uint32_t *__restrict__ work, uint32_t *__restrict__ spill,
                 const int32_t *__restrict__ proposal, const int2 *__restrict__ classNodes, uint32_t numClasses,
                 int32_t *__restrict__ instOut, int32_t *__restrict__ leafOut, uint32_t *__restrict__ visitsOut) {
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
      const RayPre rw = ray_setup(ray.ox, ray.oy, ray.oz, ray.tmin, ray.dx, ray.dy, ray.dz);
      HitState h;
      hit_init(h, ray.tmax);
      uint32_t visits = 0;
      // With a proposal: the proposed class (a subtree of one instance's own tree) first, whose hit
      // tightens t; then the top-level walk, which finds that hit again and records its leaf.
      if (!kPlain && proposal) {
        const int32_t c = proposal[ri];
        if (c >= 0 && (uint32_t)c < numClasses) {
          const int2 cn = classNodes[c];
          if (cn.x >= 0) {
            const TlasInst &ti = insts[cn.x];
            if ((ti.flags & kInstEnabled) && (!kCheck || blas_live(pool, ti))) {
              if (ti.flags & kInstIdentity) {
                visits += traverse<kMaxStack>(rw, ti.nodes, ti.tris, ti.n, cn.y, h, st, 0, (uint32_t)cn.x, -1);
              } else {
                const RayPre ro = object_ray(ti, ray);
                visits += traverse<kMaxStack>(ro, ti.nodes, ti.tris, ti.n, cn.y, h, st, 0, (uint32_t)cn.x, -1);
              }
            }
          }
        }
      }
      auto topLeaf = [&](int pos, int) __attribute__((always_inline)) {
                                  // A top-level leaf: one triangle of one instance (a copy of its
                                  // record), intersected in the instance's space.
                                  if ((uint32_t)pos >= leaves) return;
                                  const TriRecord tr = kPlain ? at32(topTris, (uint32_t)pos) : topTris[pos];
                                  const int code = __float_as_int(tr.v1.w);
                                  if (code < 0) return;
                                  const uint32_t inst = (uint32_t)code & kLeafInstMask;
                                  const int k = (int)__float_as_uint(tr.v2.w);
                                  if (kCheck && !blas_live(pool, insts[inst])) return;
                                  if ((uint32_t)code & kLeafIdentity) {
                                    intersect_record(rw, tr, inst, k, h);
                                  } else {
                                    const RayPre ro = object_ray(insts[inst], ray);
                                    intersect_record(ro, tr, inst, k, h);
                                  }
                                };
      visits += walk_batched<kMaxStack, decltype(topLeaf) &, kPlain>(rw, topNodes, leaves, leaves == 1 ? ~0 : 0, h,
                                                                     st, 0, kNoSkip, topLeaf);
      GPRTHit o;
      o.t = h.prim >= 0 ? h.t : ray.tmax;
      o.primID = h.prim;
      o.u = h.u;
      o.v = h.v;
      store_hit_stream(hits + ri, o);
      if (instOut) instOut[ri] = h.prim >= 0 ? (int32_t)h.inst : -1;
      if (leafOut) leafOut[ri] = h.prim >= 0 ? h.leaf : -1;
      if (!kPlain && visitsOut) visitsOut[ri] = visits;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Bounce rays: compact the hit rays (flags -> exclusive scan -> scatter), then ray j starts at the
// hit point of hit (j mod hits), cosine-weighted about the geometric normal (world space) facing
// the incoming ray (Malley's method; basis of Duff et al., JCGT 6(1), 2017), in FP64 with one
// rounding at the end, with streams 20 / 21 of the counter-based generator of scenes.py.
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_hit_flags(const GPRTHit *__restrict__ hits, uint32_t n,
                                                   uint32_t *__restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = hits[i].primID >= 0 ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_hit_list(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ offs,
                                                  uint32_t n, uint32_t *__restrict__ list) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) list[offs[i]] = i;
}

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double urand(uint64_t seed, uint64_t stream, uint64_t ctr) {
  const uint64_t key = mix64(seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull + 1ull);
  return (double)(mix64(key + (ctr + 1ull) * 0x9E3779B97F4A7C15ull) >> 40) * (1.0 / 16777216.0);
}

__global__ void __launch_bounds__(256) k_bounce(const TlasInst *__restrict__ insts, uint32_t ninst, TlasInst single,
                                                const BlasDesc *__restrict__ pool, const GPRTRay *__restrict__ prim,
                                                const GPRTHit *__restrict__ hits, const int32_t *__restrict__ instIds,
                                                const uint32_t *__restrict__ list,
                                                const uint32_t *__restrict__ numHits, uint32_t count, uint64_t seed,
                                                float tmin, GPRTRay *__restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const uint32_t nh = *numHits;
  GPRTRay o{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, -1.f};  // empty interval: traces as a miss
  if (nh == 0) {
    out[j] = o;
    return;
  }
  const uint32_t src = list[j % nh];
  const GPRTRay r = prim[src];
  const GPRTHit h = hits[src];
  const int32_t ii = instIds ? instIds[src] : 0;
  TlasInst d = single;
  if (insts) {
    if (ii < 0 || (uint32_t)ii >= ninst) {
      out[j] = o;
      return;
    }
    d = insts[ii];
    if (!blas_live(pool, d)) {
      out[j] = o;
      return;
    }
  }
  const uint32_t p = (uint32_t)h.primID < d.n ? (uint32_t)h.primID : d.n - 1;
  const uint32_t *tri = (const uint32_t *)(d.idx + (size_t)p * d.istride);
  const bool identity = (d.flags & kInstIdentity) != 0;
  double v[3][3];
  for (int k = 0; k < 3; ++k) {
    const uint32_t vi = tri[k] < d.nverts ? tri[k] : d.nverts - 1;
    const float *pv = (const float *)(d.verts + (size_t)vi * d.vstride);
    const double x = pv[0], y = pv[1], z = pv[2];
    for (int c = 0; c < 3; ++c) {
      const float *m = d.fwd + 4 * c;
      v[k][c] = identity ? (c == 0 ? x : c == 1 ? y : z)
                         : (((double)m[0] * x + (double)m[1] * y) + (double)m[2] * z) + (double)m[3];
    }
  }
  const double dir[3] = {r.dx, r.dy, r.dz};
  const double t = h.t;
  const double P[3] = {(double)r.ox + t * dir[0], (double)r.oy + t * dir[1], (double)r.oz + t * dir[2]};
  const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
  const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  len = len > 1e-300 ? len : 1e-300;
  for (int k = 0; k < 3; ++k) n[k] /= len;
  if (n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2] > 0) {
    for (int k = 0; k < 3; ++k) n[k] = -n[k];
  }
  const double u1 = urand(seed, 20, j), u2 = urand(seed, 21, j);
  const double rad = sqrt(u1), phi = 2.0 * 3.14159265358979323846 * u2;
  const double lx = rad * cos(phi), ly = rad * sin(phi), lz = sqrt(fmax(0.0, 1.0 - u1));
  const double sgn = n[2] >= 0.0 ? 1.0 : -1.0;
  const double a = -1.0 / (sgn + n[2]);
  const double bb = n[0] * n[1] * a;
  const double t1[3] = {1.0 + sgn * n[0] * n[0] * a, sgn * bb, -sgn * n[0]};
  const double t2[3] = {bb, sgn + n[1] * n[1] * a, -n[1]};
  double w[3];
  for (int k = 0; k < 3; ++k) w[k] = lx * t1[k] + ly * t2[k] + lz * n[k];
  const double wl = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  o.ox = (float)P[0];
  o.oy = (float)P[1];
  o.oz = (float)P[2];
  o.tmin = tmin;
  o.dx = (float)(w[0] / wl);
  o.dy = (float)(w[1] / wl);
  o.dz = (float)(w[2] / wl);
  o.tmax = 1e30f;
  out[j] = o;
}

// Proposal class depth: the largest Dc <= 6 with instances * 2^Dc <= 4096.
uint32_t class_depth(uint32_t ninst) {
  uint32_t d = 0;
  while (d < 6 && ((uint64_t)ninst << (d + 1)) <= 4096) ++d;
  return d;
}

TlasInst single_of(const Accel *a) {
  TlasInst t;
  std::memset(&t, 0, sizeof(t));
  const Geom *g = a->geom;
  t.nodes = (const BVHNode *)a->nodes;
  t.tris = (const TriRecord *)a->tris;
  t.verts = (const char *)g->vbuf->d + g->voffset;
  t.idx = (const char *)g->ibuf->d + g->ioffset;
  t.n = a->n;
  t.flags = kInstEnabled | kInstIdentity;
  t.vstride = g->vstride;
  t.istride = g->istride;
  t.nverts = g->vcount;
  t.slot = a->slot;
  for (int r = 0; r < 3; ++r) t.fwd[4 * r + r] = t.inv[4 * r + r] = 1.f;
  return t;
}

const BlasDesc kZeroDesc{};

}  // namespace

// ------------------------------------------------------------------------------------------------
// Descriptor pool
// ------------------------------------------------------------------------------------------------
static bool pool_ensure(Context *ctx) {
  if (ctx->blasPool) return true;
  BlasDesc *p = nullptr;
  if (!hip_ok(hipMalloc((void **)&p, (size_t)kBlasPoolSlots * sizeof(BlasDesc)), "hipMalloc(pool)")) return false;
  if (!hip_ok(hipMemset(p, 0, (size_t)kBlasPoolSlots * sizeof(BlasDesc)), "hipMemset(pool)")) {
    hipFree(p);
    return false;
  }
  ctx->blasPool = p;
  return true;
}

bool blas_slot_acquire(Context *ctx, Accel *a) {
  if (!pool_ensure(ctx)) return false;
  if (!ctx->blasFree.empty()) {
    a->slot = ctx->blasFree.back();
    ctx->blasFree.pop_back();
  } else if (ctx->blasNext < kBlasPoolSlots) {
    a->slot = ctx->blasNext++;
  } else {
    fail("accel create: more than %u live triangle and AABB accels in one context", kBlasPoolSlots);
    return false;
  }
  if (ctx->slotAccel.size() <= a->slot) ctx->slotAccel.resize((size_t)a->slot + 1, nullptr);
  ctx->slotAccel[a->slot] = a;
  std::memset(&a->hdesc, 0, sizeof(a->hdesc));
  return true;
}

void blas_slot_release(Accel *a) {
  if (a->slot == 0xFFFFFFFFu) return;
  Context *ctx = a->ctx;
  hip_ok(hipMemcpyAsync(ctx->blasPool + a->slot, &kZeroDesc, sizeof(BlasDesc), hipMemcpyHostToDevice, ctx->stream),
         "hipMemcpyAsync(descriptor)");
  hipStreamSynchronize(ctx->stream);
  ctx->blasFree.push_back(a->slot);
  if (a->slot < ctx->slotAccel.size()) ctx->slotAccel[a->slot] = nullptr;
  a->slot = 0xFFFFFFFFu;
  ++ctx->blasEpoch;
}

bool blas_desc_refresh(Context *ctx, hipStream_t s, Accel *a) {
  if (a->slot == 0xFFFFFFFFu || !ctx->blasPool) return true;
  BlasDesc d;
  std::memset(&d, 0, sizeof(d));
  if (a->store && a->nodes) {
    const Geom *g = a->geom;
    d.nodes = a->nodes;
    d.tris = a->tris;
    d.header = a->header;
    const bool boxes = a->kind == AccelKind::AABBs;
    d.verts = (const char *)g->vbuf->d + g->voffset;
    d.idx = boxes ? nullptr : (const char *)g->ibuf->d + g->ioffset;
    d.n = a->n;
    d.vstride = g->vstride;
    d.istride = boxes ? 0u : g->istride;
    d.nverts = g->vcount;
    d.magic = boxes ? kBoxAccelMagic : kBlasMagic;  // instance records accept triangle accels only
    d.slot = a->slot;
  }
  if (std::memcmp(&d, &a->hdesc, sizeof(d)) == 0) return true;
  a->hdesc = d;
  ++ctx->blasEpoch;  // instance accels built before this see a moved or retired triangle accel
  // Pageable source: the copy is staged before the call returns, so a->hdesc may change later.
  return hip_ok(hipMemcpyAsync(ctx->blasPool + a->slot, &a->hdesc, sizeof(BlasDesc), hipMemcpyHostToDevice, s),
                "hipMemcpyAsync(descriptor)");
}

// ------------------------------------------------------------------------------------------------
// Instance accel build / trace
// ------------------------------------------------------------------------------------------------
void tlas_release(Accel *a) {
  Tlas *t = a->tlas;
  if (!t) return;
  hipStreamSynchronize(a->ctx->stream);
  accel_detach(&t->top);
  if (t->block) hipFree(t->block);
  delete t;
  a->tlas = nullptr;
}

uint32_t tlas_leaves(const Accel *a) { return a->tlas ? a->tlas->leaves : 0; }
uint32_t tlas_classes(const Accel *a) { return a->tlas ? a->tlas->classes : 0; }
const int32_t *tlas_leaf_class(const Accel *a) { return a->tlas ? a->tlas->leafClass : nullptr; }

size_t tlas_size(const Accel *a) {
  const Tlas *t = a->tlas;
  if (!t || !a->built) return 0;
  const uint32_t L = t->leaves;
  return t->blockBytes + (size_t)(L > 1 ? L - 1 : 1) * sizeof(BVHNode) + (size_t)L * sizeof(TriRecord);
}

bool tlas_build(Context *ctx, hipStream_t s, Accel *a) {
  Tlas *t = a->tlas;
  const uint32_t N = t->numInstances;
  if ((size_t)N * sizeof(InstanceRecord) > t->instances->size) {
    fail("gprtAccelBuild: the instance buffer holds fewer than %u records", N);
    return false;
  }
  if (!pool_ensure(ctx)) return false;
  // Lay the leaves out from the records (read back: 64 bytes per instance): instance i owns leaves
  // [first[i], first[i] + n_i) for the triangle accel its address names (0 if it names none).
  t->hostRecords.resize((size_t)N * sizeof(InstanceRecord));
  if (!hip_ok(hipMemcpyAsync(t->hostRecords.data(), t->instances->d, t->hostRecords.size(), hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(instance records)") ||
      !hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize(instance records)"))
    return false;
  const uint64_t base = (uint64_t)(uintptr_t)ctx->blasPool;
  t->hostFirst.assign((size_t)N + 1, 0u);
  uint64_t total = 0;
  for (uint32_t i = 0; i < N; ++i) {
    InstanceRecord rec;
    std::memcpy(&rec, t->hostRecords.data() + (size_t)i * sizeof(InstanceRecord), sizeof(rec));
    uint32_t n = 0;
    if (rec.addr >= base && rec.addr < base + (uint64_t)kBlasPoolSlots * sizeof(BlasDesc) &&
        (rec.addr - base) % sizeof(BlasDesc) == 0) {
      const uint32_t slot = (uint32_t)((rec.addr - base) / sizeof(BlasDesc));
      const Accel *b = slot < ctx->slotAccel.size() ? ctx->slotAccel[slot] : nullptr;
      if (b && b->built && (rec.w0 >> 24) != 0) n = b->n;
    }
    t->hostFirst[i] = (uint32_t)total;
    total += n;
  }
  if (total >= (1ull << 30)) {
    fail("gprtAccelBuild: the instances hold %llu triangles; the limit is 2^30", (unsigned long long)total);
    return false;
  }
  t->hostFirst[N] = (uint32_t)total;
  const uint32_t L = total ? (uint32_t)total : 1u;  // an empty scene keeps one disabled leaf
  if (!t->block || L > t->leafCap) {
    if (t->block) {
      hipStreamSynchronize(s);
      hipFree(t->block);
      t->block = nullptr;
    }
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const uint32_t cap = L + L / 8;  // room for some growth without a new allocation
    size_t off = al((size_t)N * sizeof(TlasInst));
    const size_t oFirst = off;
    off += al(((size_t)N + 1) * 4);
    const size_t oSub = off;
    off += al((size_t)cap * sizeof(int2));
    const size_t oLeafCls = off;
    off += al((size_t)cap * 4);
    const size_t oVerts = off;
    off += al((size_t)cap * 36);
    const size_t oIdx = off;
    off += al((size_t)cap * 12);
    const size_t oCls = off;
    off += al((size_t)t->classes * sizeof(int2));
    if (!hip_ok(hipMalloc(&t->block, off), "hipMalloc(instance accel)")) return false;
    t->blockBytes = off;
    t->leafCap = cap;
    char *b = (char *)t->block;
    t->insts = (TlasInst *)b;
    t->instFirst = (uint32_t *)(b + oFirst);
    t->subroots = (int2 *)(b + oSub);
    t->leafClass = (int32_t *)(b + oLeafCls);
    t->classNodes = (int2 *)(b + oCls);
    t->vbuf.ctx = t->ibuf.ctx = ctx;
    t->vbuf.d = b + oVerts;
    t->ibuf.d = b + oIdx;
    t->gtype = GeomType{ctx, GPRT_TRIANGLES, 0};
    t->geom.ctx = ctx;
    t->geom.type = &t->gtype;
    t->geom.vbuf = &t->vbuf;
    t->geom.vstride = 12;
    t->geom.voffset = 0;
    t->geom.ibuf = &t->ibuf;
    t->geom.istride = 12;
    t->geom.ioffset = 0;
    t->top.ctx = ctx;
    t->top.geom = &t->geom;
    t->top.mode = GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE;
    t->top.recordsExternal = true;
    t->leaves = 0;
  }
  if (L != t->leaves) {
    t->leaves = L;
    t->vbuf.size = (size_t)L * 36;
    t->ibuf.size = (size_t)L * 12;
    t->geom.vcount = 3 * L;
    t->geom.icount = L;
    launch(ctx, s, "tlas_index", k_tlas_index, dim3((3 * L + 255) / 256), dim3(256), 0, (uint32_t *)t->ibuf.d,
           3 * L);
  }
  a->n = L;
  // Pageable source: the copy is staged before the call returns.
  if (!hip_ok(hipMemcpyAsync(t->instFirst, t->hostFirst.data(), ((size_t)N + 1) * 4, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(instance layout)"))
    return false;
  launch(ctx, s, "tlas_inst", k_tlas_inst, dim3((N + 255) / 256), dim3(256), 0,
         (const InstanceRecord *)t->instances->d, N, t->classDepth, (const uint32_t *)t->instFirst,
         (const BlasDesc *)ctx->blasPool, t->insts);
  launch(ctx, s, "tlas_cut", k_tlas_cut, dim3((L + 255) / 256), dim3(256), 0, (const TlasInst *)t->insts, N,
         (const uint32_t *)t->instFirst, L, t->classDepth, t->subroots, t->leafClass, (float *)t->vbuf.d);
  launch(ctx, s, "tlas_classes", k_tlas_classes, dim3((t->classes + 255) / 256), dim3(256), 0,
         (const TlasInst *)t->insts, N, t->classDepth, t->classNodes);
  Accel *top = &t->top;
  if (!lbvh_build_batch(ctx, s, &top, 1, true)) return false;
  launch(ctx, s, "tlas_tag", k_tlas_tag, dim3((L + 255) / 256), dim3(256), 0, (TriRecord *)top->tris, L,
         (const uint32_t *)(top->store->ids + top->store->desc[top->mesh].primOffset), (const int2 *)t->subroots,
         (const TlasInst *)t->insts);
  a->nodes = top->nodes;
  a->tris = top->tris;
  a->header = top->header;
  t->builtEpoch = ctx->blasEpoch;
  t->builtContentEpoch = ctx->blasContentEpoch;
  return true;
}

bool tlas_trace(Context *ctx, hipStream_t s, Accel *a, const void *rays, void *hits, uint32_t count,
                const int32_t *proposal, uint32_t numClasses, int32_t *instOut, int32_t *leafOut, uint32_t *visits,
                unsigned blocksPerCU) {
  if (count == 0) return true;
  Tlas *t = a->tlas;
  uint32_t *work = ensure_trace_scratch(ctx, s, kMaxStack - kLdsStack);
  if (!work) return false;
  const unsigned blocks = (unsigned)ctx->cuCount * (blocksPerCU ? std::min<unsigned>(blocksPerCU, kBlocksPerCU)
                                                               : (unsigned)kBlocksPerCU);
  fill_u32(ctx, s, work, 0u, 1);
  // The proposal step walks the instances' own triangle accels, the top-level walk the copies
  // taken at this accel's build: once any triangle accel has changed since then, the proposal is
  // skipped so that hits stay those of the plain trace.
  if (t->builtContentEpoch != ctx->blasContentEpoch) proposal = nullptr;
  const uint32_t nc = proposal ? std::min(numClasses, t->classes) : 0u;
  // The unchecked kernel trusts the instance table: no triangle accel of this context was
  // destroyed or moved since the instance accel was built.
  const bool checked = t->builtEpoch != ctx->blasEpoch;
  const char *name = proposal ? "trace_closest_proposal_instances" : "trace_closest_instances";
  const bool plain = !proposal && !visits && t->leaves < kSmallTreePrims;
  auto go = [&](auto kernel) {
    launch(ctx, s, name, kernel, dim3(blocks), dim3(kTraceBlock), 0, (const BVHNode *)t->top.nodes,
           (const TriRecord *)t->top.tris, t->leaves, (const TlasInst *)t->insts, (const BlasDesc *)ctx->blasPool,
           (const GPRTRay *)rays, (GPRTHit *)hits, count, work, work + 64, proposal, (const int2 *)t->classNodes, nc,
           instOut, leafOut, visits);
  };
  if (checked)
    plain ? go(k_trace_tlas<true, true>) : go(k_trace_tlas<true, false>);
  else
    plain ? go(k_trace_tlas<false, true>) : go(k_trace_tlas<false, false>);
  return true;
}

}  // namespace gprt

using namespace gprt;

GPRT_API GPRTAccel gprtInstanceAccelCreate(GPRTContext _ctx, uint32_t numInstances, GPRTBuffer _instances) {
  Context *ctx = (Context *)_ctx;
  Buffer *inst = (Buffer *)_instances;
  if (!ctx || !inst) {
    fail("gprtInstanceAccelCreate: null argument");
    return nullptr;
  }
  if (numInstances == 0) {
    fail("gprtInstanceAccelCreate: numInstances must be non-zero");
    return nullptr;
  }
  if (numInstances > (1u << 16)) {
    fail("gprtInstanceAccelCreate: %u instances exceeds the 65536 limit", numInstances);
    return nullptr;
  }
  if ((size_t)numInstances * sizeof(InstanceRecord) > inst->size) {
    fail("gprtInstanceAccelCreate: the instance buffer (%zu B) holds fewer than %u records", inst->size,
         numInstances);
    return nullptr;
  }
  Accel *a = new Accel();
  a->ctx = ctx;
  a->kind = AccelKind::Instances;
  Tlas *t = new Tlas();
  t->instances = inst;
  t->numInstances = numInstances;
  t->classDepth = class_depth(numInstances);
  t->classes = numInstances << t->classDepth;
  a->tlas = t;
  a->n = 0;
  return (GPRTAccel)a;
}

GPRT_API GPRTInstance gprtAccelGetInstance(GPRTAccel _a) {
  Accel *a = (Accel *)_a;
  GPRTInstance r;
  std::memset(&r, 0, sizeof(r));
  if (!a || a->kind != AccelKind::Triangles) {
    fail("gprtAccelGetInstance: expects a triangle accel");
    return r;
  }
  r.transform[0][0] = r.transform[1][1] = r.transform[2][2] = 1.f;
  r.mask = 0xFF;
  r.accelAddress = (uint64_t)(uintptr_t)(a->ctx->blasPool + a->slot);
  return r;
}

GPRT_API void gprtSynthBounceRays(GPRTContext _ctx, GPRTAccel _a, GPRTBuffer _rays, GPRTBuffer _hits,
                                  GPRTBuffer _instanceIDs, uint32_t n, GPRTBuffer _out, uint32_t outCount,
                                  uint64_t seed, float tmin) {
  Context *ctx = (Context *)_ctx;
  Accel *a = (Accel *)_a;
  Buffer *rays = (Buffer *)_rays, *hits = (Buffer *)_hits, *ids = (Buffer *)_instanceIDs, *out = (Buffer *)_out;
  if (!ctx || !a || !rays || !hits || !out) {
    fail("gprtSynthBounceRays: null argument");
    return;
  }
  if (!a->built) {
    fail("gprtSynthBounceRays: accel has not been built");
    return;
  }
  if (a->kind == AccelKind::Instances && !ids) {
    fail("gprtSynthBounceRays: an instance accel needs the instance IDs of the hits");
    return;
  }
  if ((size_t)n * sizeof(GPRTRay) > rays->size || (size_t)n * sizeof(GPRTHit) > hits->size ||
      (ids && (size_t)n * 4 > ids->size) || (size_t)outCount * sizeof(GPRTRay) > out->size) {
    fail("gprtSynthBounceRays: sizes exceed a buffer");
    return;
  }
  if (n == 0 || outCount == 0) return;
  hipStream_t s = ctx->stream;
  const size_t scanBytes = scan_scratch_bytes(n);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t need = al(scanBytes) + 3 * al((size_t)n * 4);
  void *sp = ctx->ensure_arena(need);
  if (!sp) return;
  char *b = (char *)sp;
  uint32_t *state = (uint32_t *)b;  // the scan's total lands in its first word
  uint32_t *flags = (uint32_t *)(b + al(scanBytes));
  uint32_t *offs = (uint32_t *)(b + al(scanBytes) + al((size_t)n * 4));
  uint32_t *list = (uint32_t *)(b + al(scanBytes) + 2 * al((size_t)n * 4));
  const unsigned g = (n + 255) / 256;
  launch(ctx, s, "synth_hit_flags", k_hit_flags, dim3(g), dim3(256), 0, (const GPRTHit *)hits->d, n, flags);
  scan_run(ctx, s, flags, offs, n, 0, false, state, nullptr);
  launch(ctx, s, "synth_hit_list", k_hit_list, dim3(g), dim3(256), 0, (const uint32_t *)flags,
         (const uint32_t *)offs, n, list);
  const bool inst = a->kind == AccelKind::Instances;
  TlasInst single;
  std::memset(&single, 0, sizeof(single));
  if (!inst) single = single_of(a);
  launch(ctx, s, "synth_bounce", k_bounce, dim3((outCount + 255) / 256), dim3(256), 0,
         inst ? (const TlasInst *)a->tlas->insts : (const TlasInst *)nullptr, inst ? a->tlas->numInstances : 0u,
         single, (const BlasDesc *)ctx->blasPool, (const GPRTRay *)rays->d, (const GPRTHit *)hits->d,
         inst ? (const int32_t *)ids->d : (const int32_t *)nullptr, (const uint32_t *)list, (const uint32_t *)state,
         outCount, seed, tmin, (GPRTRay *)out->d);
}
