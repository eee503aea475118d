// Triangle LBVH build for gfx950: scene bounds, Morton keys, key sort, then the Karras radix tree
// and its boxes in one bottom-up pass, following the stage order of GPRT's development LBVH
// (gprt/development/gprt_lbvh.h:170-195; kernels gprt/development/gprt_lbvh.hlsl:215-428).
//
// Every kernel is launched over a tile table that covers one or many meshes: tile t belongs to
// mesh tileMesh[t] and covers tileKeys consecutive primitives of it. A batch of thousands of
// uneven meshes is therefore built by the same ~10 launches as a single mesh, and the sequence
// can be captured once into a hipGraph and replayed.
//
// Differences in form from the reference (results: the same binary radix tree, the same boxes):
//  * Scene bounds: per-tile wave/workgroup min-max reduction, then one ordered-uint atomic per
//    tile and axis, instead of six float CAS atomics per triangle on one address set
//    (hlsl:215-226).
//  * Keys are defined as morton30<<32 | primID. The primIDs enter the sort in ascending order,
//    so a stable sort of the 30 Morton bits with primID payload gives exactly the order of the
//    64-bit keys, in 4 passes of 8 bits (u32 key + u32 payload) instead of 8 passes over u64.
//  * The sort is a segmented one-sweep radix sort: per-mesh digit histograms are built by the
//    key kernel, tiles never cross a mesh, the decoupled look-back stops at the mesh's first tile,
//    and a mesh that fits in one tile is sorted entirely in LDS with no look-back at all.
//  * Hierarchy: the Karras radix tree (hlsl:287-338) with the same node numbering, built bottom-up
//    together with the boxes in one pass (agglomerative construction): every node is combined
//    exactly once when its second child arrives, O(N), with no split search and no topology
//    pass, versus every leaf doing six CAS atomics on every ancestor (hlsl:400-428).
#include <algorithm>
#include <cstring>

#include "bvh.hpp"
#include "device_common.hpp"
#include "internal.hpp"

namespace gprt {

namespace {

constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;
constexpr int kMortonBits = 30;

struct Subtree {  // a finished subtree passed up a level: {box lo, ref}, {box hi, first leaf}
  float4 a, b;
};
static_assert(sizeof(Subtree) == 32, "Subtree must be 32 bytes");

struct BuildArgs {
  const MeshDesc *desc;
  const uint32_t *tileMesh;
  uint32_t tileKeys;
  uint32_t tiles;
  uint32_t *codes, *ids, *codesTmp, *idsTmp;
  BVHNode *nodes;
  TriRecord *tris;
  float *centroids;  // [prims][3]: triangle centroids in primitive order (bounds kernel -> key kernel)
  uint32_t *parentLeaf, *parentInner;
  float *headers;
  uint32_t *hist, *status, *tickets;
  // Subtrees that a level passes to the next, in leaf order, with delta(first leaf - 1) beside
  // them: list0 holds each 1024-key sub-tile's outputs at the sub-tile's own key positions (count
  // at subCount[tile * subPerTile + sub]); list1 each 65536-key region's outputs at the region's
  // key positions (count at regionCount[first tile of the region]). The top level reuses list0.
  Subtree *list0, *list1;
  int8_t *listD0, *listD1;
  uint32_t *subCount;
  uint32_t *regionCount;
  const uint32_t *regionTile;  // first tile of every 65536-key region, mesh by mesh
  const uint32_t *bigMesh;     // meshes of more than 65536 keys (the ones with a top level)
  int passes;
  int digitBits;
};

struct TileRange {
  uint32_t mesh, lt, lo, hi;
};

__device__ __forceinline__ TileRange tile_range(const BuildArgs &a, const MeshDesc &d, uint32_t t, uint32_t m) {
  TileRange r;
  r.mesh = m;
  r.lt = t - d.tileOffset;
  r.lo = r.lt * a.tileKeys;
  r.hi = min(r.lo + a.tileKeys, d.ntri);
  return r;
}

// Geometry pointers come from the mesh table; they are device (global) memory, and saying so lets
// the loads be global rather than flat loads.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *global_ptr(const char *p) {
  return (const __attribute__((address_space(1))) T *)p;
}

// Vertex indices are clamped to the vertex count, so a malformed index buffer yields a wrong
// box rather than an out-of-bounds read.
__device__ __forceinline__ float3 load_vertex(const MeshDesc &d, uint32_t i) {
  i = i < d.nverts ? i : d.nverts - 1;
  const auto *p = global_ptr<float>(d.verts + (size_t)i * d.vstride);
  return make_float3(p[0], p[1], p[2]);
}

// primIDs are clamped to the mesh as well, so no gather can leave the caller's buffers.
__device__ __forceinline__ void load_triangle(const MeshDesc &d, uint32_t prim, float3 &a, float3 &b, float3 &c) {
  prim = prim < d.ntri ? prim : d.ntri - 1;
  const auto *p = global_ptr<uint32_t>(d.idx + (size_t)prim * d.istride);
  a = load_vertex(d, p[0]);
  b = load_vertex(d, p[1]);
  c = load_vertex(d, p[2]);
}

// Primitive `prim` as three points whose bounding box is the primitive's box: a triangle's
// vertices, a box primitive's two corners {lo.xyz, hi.xyz}, an edge's two end points (the first
// repeated), or a point three times.
__device__ __forceinline__ void load_prim(const MeshDesc &d, uint32_t prim, float3 &a, float3 &b, float3 &c) {
  prim = prim < d.ntri ? prim : d.ntri - 1;
  if (d.kind == kPrimBoxes) {
    const auto *p = global_ptr<float>(d.verts + (size_t)prim * d.vstride);
    a = make_float3(p[0], p[1], p[2]);
    b = make_float3(p[3], p[4], p[5]);
    c = a;
  } else if (d.kind == kPrimEdges) {
    const auto *e = global_ptr<uint32_t>(d.idx + (size_t)prim * d.istride);
    a = load_vertex(d, e[0]);
    b = load_vertex(d, e[1]);
    c = a;
  } else if (d.kind == kPrimPoints) {
    a = load_vertex(d, prim);
    b = a;
    c = a;
  } else {
    load_triangle(d, prim, a, b, c);
  }
}

__device__ __forceinline__ float unord(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmin_ord(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax_ord(v, __shfl_xor(v, o, 64));
  return v;
}

// Header word set while every triangle of the mesh indexes its own three vertices in order
// (triangle i: vertices 3i, 3i+1, 3i+2, as triangle soups and the blob meshes are laid out): the
// bounds pass checks it on the way, and the hierarchy pass of the same build then reads the
// vertices without gathering the index triple first.
constexpr uint32_t kIdentityWord = 15;

// ------------------------------------------------------------------------------------------------
// Init: zero histograms / look-back state / edge counts; box accumulators to min/max identities.
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_init(uint32_t *__restrict__ zero, size_t words, float *__restrict__ headers,
                                              uint32_t meshes) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t j = i; j < (size_t)meshes * 16; j += stride) {
    const uint32_t q = (uint32_t)(j & 15);
    // box accumulators; word 15: "every triangle i indexes vertices 3i, 3i+1, 3i+2" until refuted
    ((uint32_t *)headers)[j] = q < 3 ? 0xFFFFFFFFu : (q == kIdentityWord ? 1u : 0u);
  }
  for (; i < words; i += stride) zero[i] = 0u;
}

// The point a primitive's key is computed from, rounded exactly as the key computations of
// gprt_lbvh.hlsl (no contraction): a triangle's centroid (p1 + p2 + p3) / 3 (:139-146), an edge's
// midpoint (p1 + p2) / 2 (:131-137), a point itself (:228-236); a box primitive's centre is
// (lo + hi) * 0.5.
__device__ __forceinline__ float3 centroid(float3 p, float3 q, float3 s) {
  return make_float3((p.x + q.x + s.x) / 3.0f, (p.y + q.y + s.y) / 3.0f, (p.z + q.z + s.z) / 3.0f);
}
__device__ __forceinline__ float3 prim_centroid(const MeshDesc &d, float3 p, float3 q, float3 s) {
  if (d.kind == kPrimBoxes) return make_float3((p.x + q.x) * 0.5f, (p.y + q.y) * 0.5f, (p.z + q.z) * 0.5f);
  if (d.kind == kPrimEdges) return make_float3((p.x + q.x) / 2.0f, (p.y + q.y) / 2.0f, (p.z + q.z) / 2.0f);
  if (d.kind == kPrimPoints) return p;
  return centroid(p, q, s);
}

// ------------------------------------------------------------------------------------------------
// Identity-index check of one triangle mesh for a refit, which runs no bounds pass: the mesh's
// identity word (set to 1 before the launch) is cleared by any tile holding a triangle i that does
// not index vertices 3i, 3i+1, 3i+2, as the bounds pass does during a build. Reads the index
// triples only (12 B per triangle, in order), so the refit's sub-tile level can skip the dependent
// index gather for triangle soups.
// ------------------------------------------------------------------------------------------------
constexpr int kIdentThreads = 1024;

__global__ void __launch_bounds__(64) k_ident_reset(float *__restrict__ headers, uint32_t m) {
  if (threadIdx.x == 0) ((uint32_t *)headers)[(size_t)m * 16 + kIdentityWord] = 1u;
}

__global__ void __launch_bounds__(kIdentThreads) k_ident(BuildArgs a, uint32_t tileBase, uint32_t m) {
  const uint32_t t = tileBase + blockIdx.x;
  const MeshDesc d = a.desc[m];
  const TileRange r = tile_range(a, d, t, m);
  bool ident = true;
  // kIdentBatch triples per thread in flight at a time (past the end a valid one is repeated).
  constexpr int kIdentBatch = 8;
  for (uint32_t i0 = r.lo + threadIdx.x; i0 < r.hi; i0 += kIdentThreads * kIdentBatch) {
    uint32_t v[kIdentBatch][3];
#pragma unroll
    for (int u = 0; u < kIdentBatch; ++u) {
      const uint32_t i = min(i0 + u * kIdentThreads, r.hi - 1);
      const auto *ix = global_ptr<uint32_t>(d.idx + (size_t)i * d.istride);
      v[u][0] = ix[0];
      v[u][1] = ix[1];
      v[u][2] = ix[2];
    }
#pragma unroll
    for (int u = 0; u < kIdentBatch; ++u) {
      const uint32_t i = min(i0 + u * kIdentThreads, r.hi - 1);
      ident &= v[u][0] == 3 * i && v[u][1] == 3 * i + 1 && v[u][2] == 3 * i + 2;
    }
  }
  if (!__syncthreads_and(ident ? 1 : 0) && threadIdx.x == 0)
    atomicAnd((uint32_t *)a.headers + (size_t)m * 16 + kIdentityWord, 0u);
}

// ------------------------------------------------------------------------------------------------
// Scene box per mesh over all vertices of its triangles (gprt_lbvh.hlsl:121-131, :215-226). The
// triangles' centroids are stored on the way, so the key kernel reads 12 bytes per triangle
// instead of gathering the triangle again.
// ------------------------------------------------------------------------------------------------
constexpr int kBoundsThreads = 1024;  // one workgroup per tile (its six atomics per mesh stay few)

__global__ void __launch_bounds__(kBoundsThreads) k_bounds(BuildArgs a, uint32_t tileBase) {
  const uint32_t t = tileBase + blockIdx.x;
  const uint32_t m = a.tileMesh[t];
  const MeshDesc d = a.desc[m];
  const TileRange r = tile_range(a, d, t, m);
  float mn[3] = {INFINITY, INFINITY, INFINITY};
  float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool ident = true;
  // kBoundsBatch primitives per thread at a time: every index load of the batch is issued, then
  // every vertex load, so a thread waits for two memory round trips per batch, not two per
  // primitive (the centroid stores would otherwise keep the next primitive's loads behind them).
  constexpr int kBoundsBatch = 4;
  for (uint32_t i0 = r.lo + threadIdx.x; i0 < r.hi; i0 += kBoundsThreads * kBoundsBatch) {
    float3 p[kBoundsBatch], q[kBoundsBatch], s[kBoundsBatch];
    if (d.kind != kPrimTriangles) {
#pragma unroll
      for (int u = 0; u < kBoundsBatch; ++u) {
        const uint32_t i = min(i0 + u * kBoundsThreads, r.hi - 1);  // past the end: repeats a valid one
        load_prim(d, i, p[u], q[u], s[u]);
      }
    } else {
      uint32_t ia[kBoundsBatch], ib[kBoundsBatch], ic[kBoundsBatch];
#pragma unroll
      for (int u = 0; u < kBoundsBatch; ++u) {
        const uint32_t i = min(i0 + u * kBoundsThreads, r.hi - 1);
        const auto *ix = global_ptr<uint32_t>(d.idx + (size_t)i * d.istride);
        ia[u] = ix[0];
        ib[u] = ix[1];
        ic[u] = ix[2];
      }
#pragma unroll
      for (int u = 0; u < kBoundsBatch; ++u) {
        const uint32_t i = min(i0 + u * kBoundsThreads, r.hi - 1);
        ident = ident && ia[u] == 3 * i && ib[u] == 3 * i + 1 && ic[u] == 3 * i + 2;
        p[u] = load_vertex(d, ia[u]);
        q[u] = load_vertex(d, ib[u]);
        s[u] = load_vertex(d, ic[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kBoundsBatch; ++u) {
      const uint32_t i = i0 + u * kBoundsThreads;
      if (i >= r.hi) break;  // uniform per thread; the min/max below only see real primitives
      const float3 c = prim_centroid(d, p[u], q[u], s[u]);
      float *cd = a.centroids + 3 * ((size_t)d.primOffset + i);
      cd[0] = c.x;
      cd[1] = c.y;
      cd[2] = c.z;
      mn[0] = fmin_ord(mn[0], fmin_ord(p[u].x, fmin_ord(q[u].x, s[u].x)));
      mn[1] = fmin_ord(mn[1], fmin_ord(p[u].y, fmin_ord(q[u].y, s[u].y)));
      mn[2] = fmin_ord(mn[2], fmin_ord(p[u].z, fmin_ord(q[u].z, s[u].z)));
      mx[0] = fmax_ord(mx[0], fmax_ord(p[u].x, fmax_ord(q[u].x, s[u].x)));
      mx[1] = fmax_ord(mx[1], fmax_ord(p[u].y, fmax_ord(q[u].y, s[u].y)));
      mx[2] = fmax_ord(mx[2], fmax_ord(p[u].z, fmax_ord(q[u].z, s[u].z)));
    }
  }
  __shared__ float sb[kBoundsThreads / 64][6];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    mn[c] = wave_min(mn[c]);
    mx[c] = wave_max(mx[c]);
  }
  if (l == 0) {
    for (int c = 0; c < 3; ++c) {
      sb[w][c] = mn[c];
      sb[w][3 + c] = mx[c];
    }
  }
  const int allIdent = __syncthreads_and(ident ? 1 : 0);
  if (threadIdx.x == 6 && !allIdent && d.kind == kPrimTriangles)
    atomicAnd((uint32_t *)a.headers + (size_t)m * 16 + kIdentityWord, 0u);
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    float v = sb[0][c];
    for (int ww = 1; ww < kBoundsThreads / 64; ++ww) v = c < 3 ? fmin_ord(v, sb[ww][c]) : fmax_ord(v, sb[ww][c]);
    uint32_t *h = (uint32_t *)a.headers + (size_t)m * 16;
    if (c < 3)
      atomicMin(&h[c], ford(v));
    else
      atomicMax(&h[c], ford(v));
  }
}

// ------------------------------------------------------------------------------------------------
// Morton keys of triangle centroids (gprt_lbvh.hlsl:139-146, :262-277) + per-mesh digit
// histograms of every sort pass (only for meshes spanning more than one tile).
// ------------------------------------------------------------------------------------------------
constexpr int kKeysThreads = 1024;  // one workgroup per tile

template <int BITS>
__global__ void __launch_bounds__(kKeysThreads) k_keys(BuildArgs a, uint32_t tileBase) {
  constexpr int R = 1 << BITS;
  extern __shared__ uint32_t s_h[];  // [passes][R]
  const uint32_t t = tileBase + blockIdx.x;
  const uint32_t m = a.tileMesh[t];
  const MeshDesc d = a.desc[m];
  const TileRange r = tile_range(a, d, t, m);
  const bool multi = d.ntiles > 1;
  if (multi) {
    for (int i = threadIdx.x; i < a.passes * R; i += kKeysThreads) s_h[i] = 0;
    __syncthreads();
  }
  const uint32_t *h = (const uint32_t *)a.headers + (size_t)m * 16;
  const float lox = unord(h[0]), loy = unord(h[1]), loz = unord(h[2]);
  const float ex = unord(h[3]) - lox, ey = unord(h[4]) - loy, ez = unord(h[5]) - loz;
  // kKeysBatch keys per thread at a time: the centroid loads of a batch are issued together (the
  // code stores would otherwise keep the next key's loads behind them).
  constexpr int kKeysBatch = 8;
  for (uint32_t i0 = r.lo + threadIdx.x; i0 < r.hi; i0 += kKeysThreads * kKeysBatch) {
    float cx[kKeysBatch], cy[kKeysBatch], cz[kKeysBatch];
#pragma unroll
    for (int u = 0; u < kKeysBatch; ++u) {
      const uint32_t i = min(i0 + u * kKeysThreads, r.hi - 1);  // past the end: repeats a valid one
      const float *cs = a.centroids + 3 * ((size_t)d.primOffset + i);
      cx[u] = cs[0];
      cy[u] = cs[1];
      cz[u] = cs[2];
    }
#pragma unroll
    for (int u = 0; u < kKeysBatch; ++u) {
      const uint32_t i = i0 + u * kKeysThreads;
      if (i >= r.hi) break;
      const float qx = fminf(fmaxf((cx[u] - lox) / ex * 1024.0f, 0.0f), 1023.0f);
      const float qy = fminf(fmaxf((cy[u] - loy) / ey * 1024.0f, 0.0f), 1023.0f);
      const float qz = fminf(fmaxf((cz[u] - loz) / ez * 1024.0f, 0.0f), 1023.0f);
      const uint32_t code = morton30((uint32_t)qx, (uint32_t)qy, (uint32_t)qz);
      a.codes[d.primOffset + i] = code;
      if (multi) {
        // Keys in primitive order are spatially coherent, so a wave's high digits mostly agree:
        // then one lane adds the wave's count instead of 64 atomics queueing on one LDS word.
        const uint64_t active = __ballot(true);
        for (int ps = 0; ps < a.passes; ++ps) {
          const int shift = ps * BITS;
          const int bits = min(BITS, kMortonBits - shift);
          const uint32_t dig = (code >> shift) & ((1u << bits) - 1u);
          const uint32_t lead = __builtin_amdgcn_readfirstlane(dig);
          if (__ballot(dig == lead) == active) {
            if (__lane_id() == (uint32_t)__builtin_ctzll(active)) atomicAdd(&s_h[ps * R + lead], (uint32_t)__popcll(active));
          } else {
            atomicAdd(&s_h[ps * R + dig], 1u);
          }
        }
      }
    }
  }
  if (multi) {
    __syncthreads();
    uint32_t *gh = a.hist + (size_t)m * a.passes * R;
    for (int i = threadIdx.x; i < a.passes * R; i += kKeysThreads) {
      const uint32_t c = s_h[i];
      if (c) atomicAdd(&gh[i], c);
    }
  }
}

// Exclusive scan of every pass's per-mesh histogram (meshes spanning more than one tile).
template <int BITS>
__global__ void __launch_bounds__(256) k_hist_scan(BuildArgs a) {
  constexpr int R = 1 << BITS;
  const uint32_t m = blockIdx.x;
  if (a.desc[m].ntiles <= 1) return;
  __shared__ uint32_t s_w[4];
  const int tid = threadIdx.x;
  uint32_t *gh = a.hist + (size_t)m * a.passes * R;
  for (int p = 0; p < a.passes; ++p) {
    const uint32_t c = tid < R ? gh[p * R + tid] : 0u;
    const uint32_t incl = wave_incl_sum(c);
    if ((tid & 63) == 63) s_w[tid >> 6] = incl;
    __syncthreads();
    uint32_t off = incl - c;
    for (int w = 0; w < (tid >> 6); ++w) off += s_w[w];
    if (tid < R) gh[p * R + tid] = off;
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------
// Segmented one-sweep pass over (u32 Morton code, u32 primID).
// ------------------------------------------------------------------------------------------------
template <int BITS, int THREADS, int ITEMS>
__global__ void __launch_bounds__(THREADS) k_onesweep(BuildArgs a, int pass) {
  constexpr int R = 1 << BITS;
  constexpr int W = THREADS / 64;
  constexpr int TILE = THREADS * ITEMS;
  constexpr int DPT = (R + THREADS - 1) / THREADS;
  // Keys and payloads in tile order; until the reorder the front of s_buf holds the per-wave
  // digit counts instead (W x R words), so the tile needs no more LDS than its two buffers.
  __shared__ __attribute__((aligned(16))) uint32_t s_buf[TILE];
  __shared__ __attribute__((aligned(16))) uint32_t s_vbuf[TILE];
  static_assert(W * R <= TILE, "per-wave digit counts must fit in the key buffer");
  uint32_t(*s_whist)[R] = reinterpret_cast<uint32_t(*)[R]>(s_buf);
  __shared__ uint32_t s_tdigit[R];
  __shared__ uint32_t s_gbase[R];
  __shared__ uint32_t s_wtot[W];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int l = tid & 63;
  const int shift = pass * BITS;
  const int bits = min(BITS, kMortonBits - shift);
  const uint32_t dmask = (1u << bits) - 1u;
  const uint32_t *__restrict__ src = (pass & 1) ? a.codesTmp : a.codes;
  uint32_t *__restrict__ dst = (pass & 1) ? a.codes : a.codesTmp;
  const uint32_t *__restrict__ vsrc = (pass & 1) ? a.idsTmp : a.ids;
  uint32_t *__restrict__ vdst = (pass & 1) ? a.ids : a.idsTmp;

  if (tid == 0) s_tile = ticket(&a.tickets[pass]);
  for (int i = tid; i < W * R; i += THREADS) (&s_whist[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t t = s_tile;
  const uint32_t m = a.tileMesh[t];
  const MeshDesc d = a.desc[m];
  const uint32_t lt = t - d.tileOffset;
  const uint32_t lo = lt * TILE;
  const uint32_t valid = min((uint32_t)TILE, d.ntri - lo);
  const uint32_t base = d.primOffset + lo;
  const uint32_t segEnd = d.primOffset + d.ntri;
  const bool single = d.ntiles == 1;

  // Keys and payloads are loaded together up front, so the payload loads overlap the ranking and
  // the look-back; digits are recomputed from the keys (fewer live registers). The first pass's
  // input is in primitive order, so its payload (the primID) is the key's position.
  auto digit = [&](uint32_t x) { return (x >> shift) & dmask; };
  uint32_t k[ITEMS], v[ITEMS], pos[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t q = (uint32_t)(w * ITEMS * 64 + j * 64 + l);
    k[j] = q < valid ? src[base + q] : 0xFFFFFFFFu;
    v[j] = q < valid ? (pass == 0 ? lo + q : vsrc[base + q]) : 0u;
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t dj = digit(k[j]);
    const uint64_t mm = match_digit<BITS>(dj, ~0ull);
    const uint32_t rnk = mask_rank(mm);
    const uint32_t c = (uint32_t)__popcll(mm);
    const uint32_t pre = s_whist[w][dj];
    if (rnk == 0) s_whist[w][dj] = pre + c;
    pos[j] = pre + rnk;
  }
  __syncthreads();

  uint32_t cnt[DPT];
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int dd = tid * DPT + q;
    cnt[q] = 0;
    if (dd < R) {
      uint32_t sum = 0;
#pragma unroll
      for (int ww = 0; ww < W; ++ww) {
        const uint32_t c = s_whist[ww][dd];
        s_whist[ww][dd] = sum;
        sum += c;
      }
      cnt[q] = sum;
      if (!single)
        st_agent(&a.status[((size_t)pass * a.tiles + t) * R + dd], (lt == 0 ? kFlagInc : kFlagAgg) | sum);
    }
  }
  {
    uint32_t tsum = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) tsum += cnt[q];
    const uint32_t incl = wave_incl_sum(tsum);
    if (l == 63) s_wtot[w] = incl;
    __syncthreads();
    uint32_t off = incl - tsum;
    for (int ww = 0; ww < w; ++ww) off += s_wtot[ww];
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const int dd = tid * DPT + q;
      if (dd < R) s_tdigit[dd] = off;
      off += cnt[q];
    }
  }
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int dd = tid * DPT + q;
    if (dd >= R) continue;
    if (single) {
      s_gbase[dd] = base;
      continue;
    }
    uint32_t excl = 0;
    if (lt > 0) {
      // Look back four predecessors per step (four loads in flight): sum aggregates up to the
      // first inclusive value; an unpublished tile is waited for and the walk resumes there.
      int64_t tt = (int64_t)t - 1;
      const int64_t first = d.tileOffset;
      while (tt >= first) {
        const int nb = (int)min((int64_t)4, tt - first + 1);
        uint32_t sv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (i < nb) sv[i] = ld_agent(&a.status[((size_t)pass * a.tiles + (tt - i)) * R + dd]);
        int used = 0;
        bool done = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (done || i >= nb || used < i) continue;
          const uint32_t flag = sv[i] & ~kValueMask;
          if (flag == 0u) continue;  // not published: stop consuming here
          excl += sv[i] & kValueMask;
          used = i + 1;
          done = flag == kFlagInc;
        }
        if (done) break;
        tt -= used;
        if (used < nb) __builtin_amdgcn_s_sleep(1);
      }
      st_agent(&a.status[((size_t)pass * a.tiles + t) * R + dd], kFlagInc | (excl + cnt[q]));
    }
    const uint32_t gstart = a.hist[((size_t)m * a.passes + pass) * R + dd];
    s_gbase[dd] = d.primOffset + gstart + excl - s_tdigit[dd];
  }
  // s_tdigit / s_gbase of every digit must be visible before any lane reads another thread's.
  __syncthreads();

#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t dj = digit(k[j]);
    pos[j] = s_tdigit[dj] + s_whist[w][dj] + pos[j];
  }
  __syncthreads();  // the counts are read; the buffer takes the keys
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    s_buf[pos[j]] = k[j];
    s_vbuf[pos[j]] = v[j];
  }
  __syncthreads();
  // Runs of equal digits are consecutive in the tile order, so the writes coalesce per run. Every
  // position of the tile holds a key and payload (the padding included): all LDS reads are issued
  // before the first write, so the wave waits once instead of twice per key.
  uint32_t sk[ITEMS], sv[ITEMS], gg[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    sk[j] = s_buf[j * THREADS + tid];
    sv[j] = s_vbuf[j * THREADS + tid];
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) gg[j] = s_gbase[digit(sk[j])] + (uint32_t)(j * THREADS + tid);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    // The range test guards the write; it always holds for the tile's own keys when consistent.
    if ((uint32_t)(j * THREADS + tid) < valid && gg[j] >= d.primOffset && gg[j] < segEnd) {
      dst[gg[j]] = sk[j];
      vdst[gg[j]] = sv[j];
    }
  }
}

__global__ void __launch_bounds__(256) k_copy_back(BuildArgs a, uint32_t prims) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < prims; i += stride) {
    a.codes[i] = a.codesTmp[i];
    a.ids[i] = a.idsTmp[i];
  }
}

// ------------------------------------------------------------------------------------------------
// Hierarchy and boxes in one bottom-up pass (agglomerative LBVH construction, Apetrei 2014).
//
// delta(i) = clz64(key_i ^ key_{i+1}) over adjacent sorted keys, -1 past either end of the mesh
// (the delta of gprt_lbvh.hlsl:45-59 restricted to neighbours). The node covering leaves [l, r]
// has as parent the node split at r when delta(r) > delta(l-1) (it is then the left child), else
// the node split at l-1 (right child); the two never tie, since the keys are distinct. This is
// the binary radix tree of the Karras split search (gprt_lbvh.hlsl:64-103, :287-338), built
// without the searches, and the Karras index of a node follows from its side: r for a left
// child, l for a right child, 0 for the root. Both children meet at their parent's split
// position: the first to arrive leaves its record there and stops, the second merges both boxes,
// writes the finished 64-byte node at its Karras index and climbs on (the atomic refit of
// gprt_lbvh.hlsl:400-428, with every node combined exactly once).
//
// Three levels, each merging in LDS within one workgroup (no arrival counter in global memory):
//  1. k_build_sub: one 1024-thread workgroup per 1024-key sub-tile. The sub-tile's deltas, leaf
//     boxes, finished node boxes and the meeting slots of its interior splits live in LDS. The
//     climb runs in rounds separated by barriers: round 0 sends every leaf to its parent's split,
//     and each later round merges the splits whose second child arrived in the previous round,
//     one per thread from a queue, so the shrinking work stays packed in the first waves instead
//     of leaving one active lane in every wave. The subtrees whose parents span sub-tiles (their
//     parent's split lies on the sub-tile's edge, or their sibling never arrived) partition the
//     sub-tile; they are written in leaf order as its outputs.
//  2. k_build_region: one workgroup per 65536-key region treats its sub-tiles' outputs as leaves
//     and merges them by the same rule (merge_sequence below); its outputs partition the region.
//  3. k_build_top: one workgroup per mesh merges its regions' outputs up to the root.
// Kernel boundaries order the levels.
// ------------------------------------------------------------------------------------------------
constexpr int kBuildSub = 1024;
constexpr uint32_t kBuildRegion = 65536;  // keys per level-2 region (a multiple of every tile size)
constexpr int kNoOutput = 0x7FFFFFFF;  // no output subtree starts at this leaf

__device__ __forceinline__ void lds_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local"); }
__device__ __forceinline__ void lds_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local"); }
// A workgroup barrier that orders LDS only: unlike __syncthreads it does not wait for the
// workgroup's outstanding global stores (the finished nodes), which no one reads in this kernel.
__device__ __forceinline__ void lds_barrier() {
  lds_release();
  __builtin_amdgcn_s_barrier();
  lds_acquire();
}

__device__ __forceinline__ void load_child_box(const BVHNode *nodes, uint32_t node, uint32_t side, float b[6]) {
  const float2 *p = (const float2 *)((const float *)&nodes[node] + side * 6);
  const float2 x = p[0], y = p[1], z = p[2];
  b[0] = x.x;
  b[1] = x.y;
  b[2] = y.x;
  b[3] = y.y;
  b[4] = z.x;
  b[5] = z.y;
}

__device__ __forceinline__ void merge_box(float box[6], const float sib[6]) {
  box[0] = fmin_ord(box[0], sib[0]);
  box[1] = fmin_ord(box[1], sib[1]);
  box[2] = fmin_ord(box[2], sib[2]);
  box[3] = fmax_ord(box[3], sib[3]);
  box[4] = fmax_ord(box[4], sib[4]);
  box[5] = fmax_ord(box[5], sib[5]);
}

// Writes a finished node whole: child boxes, child references, the other end of its leaf range.
__device__ __forceinline__ void write_node(BVHNode *node, const float b0[6], const float b1[6], int c0, int c1,
                                           uint32_t other) {
  float4 *dst = (float4 *)node;
  dst[0] = make_float4(b0[0], b0[1], b0[2], b0[3]);
  dst[1] = make_float4(b0[4], b0[5], b1[0], b1[1]);
  dst[2] = make_float4(b1[2], b1[3], b1[4], b1[5]);
  dst[3] = make_float4(__int_as_float(c0), __int_as_float(c1), __uint_as_float(other), 0.0f);
}

__device__ __forceinline__ void put_entry(Subtree *p, const float b[6], int ref, uint32_t first) {
  p->a = make_float4(b[0], b[1], b[2], __int_as_float(ref));
  p->b = make_float4(b[3], b[4], b[5], __uint_as_float(first));
}

// identOk: the mesh's identity-index word is current (set by this build's bounds pass, or by the
// refit's own check of the index buffer).
__global__ void __launch_bounds__(kBuildSub) k_build_sub(BuildArgs a, uint32_t tileBase, uint32_t subPerTile,
                                                         uint32_t identOk) {
  __shared__ int8_t s_delta[kBuildSub + 4];        // delta(lo - 1 + q), in [-1, 63]
  __shared__ uint16_t s_split[kBuildSub];          // split of finished local node lo + q (0xFFFF: none)
  __shared__ float s_leaf[kBuildSub][6];           // leaf boxes (position - lo)
  __shared__ float s_box[kBuildSub][6];            // finished inner node boxes (index - lo)
  __shared__ uint32_t s_meet[kBuildSub];           // sides arrived at split lo + q: bit 0 left, bit 1 right
  __shared__ int s_ref[kBuildSub][2];              // the arrived child per side (inner index or ~leaf)
  __shared__ uint16_t s_end[kBuildSub][2];         // its far leaf - lo (l of a left child, r of a right child)
  __shared__ uint16_t s_ready[2][kBuildSub / 2];   // splits whose two children have arrived
  __shared__ uint32_t s_nready[3];                 // ready-list sizes by round
  __shared__ int s_out[kBuildSub];                 // reference of the output subtree starting at lo + q
  __shared__ uint32_t s_wsum[kBuildSub / 64 + 1];  // output counts per wave, then their prefix
  const uint32_t t = tileBase + blockIdx.x / subPerTile, sub = blockIdx.x % subPerTile;
  const uint32_t m = a.tileMesh[t];
  const MeshDesc d = a.desc[m];
  const TileRange r = tile_range(a, d, t, m);
  const uint32_t lo = r.lo + sub * kBuildSub;
  uint32_t *__restrict__ subCount = a.subCount + (size_t)t * subPerTile + sub;
  if (lo >= r.hi) {  // uniform over the workgroup
    if (threadIdx.x == 0) *subCount = 0u;
    return;
  }
  const uint32_t hi = min(lo + (uint32_t)kBuildSub, r.hi);
  const uint32_t n = d.ntri;
  const uint32_t *__restrict__ codes = a.codes + d.primOffset;
  const uint32_t *__restrict__ ids = a.ids + d.primOffset;
  BVHNode *__restrict__ nodes = a.nodes + d.innerOffset;
  float *__restrict__ header = a.headers + (size_t)m * 16;
  const uint32_t tid = threadIdx.x;
  const uint32_t k = lo + tid;  // leaf position
  // The leaf's key pair (its delta) and primID are loaded in one round trip, then its triangle.
  uint32_t prim = 0, ck = 0, cn = 0, in = 0, cp = 0, ip = 0;
  if (k < hi) {
    prim = ids[k];
    ck = codes[k];
    if (k + 1 < n) {
      cn = codes[k + 1];
      in = ids[k + 1];
    }
  }
  if (tid == 0 && lo > 0) {
    cp = codes[lo - 1];
    ip = ids[lo - 1];
  }
  if (tid < 3) s_nready[tid] = 0u;
  s_meet[tid] = 0u;
  s_split[tid] = 0xFFFFu;
  s_out[tid] = kNoOutput;
  float box[6];
  const bool ident = identOk && d.kind == kPrimTriangles && ((const uint32_t *)header)[kIdentityWord] != 0u;
  if (k < hi) {
    float3 p, q, s;
    if (ident) {
      p = load_vertex(d, 3 * prim);
      q = load_vertex(d, 3 * prim + 1);
      s = load_vertex(d, 3 * prim + 2);
    } else {
      load_prim(d, prim, p, q, s);
    }
    s_delta[tid + 1] = k + 1 < n ? __clzll((((uint64_t)ck << 32) | prim) ^ (((uint64_t)cn << 32) | in)) : -1;
    if (tid == 0) s_delta[0] = lo > 0 ? __clzll((((uint64_t)cp << 32) | ip) ^ (((uint64_t)ck << 32) | prim)) : -1;
    box[0] = fmin_ord(p.x, fmin_ord(q.x, s.x));
    box[1] = fmin_ord(p.y, fmin_ord(q.y, s.y));
    box[2] = fmin_ord(p.z, fmin_ord(q.z, s.z));
    box[3] = fmax_ord(p.x, fmax_ord(q.x, s.x));
    box[4] = fmax_ord(p.y, fmax_ord(q.y, s.y));
    box[5] = fmax_ord(p.z, fmax_ord(q.z, s.z));
    // Leaf record: a triangle {v0, primID} {v1} {v2}, an edge {p1, primID} {p2} {0}, or a box
    // primitive's (a point's) ordered corners {lo, primID} {hi} {0}.
    TriRecord rec;
    if (d.kind == kPrimBoxes || d.kind == kPrimPoints) {
      rec.v0 = make_float4(box[0], box[1], box[2], __uint_as_float(prim));
      rec.v1 = make_float4(box[3], box[4], box[5], 0.0f);
      rec.v2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else if (d.kind == kPrimEdges) {
      rec.v0 = make_float4(p.x, p.y, p.z, __uint_as_float(prim));
      rec.v1 = make_float4(q.x, q.y, q.z, 0.0f);
      rec.v2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
      rec.v0 = make_float4(p.x, p.y, p.z, __uint_as_float(prim));
      rec.v1 = make_float4(q.x, q.y, q.z, 0.0f);
      rec.v2 = make_float4(s.x, s.y, s.z, 0.0f);
    }
    if (!d.noRecords) a.tris[d.primOffset + k] = rec;
    for (int c = 0; c < 6; ++c) s_leaf[tid][c] = box[c];
  }
  lds_barrier();

  // The node [l, rr] (reference `ref`, box `b`, already in LDS) arrives at its parent's split: the
  // root ends the climb, a split on the sub-tile's edge makes the node an output of the sub-tile,
  // and the second arrival at an interior split queues it for the next round.
  auto arrive = [&](uint32_t l, uint32_t rr, int ref, const float b[6], uint32_t nextQ) {
    const int dl = s_delta[l - lo], dr = s_delta[rr - lo + 1];
    if (dl < 0 && dr < 0) {  // [l, rr] = [0, n - 1]: the root (or the only leaf)
      for (int c = 0; c < 6; ++c) header[8 + c] = b[c];
      return;
    }
    const bool leftChild = dr > dl;
    const uint32_t split = leftChild ? rr : l - 1u;
    const uint32_t side = leftChild ? 0u : 1u;
    if (split < lo || split + 1u >= hi) {
      s_out[l - lo] = ref;
      return;
    }
    const uint32_t g = split - lo;
    s_ref[g][side] = ref;
    s_end[g][side] = (uint16_t)((leftChild ? l : rr) - lo);
    const uint32_t old =
        __hip_atomic_fetch_or(&s_meet[g], 1u << side, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old != 0u) {
      const uint32_t slot =
          __hip_atomic_fetch_add(&s_nready[nextQ % 3], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      s_ready[nextQ & 1][slot] = (uint16_t)g;
    }
  };

  // Round 0: the leaves. Round i > 0 merges the splits queued in round i - 1, so the active climbs
  // are packed into the first waves; the barrier between rounds orders every LDS write.
  if (k < hi) arrive(k, k, ~(int)k, box, 1u);
  lds_barrier();
  // A split is queued once, so rounds never exceed the sub-tile's keys; the bound only guards
  // against malformed keys.
  for (uint32_t round = 1; round <= kBuildSub; ++round) {
    const uint32_t count = s_nready[round % 3];
    if (count == 0u) break;  // uniform: read after the barrier
    if (tid == 0) s_nready[(round + 2) % 3] = 0u;  // the list the next round fills
    if (tid < count) {
      const uint32_t g = s_ready[round & 1][tid];
      const int c0 = s_ref[g][0], c1 = s_ref[g][1];
      const uint32_t pl = lo + s_end[g][0], pr = lo + s_end[g][1];
      const float *p0 = c0 < 0 ? s_leaf[(uint32_t)(~c0) - lo] : s_box[(uint32_t)c0 - lo];
      const float *p1 = c1 < 0 ? s_leaf[(uint32_t)(~c1) - lo] : s_box[(uint32_t)c1 - lo];
      float b0[6], b1[6];
      for (int c = 0; c < 6; ++c) {
        b0[c] = p0[c];
        b1[c] = p1[c];
      }
      const int pdl = s_delta[pl - lo], pdr = s_delta[pr - lo + 1];
      const uint32_t P = (pdl < 0 && pdr < 0) ? 0u : (pdr > pdl ? pr : pl);
      s_split[P - lo] = (uint16_t)g;  // written whole in the final pass
      merge_box(b0, b1);
      for (int c = 0; c < 6; ++c) s_box[P - lo][c] = b0[c];
      arrive(pl, pr, (int)P, b0, round + 1u);
    }
    lds_barrier();
  }
  // Finished local nodes, written whole and coalesced (node lo + tid by thread tid).
  {
    const uint32_t g = s_split[tid];
    if (g != 0xFFFFu) {
      const int c0 = s_ref[g][0], c1 = s_ref[g][1];
      const uint32_t pl = lo + s_end[g][0], pr = lo + s_end[g][1];
      const float *p0 = c0 < 0 ? s_leaf[(uint32_t)(~c0) - lo] : s_box[(uint32_t)c0 - lo];
      const float *p1 = c1 < 0 ? s_leaf[(uint32_t)(~c1) - lo] : s_box[(uint32_t)c1 - lo];
      float b0[6], b1[6];
      for (int c = 0; c < 6; ++c) {
        b0[c] = p0[c];
        b1[c] = p1[c];
      }
      const uint32_t P = lo + tid;
      write_node(&nodes[P], b0, b1, c0, c1, P == pl ? pr : pl);
    }
  }
  // Interior splits [lo, hi - 2] with one child arrived: the other child spans sub-tiles, so the
  // arrived one is an output too.
  if (tid + 1u < hi - lo) {
    const uint32_t meet = s_meet[tid];
    if (meet == 1u || meet == 2u) {
      const uint32_t side = meet >> 1;
      s_out[side == 0 ? s_end[tid][0] : tid + 1u] = s_ref[tid][side];
    }
  }
  lds_barrier();
  // The outputs partition the sub-tile's leaves; written in leaf order, the next level reads them
  // as a sequence.
  {
    const int ref = s_out[tid];
    const bool out = ref != kNoOutput;
    const uint64_t mask = __ballot(out);
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    if (lane == 0) s_wsum[wave] = (uint32_t)__popcll(mask);
    lds_barrier();
    if (tid == 0) {
      uint32_t run = 0;
      for (uint32_t w = 0; w < kBuildSub / 64; ++w) {
        const uint32_t c = s_wsum[w];
        s_wsum[w] = run;
        run += c;
      }
      s_wsum[kBuildSub / 64] = run;
      *subCount = run;
    }
    lds_barrier();
    if (out) {
      const uint32_t rank = s_wsum[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
      const float *b = ref < 0 ? s_leaf[(uint32_t)(~ref) - lo] : s_box[(uint32_t)ref - lo];
      float cb[6];
      for (int c = 0; c < 6; ++c) cb[c] = b[c];
      const size_t at = (size_t)d.primOffset + lo + rank;
      put_entry(&a.list0[at], cb, ref, lo + tid);
      a.listD0[at] = s_delta[tid];
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Levels 2 and 3: merging a sequence of subtrees in LDS.
//
// A level's input is the outputs of the level below, in leaf order: subtrees that partition its
// range, each with its box, its reference (inner index or ~leaf), its first leaf L and delta(L-1).
// Treating each subtree as one leaf, the same rule builds the nodes above them: with D[i] =
// delta(L_i - 1) and D[S] the delta after the range, the node over subtrees [a, b] goes to the
// boundary after b when D[b+1] > D[a] (a left child), else to the boundary before a. Merges run
// in rounds in LDS as in the sub-tile; nodes whose parent lies outside the window, and nodes still
// waiting for a sibling at the end, are the window's outputs, again in leaf order.
//
// A window holds at most kMergeCap subtrees. A longer sequence is merged window by window, the
// outputs of each window replacing its inputs at the front of the same list, until one window
// holds the rest (a region's sub-tiles leave about 12 subtrees each, so one window usually takes a
// whole region): a window of more than 130 subtrees always merges some (adjacent boundary deltas
// of subtrees differ, and a sequence of values in [-1, 63] without an interior local maximum has
// at most 130 entries), and its outputs are at most 2 x 65 subtrees (one per tree level on each
// side).
// ------------------------------------------------------------------------------------------------
// Regions: windows of 1024 subtrees (512 threads, 75 KB of LDS: two workgroups per CU). The top
// level runs one workgroup per large mesh and takes windows of 2048 (1024 threads, 150 KB), so a
// mesh of up to ~128 regions merges in one window.
constexpr uint32_t kRegionCap = 1024, kTopCap = 2048;

template <uint32_t kMergeCap>
struct MergeLds {
  static constexpr int kMergeThreads = kMergeCap / 2;  // the output pass takes two subtrees per thread
  float leafBox[kMergeCap][6];   // input subtrees' boxes
  float nodeBox[kMergeCap][6];   // merged nodes' boxes, by boundary
  int ref[kMergeCap];            // input subtrees' references
  uint32_t node[kMergeCap];      // merged nodes' inner indices, by boundary
  uint32_t L[kMergeCap + 1];     // first leaf of each input subtree; [S]: the position after the window
  int8_t D[kMergeCap + 4];       // delta(L - 1) of each input subtree; [S]: the delta after the window
  uint32_t meet[kMergeCap / 4];  // sides arrived per boundary, 8 bits each
  uint16_t child[kMergeCap][2];  // the arrived child per side: local id (< kMergeCap input, else node)
  uint16_t far[kMergeCap][2];    // its far input subtree (first of a left child, last of a right child)
  uint16_t ready[2][kMergeCap / 2];
  uint16_t out[kMergeCap];       // local id of the output starting at each input subtree (0xFFFF: none)
  uint32_t nready[3];
  uint32_t wsum[kMergeThreads / 64 + 1];
};

// Merges input subtrees [0, S) (already in LDS: leafBox, ref, L[0..S], D[0..S]) and appends the
// window's outputs at out[*outCount...] (outD: their deltas). Called by every thread.
template <uint32_t kMergeCap>
__device__ void merge_window(MergeLds<kMergeCap> &sm, uint32_t S, uint32_t n, BVHNode *__restrict__ nodes, float *__restrict__ header,
                             Subtree *__restrict__ out, int8_t *__restrict__ outD, uint32_t outBase,
                             uint32_t &outCount) {
  constexpr int kMergeThreads = MergeLds<kMergeCap>::kMergeThreads;
  const uint32_t tid = threadIdx.x;
  auto box_of = [&](uint32_t id) -> const float * { return id < kMergeCap ? sm.leafBox[id] : sm.nodeBox[id - kMergeCap]; };
  auto ref_of = [&](uint32_t id) -> int { return id < kMergeCap ? sm.ref[id] : (int)sm.node[id - kMergeCap]; };
  // The node over input subtrees [fa, fb] (local id `id`, box in LDS) arrives at its parent's
  // boundary.
  auto arrive = [&](uint32_t fa, uint32_t fb, uint32_t id, uint32_t nextQ) {
    const int dl = sm.D[fa], dr = sm.D[fb + 1];
    if (dl < 0 && dr < 0) {  // the whole mesh: the root
      const float *b = box_of(id);
      for (int c = 0; c < 6; ++c) header[8 + c] = b[c];
      return;
    }
    const bool leftChild = dr > dl;
    if (leftChild ? fb + 1u >= S : fa == 0u) {  // the parent lies outside the window
      sm.out[fa] = (uint16_t)id;
      return;
    }
    const uint32_t g = leftChild ? fb : fa - 1u, side = leftChild ? 0u : 1u;
    sm.child[g][side] = (uint16_t)id;
    sm.far[g][side] = (uint16_t)(leftChild ? fa : fb);
    const uint32_t shift = (g & 3u) * 8u;
    const uint32_t old = __hip_atomic_fetch_or(&sm.meet[g >> 2], (1u << side) << shift, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
    if ((old >> shift) & 3u) {
      const uint32_t slot =
          __hip_atomic_fetch_add(&sm.nready[nextQ % 3], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      sm.ready[nextQ & 1][slot] = (uint16_t)g;
    }
  };
  for (uint32_t i = tid; i < kMergeCap / 4; i += kMergeThreads) sm.meet[i] = 0u;
  for (uint32_t i = tid; i < S; i += kMergeThreads) sm.out[i] = 0xFFFFu;
  if (tid < 3) sm.nready[tid] = 0u;
  lds_barrier();
  for (uint32_t i = tid; i < S; i += kMergeThreads) arrive(i, i, i, 1u);
  lds_barrier();
  for (uint32_t round = 1; round <= kMergeCap; ++round) {
    const uint32_t count = sm.nready[round % 3];
    if (count == 0u) break;  // uniform: read after the barrier
    if (tid == 0) sm.nready[(round + 2) % 3] = 0u;
    for (uint32_t q = tid; q < count; q += kMergeThreads) {
      const uint32_t g = sm.ready[round & 1][q];
      const uint32_t i0 = sm.child[g][0], i1 = sm.child[g][1];
      const uint32_t fa = sm.far[g][0], fb = sm.far[g][1];
      const uint32_t pl = sm.L[fa], pr = sm.L[fb + 1] - 1u;
      const int pdl = sm.D[fa], pdr = sm.D[fb + 1];
      const uint32_t P = (pdl < 0 && pdr < 0) ? 0u : (pdr > pdl ? pr : pl);
      const float *p0 = box_of(i0), *p1 = box_of(i1);
      float b0[6], b1[6];
      for (int c = 0; c < 6; ++c) {
        b0[c] = p0[c];
        b1[c] = p1[c];
      }
      if (P + 1u < n) write_node(&nodes[P], b0, b1, ref_of(i0), ref_of(i1), P == pl ? pr : pl);
      merge_box(b0, b1);
      for (int c = 0; c < 6; ++c) sm.nodeBox[g][c] = b0[c];
      sm.node[g] = P;
      arrive(fa, fb, kMergeCap + g, round + 1u);
    }
    lds_barrier();
  }
  // Boundaries with one child arrived: its sibling lies partly outside the window.
  for (uint32_t g = tid; g + 1u < S; g += kMergeThreads) {
    const uint32_t meet = (sm.meet[g >> 2] >> ((g & 3u) * 8u)) & 3u;
    if (meet == 1u || meet == 2u) {
      const uint32_t side = meet >> 1;
      sm.out[side == 0 ? sm.far[g][0] : g + 1u] = sm.child[g][side];
    }
  }
  lds_barrier();
  // Outputs in leaf order: thread t takes input subtrees [2t, 2t + 1].
  const uint32_t i0 = 2 * tid, i1 = 2 * tid + 1;
  const bool o0 = i0 < S && sm.out[i0] != 0xFFFFu, o1 = i1 < S && sm.out[i1] != 0xFFFFu;
  const uint32_t mine = (uint32_t)o0 + (uint32_t)o1;
  const uint32_t lane = tid & 63u, wave = tid >> 6;
  const uint32_t incl = wave_incl_sum(mine);
  if (lane == 63u) sm.wsum[wave] = incl;
  lds_barrier();
  if (tid == 0) {
    uint32_t run = 0;
    for (uint32_t w = 0; w < kMergeThreads / 64; ++w) {
      const uint32_t c = sm.wsum[w];
      sm.wsum[w] = run;
      run += c;
    }
    sm.wsum[kMergeThreads / 64] = run;
  }
  lds_barrier();
  uint32_t rank = outBase + outCount + sm.wsum[wave] + incl - mine;
  for (uint32_t j = 0; j < 2; ++j) {
    const uint32_t i = i0 + j;
    if (!(j == 0 ? o0 : o1)) continue;
    const uint32_t id = sm.out[i];
    const float *pb = box_of(id);
    float b[6];
    for (int c = 0; c < 6; ++c) b[c] = pb[c];
    put_entry(&out[rank], b, ref_of(id), sm.L[i]);
    outD[rank] = sm.D[i];
    ++rank;
  }
  outCount += sm.wsum[kMergeThreads / 64];
  // The outputs may overwrite inputs of this list that the next window or pass loads.
  __syncthreads();
}

// Loads input subtrees [w0, w0 + S) of a sequence into LDS; `at(e)` gives entry e's index in
// `in` / `inD`. endL / endD describe what follows the sequence when the window reaches its end.
template <uint32_t kMergeCap, class At>
__device__ void load_window(MergeLds<kMergeCap> &sm, const Subtree *__restrict__ in, const int8_t *__restrict__ inD, At at,
                            uint32_t w0, uint32_t S, uint32_t total, uint32_t endL, int endD) {
  for (uint32_t i = threadIdx.x; i <= S; i += MergeLds<kMergeCap>::kMergeThreads) {
    const uint32_t e = w0 + i;
    if (e < total) {
      const size_t x = at(e);
      const float4 p = in[x].a, q = in[x].b;
      sm.L[i] = __float_as_uint(q.w);
      sm.D[i] = inD[x];
      if (i < S) {
        sm.leafBox[i][0] = p.x;
        sm.leafBox[i][1] = p.y;
        sm.leafBox[i][2] = p.z;
        sm.leafBox[i][3] = q.x;
        sm.leafBox[i][4] = q.y;
        sm.leafBox[i][5] = q.z;
        sm.ref[i] = __float_as_int(p.w);
      }
    } else {  // i == S at the end of the sequence
      sm.L[i] = endL;
      sm.D[i] = (int8_t)endD;
    }
  }
  lds_barrier();
}

// Merges a whole sequence of `total` input subtrees: window by window (outputs compacted into
// `tmp` / `tmpD` from its start) until one window holds what is left, whose outputs go to `dst`.
// `at` maps entry e of the first pass to its index in `in`. Returns the number of outputs.
template <uint32_t kMergeCap, class At>
__device__ uint32_t merge_sequence(MergeLds<kMergeCap> &sm, const Subtree *in, const int8_t *inD, At at, uint32_t total,
                                   uint32_t endL, int endD, Subtree *tmp, int8_t *tmpD, Subtree *dst, int8_t *dstD,
                                   uint32_t n, BVHNode *nodes, float *header) {
  bool first = true;
  // Each pass leaves at most 130 outputs per window of kMergeCap, so passes end; the bound only
  // guards against malformed keys.
  for (int pass = 0; pass < 16 && total > 0; ++pass) {
    const bool last = total <= kMergeCap;
    uint32_t outCount = 0;
    for (uint32_t w0 = 0; w0 < total; w0 += kMergeCap) {
      const uint32_t S = min(kMergeCap, total - w0);
      if (first) load_window(sm, in, inD, at, w0, S, total, endL, endD);
      else load_window(sm, tmp, tmpD, [](uint32_t e) { return (size_t)e; }, w0, S, total, endL, endD);
      merge_window(sm, S, n, nodes, header, last ? dst : tmp, last ? dstD : tmpD, 0u, outCount);
    }
    if (last) return outCount;
    first = false;
    total = outCount;
  }
  return 0u;
}

// Level 2: one workgroup per 65536-key region merges the outputs of its sub-tiles (each sub-tile's
// list at its own key positions in list0) into list1 at the region's key positions.
__global__ void __launch_bounds__(kRegionCap / 2) k_build_region(BuildArgs a, uint32_t regionBase) {
  __shared__ MergeLds<kRegionCap> sm;
  __shared__ uint32_t s_pre[kBuildRegion / kBuildSub + 1];
  const uint32_t t = a.regionTile[regionBase + blockIdx.x];  // the region's first tile
  const uint32_t m = a.tileMesh[t];
  const MeshDesc d = a.desc[m];
  const uint32_t lt = t - d.tileOffset;
  const uint32_t n = d.ntri;
  const uint32_t regionLo = lt * a.tileKeys;
  const uint32_t regionHi = min(regionLo + kBuildRegion, n);
  constexpr uint32_t kSubs = kBuildRegion / kBuildSub;
  const uint32_t subPerTile = a.tileKeys / kBuildSub;
  const uint32_t nsub = (regionHi - regionLo + kBuildSub - 1) / kBuildSub;
  if (threadIdx.x < 64) {
    // a sub-tile's outputs partition its keys: never more than it holds
    const uint32_t c = threadIdx.x < nsub ? min(a.subCount[(size_t)t * subPerTile + threadIdx.x],
                                                min((uint32_t)kBuildSub, regionHi - regionLo - threadIdx.x * kBuildSub))
                                          : 0u;
    s_pre[threadIdx.x + 1] = wave_incl_sum(c);
    if (threadIdx.x == 0) s_pre[0] = 0u;
  }
  static_assert(kSubs == 64, "one wave gathers the sub-tile counts");
  // delta(regionHi - 1): what follows the region's last subtree.
  int endD = -1;
  if (regionHi < n) {
    const uint32_t *codes = a.codes + d.primOffset, *ids = a.ids + d.primOffset;
    endD = __clzll((((uint64_t)codes[regionHi - 1] << 32) | ids[regionHi - 1]) ^
                   (((uint64_t)codes[regionHi] << 32) | ids[regionHi]));
  }
  __syncthreads();
  const uint32_t total = s_pre[nsub];
  const size_t base = (size_t)d.primOffset + regionLo;
  auto at = [&](uint32_t e) -> size_t {
    uint32_t j = 0, jh = nsub;  // the sub-tile j with s_pre[j] <= e < s_pre[j + 1]
    while (jh - j > 1) {
      const uint32_t mid = (j + jh) >> 1;
      if (s_pre[mid] <= e) j = mid;
      else jh = mid;
    }
    return base + (size_t)j * kBuildSub + (e - s_pre[j]);
  };
  const uint32_t outs = merge_sequence(sm, a.list0, a.listD0, at, total, regionHi, endD, a.list1 + base,
                                       a.listD1 + base, a.list1 + base, a.listD1 + base, n, a.nodes + d.innerOffset,
                                       a.headers + (size_t)m * 16);
  if (threadIdx.x == 0) a.regionCount[t] = outs;
}

// Level 3: one workgroup per mesh gathers its regions' outputs into list0 (from the mesh's start)
// and merges them up to the root.
__global__ void __launch_bounds__(kTopCap / 2) k_build_top(BuildArgs a, uint32_t fromList, uint32_t meshBase) {
  constexpr int kMergeThreads = kTopCap / 2;
  __shared__ MergeLds<kTopCap> sm;
  const uint32_t m = fromList ? a.bigMesh[blockIdx.x] : meshBase + blockIdx.x;
  const MeshDesc d = a.desc[m];
  const uint32_t n = d.ntri;
  if (n <= kBuildRegion) return;  // the region level reached the root
  const uint32_t perRegion = kBuildRegion / a.tileKeys;
  const uint32_t regions = (n + kBuildRegion - 1) / kBuildRegion;
  const size_t base = d.primOffset;
  uint32_t total = 0;
  uint32_t *pre = sm.L;  // a chunk's region-count prefix (kMergeThreads + 1 words)
  for (uint32_t r0 = 0; r0 < regions; r0 += kMergeThreads) {
    const uint32_t r = r0 + threadIdx.x;
    const uint32_t c = r < regions ? min(a.regionCount[d.tileOffset + r * perRegion], min(kBuildRegion, n - r * kBuildRegion)) : 0u;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_sum(c);
    if (lane == 63u) sm.wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (uint32_t w = 0; w < kMergeThreads / 64; ++w) {
        const uint32_t v = sm.wsum[w];
        sm.wsum[w] = run;
        run += v;
      }
      sm.wsum[kMergeThreads / 64] = run;
    }
    __syncthreads();
    pre[threadIdx.x + 1] = sm.wsum[wave] + incl;
    if (threadIdx.x == 0) pre[0] = 0u;
    __syncthreads();
    const uint32_t chunk = pre[kMergeThreads], nr = min(regions - r0, (uint32_t)kMergeThreads);
    for (uint32_t e = threadIdx.x; e < chunk; e += kMergeThreads) {
      uint32_t j = 0, jh = nr;
      while (jh - j > 1) {
        const uint32_t mid = (j + jh) >> 1;
        if (pre[mid] <= e) j = mid;
        else jh = mid;
      }
      const size_t src = base + (size_t)(r0 + j) * kBuildRegion + (e - pre[j]);
      const size_t dst = base + total + e;
      a.list0[dst] = a.list1[src];
      a.listD0[dst] = a.listD1[src];
    }
    total += chunk;
    __syncthreads();
  }
  merge_sequence(sm, a.list0 + base, a.listD0 + base, [](uint32_t e) { return (size_t)e; }, total, n, -1,
                 a.list0 + base, a.listD0 + base, a.list0 + base, a.listD0 + base, n, a.nodes + d.innerOffset,
                 a.headers + (size_t)m * 16);
}

// Parent links (inner index << 1 | side; 0xFFFFFFFF at the root) of one mesh, derived from the
// finished nodes when a consumer asks for them (the reference-layout export, proposal classes).
__global__ void __launch_bounds__(256) k_parents(const BVHNode *__restrict__ nodes, uint32_t n,
                                                 uint32_t *__restrict__ parentLeaf,
                                                 uint32_t *__restrict__ parentInner) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (n <= 1) {
    if (i == 0 && n == 1) parentLeaf[0] = 0xFFFFFFFFu;
    return;
  }
  if (i >= n - 1) return;
  if (i == 0) parentInner[0] = 0xFFFFFFFFu;
  const int4 c = nodes[i].d;
  const int ch[2] = {c.x, c.y};
  for (uint32_t side = 0; side < 2; ++side) {
    const int x = ch[side];
    if (x < 0) {
      if ((uint32_t)(~x) < n) parentLeaf[~x] = (i << 1) | side;
    } else if ((uint32_t)x < n - 1) {
      parentInner[x] = (i << 1) | side;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Reference-layout export of one mesh (gprt/development/gprt_lbvh.h:28-55).
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_export(const BVHNode *__restrict__ nodes, const uint32_t *__restrict__ codes,
                                                const uint32_t *__restrict__ ids,
                                                const uint32_t *__restrict__ parentLeaf,
                                                const uint32_t *__restrict__ parentInner,
                                                const float *__restrict__ header, uint32_t n,
                                                int4 *__restrict__ outNodes, float *__restrict__ outBoxes,
                                                uint64_t *__restrict__ outKeys) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t total = 2 * n - 1;
  if (x >= total) return;
  float b[6];
  if (x < n - 1) {
    const BVHNode nd = nodes[x];
    const int c0 = nd.d.x, c1 = nd.d.y;
    const int l = c0 >= 0 ? c0 : (int)(n - 1) + ~c0;
    const int r = c1 >= 0 ? c1 : (int)(n - 1) + ~c1;
    const uint32_t pl = parentInner[x];
    outNodes[x] = make_int4(l, r, pl == 0xFFFFFFFFu ? -1 : (int)(pl >> 1), -1);
    float b0[6], b1[6];
    load_child_box(nodes, x, 0, b0);
    load_child_box(nodes, x, 1, b1);
    for (int q = 0; q < 3; ++q) {
      b[q] = fmin_ord(b0[q], b1[q]);
      b[3 + q] = fmax_ord(b0[3 + q], b1[3 + q]);
    }
  } else {
    const uint32_t k = x - (n - 1);
    if (n == 1) {
      outNodes[x] = make_int4(-1, -1, -1, (int)ids[k]);
      for (int q = 0; q < 6; ++q) b[q] = header[8 + q];
    } else {
      const uint32_t pl = parentLeaf[k];
      outNodes[x] = make_int4(-1, -1, (int)(pl >> 1), (int)ids[k]);
      if ((pl >> 1) < n - 1) {
        load_child_box(nodes, pl >> 1, pl & 1u, b);
      } else {
        for (int q = 0; q < 6; ++q) b[q] = __uint_as_float(0x7FC00000u);
      }
    }
    if (outKeys) outKeys[k] = ((uint64_t)codes[k] << 32) | ids[k];
  }
  for (int q = 0; q < 6; ++q) outBoxes[(size_t)x * 6 + q] = b[q];
}

// ------------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------------
unsigned grid_for(size_t n, unsigned block, unsigned cap) {
  size_t b = (n + block - 1) / block;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

BuildArgs args_of(BVHStore *st) {
  BuildArgs a;
  a.desc = st->dDesc;
  a.tileMesh = st->dTileMesh;
  a.tileKeys = (uint32_t)st->cfg.tileKeys;
  a.tiles = st->tiles;
  a.codes = st->codes;
  a.ids = st->ids;
  a.codesTmp = st->codesTmp;
  a.idsTmp = st->idsTmp;
  a.nodes = (BVHNode *)st->nodes;
  a.tris = (TriRecord *)st->tris;
  a.centroids = (float *)st->centroids;
  a.parentLeaf = st->parentLeaf;
  a.parentInner = st->parentInner;
  a.headers = st->headers;
  a.hist = st->hist;
  a.status = st->status;
  a.tickets = st->tickets;
  a.list0 = (Subtree *)st->list0;
  a.list1 = (Subtree *)st->list1;
  a.listD0 = (int8_t *)st->listD0;
  a.listD1 = (int8_t *)st->listD1;
  a.subCount = st->subCount;
  a.regionCount = st->regionCount;
  a.regionTile = st->dTileMesh + st->tiles;
  a.bigMesh = a.regionTile + st->regions;
  a.passes = st->passes;
  a.digitBits = st->cfg.digitBits;
  return a;
}

// The hierarchy of meshes [meshBase, meshBase + meshes) (all of the store's, or one): one
// workgroup per sub-tile, per region, and per mesh that has more than one region.
void launch_refit(Context *ctx, hipStream_t s, BVHStore *st, const BuildArgs &a, uint32_t tileBase, uint32_t tiles,
                  uint32_t meshBase, uint32_t meshes, bool afterBounds) {
  const uint32_t subPerTile = (uint32_t)st->cfg.tileKeys / kBuildSub;
  launch(ctx, s, "lbvh_build_local", k_build_sub, dim3(tiles * subPerTile), dim3(kBuildSub), 0, a, tileBase,
         subPerTile, afterBounds ? 1u : 0u);
  const uint32_t r0 = st->regionFirst[meshBase], r1 = st->regionFirst[meshBase + meshes];
  if (r1 > r0) launch(ctx, s, "lbvh_build_region", k_build_region, dim3(r1 - r0), dim3(kRegionCap / 2), 0, a, r0);
  if (meshes == st->meshes) {
    if (st->bigMeshes)
      launch(ctx, s, "lbvh_build_top", k_build_top, dim3(st->bigMeshes), dim3(kTopCap / 2), 0, a, 1u, 0u);
  } else {
    for (uint32_t m = meshBase; m < meshBase + meshes; ++m)
      if (st->desc[m].ntri > kBuildRegion)
        launch(ctx, s, "lbvh_build_top", k_build_top, dim3(1), dim3(kTopCap / 2), 0, a, 0u, m);
  }
}

// Extra dynamic LDS that caps a sort pass at `wavesPerCU` resident waves per CU (160 KB of LDS
// per CU shared by the workgroups); 0 when no cap is asked or the cap is above the natural one.
template <int BITS, int THREADS, int ITEMS>
size_t sort_lds_pad(int wavesPerCU) {
  if (wavesPerCU <= 0) return 0;
  constexpr int R = 1 << BITS, W = THREADS / 64;
  const size_t natural = (size_t)(2 * THREADS * ITEMS + 2 * R + W + 1) * 4;
  const int wgs = std::max(1, wavesPerCU / W);
  const size_t target = std::min<size_t>((160u * 1024u) / wgs, 64u * 1024u);
  return target > natural ? target - natural : 0;
}

template <int BITS>
void launch_sort_passes(Context *ctx, hipStream_t s, BVHStore *st, const BuildArgs &a) {
  const int wpc = st->cfg.wavesPerCU;
  for (int p = 0; p < st->passes; ++p) {
    if (st->cfg.tileKeys == 2048)
      launch(ctx, s, "lbvh_sort_pass", k_onesweep<BITS, 256, 8>, dim3(st->tiles), dim3(256),
             sort_lds_pad<BITS, 256, 8>(wpc), a, p);
    else if (st->cfg.tileKeys == 4096)
      launch(ctx, s, "lbvh_sort_pass", k_onesweep<BITS, 256, 16>, dim3(st->tiles), dim3(256),
             sort_lds_pad<BITS, 256, 16>(wpc), a, p);
    else if (st->cfg.tileKeys == 16384)  // one 130 KB workgroup per CU
      launch(ctx, s, "lbvh_sort_pass", k_onesweep<BITS, 1024, 16>, dim3(st->tiles), dim3(1024), 0, a, p);
    else
      launch(ctx, s, "lbvh_sort_pass", k_onesweep<BITS, 512, 16>, dim3(st->tiles), dim3(512),
             sort_lds_pad<BITS, 512, 16>(wpc), a, p);
  }
}

// The whole build of every mesh in the store: no host synchronisation inside, so it can be
// captured into a hipGraph.
void enqueue_build(Context *ctx, hipStream_t s, BVHStore *st) {
  const BuildArgs a = args_of(st);
  const int R = 1 << st->cfg.digitBits;
  launch(ctx, s, "lbvh_init", k_init, dim3(grid_for(std::max<size_t>(st->zeroWords, st->meshes * 16ull), 256, 4096)),
         dim3(256), 0, st->zeroBase, st->zeroWords, st->headers, st->meshes);
  launch(ctx, s, "lbvh_bounds", k_bounds, dim3(st->tiles), dim3(kBoundsThreads), 0, a, 0u);
  const size_t histLds = (size_t)st->passes * R * 4;
  switch (st->cfg.digitBits) {
    case 4:
      launch(ctx, s, "lbvh_keys", k_keys<4>, dim3(st->tiles), dim3(kKeysThreads), histLds, a, 0u);
      launch(ctx, s, "lbvh_hist_scan", k_hist_scan<4>, dim3(st->meshes), dim3(256), 0, a);
      launch_sort_passes<4>(ctx, s, st, a);
      break;
    case 6:
      launch(ctx, s, "lbvh_keys", k_keys<6>, dim3(st->tiles), dim3(kKeysThreads), histLds, a, 0u);
      launch(ctx, s, "lbvh_hist_scan", k_hist_scan<6>, dim3(st->meshes), dim3(256), 0, a);
      launch_sort_passes<6>(ctx, s, st, a);
      break;
    default:
      launch(ctx, s, "lbvh_keys", k_keys<8>, dim3(st->tiles), dim3(kKeysThreads), histLds, a, 0u);
      launch(ctx, s, "lbvh_hist_scan", k_hist_scan<8>, dim3(st->meshes), dim3(256), 0, a);
      launch_sort_passes<8>(ctx, s, st, a);
      break;
  }
  if (st->passes & 1)
    launch(ctx, s, "lbvh_copy_back", k_copy_back, dim3(grid_for(st->prims, 256, 8192)), dim3(256), 0, a, st->prims);
  launch_refit(ctx, s, st, a, 0u, st->tiles, 0u, st->meshes, true);
}

bool same_plan(const BVHStore *st, const std::vector<MeshDesc> &desc, const SortConfig &cfg) {
  if (st->cfg.digitBits != cfg.digitBits || st->cfg.tileKeys != cfg.tileKeys || st->cfg.wavesPerCU != cfg.wavesPerCU)
    return false;
  if (st->desc.size() != desc.size()) return false;
  return std::memcmp(st->desc.data(), desc.data(), desc.size() * sizeof(MeshDesc)) == 0;
}

void free_block(BVHStore *st) {
  if (st->graph) {
    hipGraphExecDestroy(st->graph);
    st->graph = nullptr;
  }
  st->graphValid = false;
  if (st->block) hipFree(st->block);
  st->block = nullptr;
  st->blockBytes = 0;
}

// Lays the arrays of `st` out for the given plan (reusing the allocation when it is big enough).
bool store_layout(BVHStore *st, const std::vector<MeshDesc> &desc, const std::vector<uint32_t> &tileMesh,
                  const SortConfig &cfg, uint32_t prims) {
  const uint32_t M = (uint32_t)desc.size(), T = (uint32_t)tileMesh.size();
  const uint32_t I = prims - M;  // every mesh has ntri - 1 inner nodes
  const int R = 1 << cfg.digitBits;
  const int passes = (kMortonBits + cfg.digitBits - 1) / cfg.digitBits;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t oDesc = off;
  off += al((size_t)M * sizeof(MeshDesc));
  // tile -> mesh, then the region table (first tile of each region) and the meshes with a top
  // level: one upload
  std::vector<uint32_t> regionTile, bigMesh, regionFirst(M + 1, 0u);
  for (uint32_t m = 0; m < M; ++m) {
    const uint32_t regions = (desc[m].ntri + kBuildRegion - 1) / kBuildRegion;
    regionFirst[m] = (uint32_t)regionTile.size();
    for (uint32_t r = 0; r < regions; ++r) regionTile.push_back(desc[m].tileOffset + r * (kBuildRegion / cfg.tileKeys));
    if (desc[m].ntri > kBuildRegion) bigMesh.push_back(m);
  }
  regionFirst[M] = (uint32_t)regionTile.size();
  std::vector<uint32_t> plan(tileMesh);
  plan.insert(plan.end(), regionTile.begin(), regionTile.end());
  plan.insert(plan.end(), bigMesh.begin(), bigMesh.end());
  const size_t oTile = off;
  off += al(plan.size() * 4);
  const size_t oCodes = off;
  off += al((size_t)prims * 4);
  const size_t oIds = off;
  off += al((size_t)prims * 4);
  const size_t oCodesT = off;
  off += al((size_t)prims * 4);
  const size_t oIdsT = off;
  off += al((size_t)prims * 4);
  const size_t oNodes = off;
  off += al((size_t)std::max<uint32_t>(I, 1) * sizeof(BVHNode));
  const size_t oTris = off;
  off += al((size_t)prims * sizeof(TriRecord));
  const size_t oCent = off;
  off += al((size_t)prims * 12);
  const size_t oPL = off;
  off += al((size_t)prims * 4);
  const size_t oPI = off;
  off += al((size_t)std::max<uint32_t>(I, 1) * 4);
  const size_t oHdr = off;
  off += al((size_t)M * 16 * 4);
  const size_t oList0 = off;
  off += al((size_t)prims * sizeof(Subtree));
  const size_t oList1 = off;
  off += al((size_t)prims * sizeof(Subtree));
  const size_t oListD0 = off;
  off += al((size_t)prims);
  const size_t oListD1 = off;
  off += al((size_t)prims);
  const size_t oZero = off;
  const size_t histWords = (size_t)M * passes * R;
  const size_t statWords = (size_t)passes * T * R;
  const size_t tickWords = 64;
  const size_t subWords = (size_t)T * ((size_t)cfg.tileKeys / kBuildSub);
  const size_t regionWords = T;
  const size_t zeroWords = histWords + statWords + tickWords + subWords + regionWords;
  off += al(zeroWords * 4);
  if (off > st->blockBytes) {
    free_block(st);
    if (!hip_ok(hipMalloc(&st->block, off), "hipMalloc(lbvh)")) return false;
    st->blockBytes = off;
  }
  char *b = (char *)st->block;
  st->meshes = M;
  st->prims = prims;
  st->inner = I;
  st->tiles = T;
  st->cfg = cfg;
  st->passes = passes;
  st->dDesc = (MeshDesc *)(b + oDesc);
  st->dTileMesh = (uint32_t *)(b + oTile);
  st->codes = (uint32_t *)(b + oCodes);
  st->ids = (uint32_t *)(b + oIds);
  st->codesTmp = (uint32_t *)(b + oCodesT);
  st->idsTmp = (uint32_t *)(b + oIdsT);
  st->nodes = b + oNodes;
  st->tris = b + oTris;
  st->centroids = b + oCent;
  st->parentLeaf = (uint32_t *)(b + oPL);
  st->parentInner = (uint32_t *)(b + oPI);
  st->headers = (float *)(b + oHdr);
  st->list0 = b + oList0;
  st->list1 = b + oList1;
  st->listD0 = b + oListD0;
  st->listD1 = b + oListD1;
  st->zeroBase = (uint32_t *)(b + oZero);
  st->zeroWords = zeroWords;
  st->hist = st->zeroBase;
  st->status = st->hist + histWords;
  st->tickets = st->status + statWords;
  st->subCount = st->tickets + tickWords;
  st->regionCount = st->subCount + subWords;
  st->desc = desc;
  st->tileMesh = plan;
  st->regions = (uint32_t)regionTile.size();
  st->bigMeshes = (uint32_t)bigMesh.size();
  st->regionFirst = regionFirst;
  return true;
}

// The geometry fields of a mesh descriptor (where the primitives are read from).
void geom_fields(const Accel *acc, MeshDesc &d) {
  const Geom *g = acc->geom;
  d.verts = (const char *)g->vbuf->d + g->voffset;
  d.vstride = g->vstride;
  d.nverts = g->vcount;
  d.ntri = g->icount;
  d.noRecords = acc->recordsExternal ? 1u : 0u;
  if (acc->kind == AccelKind::AABBs || acc->kind == AccelKind::Points) {
    d.kind = acc->kind == AccelKind::AABBs ? kPrimBoxes : kPrimPoints;  // read at verts + i * vstride
    d.idx = nullptr;
    d.istride = 0;
  } else {
    d.kind = acc->kind == AccelKind::Edges ? kPrimEdges : kPrimTriangles;
    d.idx = (const char *)g->ibuf->d + g->ioffset;
    d.istride = g->istride;
  }
}

void set_views(Accel *a) {
  BVHStore *st = a->store;
  const MeshDesc &d = st->desc[a->mesh];
  a->n = d.ntri;
  a->nodes = (char *)st->nodes + (size_t)d.innerOffset * sizeof(BVHNode);
  a->tris = (char *)st->tris + (size_t)d.primOffset * sizeof(TriRecord);
  a->header = st->headers + (size_t)a->mesh * 16;
}

}  // namespace

void store_release(BVHStore *st) {
  if (!st) return;
  free_block(st);
  delete st;
}

void accel_detach(Accel *a) {
  if (!a->store) return;
  BVHStore *st = a->store;
  a->store = nullptr;
  a->built = false;
  blas_desc_refresh(a->ctx, a->ctx->stream, a);  // instance accels stop reading this storage
  if (--st->refs == 0) {
    hipStreamSynchronize(a->ctx->stream);
    store_release(st);
  }
}

// The automatic tile size: 16384-key tiles (one 1024-thread sort workgroup per CU) for builds
// averaging at least 2^20 primitives per mesh, whose sort passes then run a single short chain of
// tiles (4,194,304 triangles: 33 instead of 38 us per pass); 8192 otherwise, where more tiles
// per mesh keep the look-backs short.
static int auto_tile_keys(Accel **accels, uint32_t count) {
  uint64_t prims = 0;
  for (uint32_t m = 0; m < count; ++m) prims += accels[m]->geom->icount;
  return prims >= (uint64_t)count << 20 ? 16384 : 8192;
}

bool lbvh_build_batch(Context *ctx, hipStream_t s, Accel **accels, uint32_t count, bool useGraph,
                      bool layoutOnly) {
  SortConfig cfg = ctx->sortCfg;
  if (cfg.tileKeys == 0) cfg.tileKeys = auto_tile_keys(accels, count);
  const uint32_t tileKeys = (uint32_t)cfg.tileKeys;
  std::vector<MeshDesc> desc(count);
  std::vector<uint32_t> tileMesh;
  uint64_t prims = 0, tiles = 0;
  for (uint32_t m = 0; m < count; ++m) {
    Geom *g = accels[m]->geom;
    MeshDesc &d = desc[m];
    std::memset(&d, 0, sizeof(d));
    geom_fields(accels[m], d);
    d.primOffset = (uint32_t)prims;
    d.innerOffset = (uint32_t)(prims - m);
    d.tileOffset = (uint32_t)tiles;
    d.ntiles = (g->icount + tileKeys - 1) / tileKeys;
    prims += g->icount;
    tiles += d.ntiles;
  }
  if (prims >= (1ull << 30)) {
    fail("gprtAccelBuild: %llu triangles in one build exceeds the 2^30 limit", (unsigned long long)prims);
    return false;
  }
  tileMesh.resize(tiles);
  for (uint32_t m = 0; m < count; ++m)
    for (uint32_t t = 0; t < desc[m].ntiles; ++t) tileMesh[desc[m].tileOffset + t] = m;

  // Reuse the store when exactly these accels own it and the plan is unchanged: the cached graph
  // then replays the whole build with one launch.
  BVHStore *st = accels[0]->store;
  bool reuse = st && st->refs == (int)count;
  for (uint32_t m = 0; reuse && m < count; ++m) reuse = accels[m]->store == st && accels[m]->mesh == m;
  const bool samePlan = reuse && same_plan(st, desc, cfg);
  if (!reuse) {
    for (uint32_t m = 0; m < count; ++m) accel_detach(accels[m]);
    st = new BVHStore();
    st->refs = (int)count;
    for (uint32_t m = 0; m < count; ++m) {
      accels[m]->store = st;
      accels[m]->mesh = m;
    }
  }
  if (!samePlan) {
    if (st->graph) {
      hipGraphExecDestroy(st->graph);
      st->graph = nullptr;
    }
    st->graphValid = false;
    if (!store_layout(st, desc, tileMesh, cfg, (uint32_t)prims)) return false;
    if (!hip_ok(hipMemcpyAsync(st->dDesc, st->desc.data(), desc.size() * sizeof(MeshDesc), hipMemcpyHostToDevice, s),
                "hipMemcpyAsync(mesh table)"))
      return false;
    if (!hip_ok(hipMemcpyAsync(st->dTileMesh, st->tileMesh.data(), st->tileMesh.size() * 4, hipMemcpyHostToDevice, s),
                "hipMemcpyAsync(tile table)"))
      return false;
  }
  for (uint32_t m = 0; m < count; ++m) {
    set_views(accels[m]);
    if (!blas_desc_refresh(ctx, s, accels[m])) return false;
    if (accels[m]->slot != 0xFFFFFFFFu) ++ctx->blasContentEpoch;
  }
  if (layoutOnly) return true;  // the caller fills the arrays (a deserialised accel)

  if (useGraph) {
    if (!st->graphValid) {
      // The table uploads above must land before capture starts.
      if (!hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize(pre-capture)")) return false;
      const bool timer = ctx->timer.enabled;
      ctx->timer.enabled = false;
      hipGraph_t graph = nullptr;
      if (!hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) {
        ctx->timer.enabled = timer;
        return false;
      }
      enqueue_build(ctx, s, st);
      const bool ok = hip_ok(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
      ctx->timer.enabled = timer;
      if (!ok) return false;
      if (!hip_ok(hipGraphInstantiate(&st->graph, graph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
        hipGraphDestroy(graph);
        return false;
      }
      hipGraphDestroy(graph);
      st->graphValid = true;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timer.enabled) {
      e0 = ctx->timer.get();
      e1 = ctx->timer.get();
      hipEventRecord(e0, s);
    }
    if (!hip_ok(hipGraphLaunch(st->graph, s), "hipGraphLaunch")) return false;
    if (ctx->timer.enabled) {
      hipEventRecord(e1, s);
      ctx->timer.pending.push_back({"lbvh_batch_graph", e0, e1});
    }
  } else {
    enqueue_build(ctx, s, st);
  }
  for (uint32_t m = 0; m < count; ++m) accels[m]->built = true;
  return true;
}

bool lbvh_refit(Context *ctx, hipStream_t s, Accel *acc) {
  BVHStore *st = acc->store;
  MeshDesc &d = st->desc[acc->mesh];
  // The geometry may sit in another buffer, or the same buffer at a new address (a resize),
  // since the build: the refit reads it where it is now.
  MeshDesc now = d;
  geom_fields(acc, now);
  if (std::memcmp(&now, &d, sizeof(d)) != 0) {
    d = now;
    if (!hip_ok(hipMemcpyAsync(st->dDesc + acc->mesh, &st->desc[acc->mesh], sizeof(MeshDesc), hipMemcpyHostToDevice,
                               s),
                "hipMemcpyAsync(mesh descriptor)"))
      return false;
    if (!blas_desc_refresh(ctx, s, acc)) return false;
  }
  const BuildArgs a = args_of(st);
  if (acc->slot != 0xFFFFFFFFu) ++ctx->blasContentEpoch;
  // The index buffer may have changed since the build: its identity layout is checked again (the
  // word reset by a kernel rather than a memset, so a captured refit replays correctly).
  const bool tris = d.kind == kPrimTriangles;
  if (tris) {
    launch(ctx, s, "lbvh_ident", k_ident_reset, dim3(1), dim3(64), 0, st->headers, acc->mesh);
    launch(ctx, s, "lbvh_ident", k_ident, dim3(d.ntiles), dim3(kIdentThreads), 0, a, d.tileOffset, acc->mesh);
  }
  launch_refit(ctx, s, st, a, d.tileOffset, d.ntiles, acc->mesh, 1u, tris);
  return true;
}

void lbvh_parents(Context *ctx, hipStream_t s, Accel *acc) {
  BVHStore *st = acc->store;
  const MeshDesc &d = st->desc[acc->mesh];
  launch(ctx, s, "lbvh_parents", k_parents, dim3(grid_for(d.ntri, 256, 1u << 22)), dim3(256), 0,
         (const BVHNode *)acc->nodes, d.ntri, st->parentLeaf + d.primOffset, st->parentInner + d.innerOffset);
}

bool lbvh_export(Context *ctx, hipStream_t s, Accel *acc, int32_t *nodes, float *aabbs, uint64_t *keys) {
  BVHStore *st = acc->store;
  const MeshDesc &d = st->desc[acc->mesh];
  const uint32_t n = d.ntri;
  lbvh_parents(ctx, s, acc);
  launch(ctx, s, "lbvh_export", k_export, dim3(grid_for(2 * (size_t)n - 1, 256, 1u << 22)), dim3(256), 0,
         (const BVHNode *)acc->nodes, (const uint32_t *)(st->codes + d.primOffset),
         (const uint32_t *)(st->ids + d.primOffset), (const uint32_t *)(st->parentLeaf + d.primOffset),
         (const uint32_t *)(st->parentInner + d.innerOffset), (const float *)acc->header, n, (int4 *)nodes, aabbs,
         keys);
  return true;
}

}  // namespace gprt
