// Triangle LBVH build for gfx950: scene bounds, Morton keys, key sort, Karras split search and
// bottom-up refit, following the stage order of GPRT's development LBVH
// (gprt/development/gprt_lbvh.h:170-195; kernels gprt/development/gprt_lbvh.hlsl:215-428).
//
// Differences in form (results are the same binary radix tree and the same exact boxes):
//  * Scene bounds: wave/block min-max reduction, then one ordered-uint atomic per block and
//    axis, instead of six float CAS atomics per triangle on one address set (hlsl:215-226).
//  * Keys: the reference sorts 32-bit Morton codes with primID payload; ties are broken by the
//    primitive index (hlsl:56-58). Here keys are defined as morton30<<32 | primID. Because the
//    primIDs enter the sort in ascending order, a stable sort of the 30 Morton bits with primID
//    payload yields exactly the order of the 64-bit keys, at 4 passes of 8 bits instead of 8.
//  * Karras: delta(i,j) = clz64(key_i ^ key_j), -1 outside [0, N-1] (hlsl:45-59); the split
//    search is the binary search of hlsl:84-97 without its trailing t=1 iterations (they cannot
//    move s once the search has converged).
//  * Refit: one thread per leaf climbs while it is the second child to arrive at a node (atomic
//    arrival counter, acquire/release at device scope), so every node is combined exactly once:
//    O(N) work, versus every leaf doing six CAS atomics on every ancestor (hlsl:400-428).
#include "bvh.hpp"
#include "device_common.hpp"
#include "internal.hpp"

namespace gprt {

namespace {

struct GeomView {
  const char *verts;
  uint32_t vstride;
  const char *idx;
  uint32_t istride;
  uint32_t nverts;
};

// Vertex indices are clamped to the vertex count, so a malformed index buffer yields a wrong
// box rather than an out-of-bounds read.
__device__ __forceinline__ float3 load_vertex(const GeomView &g, uint32_t i) {
  i = i < g.nverts ? i : g.nverts - 1;
  const float *p = (const float *)(g.verts + (size_t)i * g.vstride);
  return make_float3(p[0], p[1], p[2]);
}

__device__ __forceinline__ uint3 load_tri(const GeomView &g, uint32_t prim) {
  const uint32_t *p = (const uint32_t *)(g.idx + (size_t)prim * g.istride);
  return make_uint3(p[0], p[1], p[2]);
}

// Zero refit counters; set the scene box accumulators to the identities of min/max.
__global__ void __launch_bounds__(256) k_build_init(uint32_t *__restrict__ counters, uint32_t nInner,
                                                    uint32_t *__restrict__ header) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  if (i < 16) header[i] = (i < 3) ? 0xFFFFFFFFu : 0u;
  for (; i < nInner; i += stride) counters[i] = 0u;
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

// Scene box over all vertices of all triangles (gprt_lbvh.hlsl:121-131, :215-226).
__global__ void __launch_bounds__(256) k_bounds(GeomView g, uint32_t n, uint32_t *__restrict__ header) {
  float mn[3] = {INFINITY, INFINITY, INFINITY};
  float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint3 t = load_tri(g, (uint32_t)i);
    const float3 a = load_vertex(g, t.x), b = load_vertex(g, t.y), c = load_vertex(g, t.z);
    mn[0] = fmin_ord(mn[0], fmin_ord(a.x, fmin_ord(b.x, c.x)));
    mn[1] = fmin_ord(mn[1], fmin_ord(a.y, fmin_ord(b.y, c.y)));
    mn[2] = fmin_ord(mn[2], fmin_ord(a.z, fmin_ord(b.z, c.z)));
    mx[0] = fmax_ord(mx[0], fmax_ord(a.x, fmax_ord(b.x, c.x)));
    mx[1] = fmax_ord(mx[1], fmax_ord(a.y, fmax_ord(b.y, c.y)));
    mx[2] = fmax_ord(mx[2], fmax_ord(a.z, fmax_ord(b.z, c.z)));
  }
  __shared__ float s[4][6];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    mn[a] = wave_min(mn[a]);
    mx[a] = wave_max(mx[a]);
  }
  if (l == 0) {
    for (int a = 0; a < 3; ++a) {
      s[w][a] = mn[a];
      s[w][3 + a] = mx[a];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    float v = s[0][a];
    for (int ww = 1; ww < 4; ++ww) v = (a < 3) ? fmin_ord(v, s[ww][a]) : fmax_ord(v, s[ww][a]);
    if (a < 3)
      atomicMin(&header[a], ford(v));
    else
      atomicMax(&header[a], ford(v));
  }
}

__device__ __forceinline__ float unord(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Morton keys of triangle centroids (gprt_lbvh.hlsl:139-146, :262-277). Keys are written as the
// 30-bit code with the primID as payload; ids enter the sort in ascending order.
__global__ void __launch_bounds__(256) k_keys(GeomView g, uint32_t n, const uint32_t *__restrict__ header,
                                              uint32_t *__restrict__ codes, uint32_t *__restrict__ ids) {
  const float lox = unord(header[0]), loy = unord(header[1]), loz = unord(header[2]);
  const float hix = unord(header[3]), hiy = unord(header[4]), hiz = unord(header[5]);
  const float ex = hix - lox, ey = hiy - loy, ez = hiz - loz;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint3 t = load_tri(g, (uint32_t)i);
    const float3 a = load_vertex(g, t.x), b = load_vertex(g, t.y), c = load_vertex(g, t.z);
    const float cx = (a.x + b.x + c.x) / 3.0f;
    const float cy = (a.y + b.y + c.y) / 3.0f;
    const float cz = (a.z + b.z + c.z) / 3.0f;
    const float px = (cx - lox) / ex, py = (cy - loy) / ey, pz = (cz - loz) / ez;
    const float qx = fminf(fmaxf(px * 1024.0f, 0.0f), 1023.0f);
    const float qy = fminf(fmaxf(py * 1024.0f, 0.0f), 1023.0f);
    const float qz = fminf(fmaxf(pz * 1024.0f, 0.0f), 1023.0f);
    codes[i] = morton30((uint32_t)qx, (uint32_t)qy, (uint32_t)qz);
    ids[i] = (uint32_t)i;
  }
}

__device__ __forceinline__ uint64_t key_at(const uint32_t *__restrict__ codes, const uint32_t *__restrict__ ids,
                                           int64_t j) {
  return ((uint64_t)codes[j] << 32) | ids[j];
}

__device__ __forceinline__ int delta64(const uint32_t *__restrict__ codes, const uint32_t *__restrict__ ids,
                                       uint64_t ki, int64_t j, int64_t n) {
  if (j < 0 || j >= n) return -1;
  return __clzll(ki ^ key_at(codes, ids, j));
}

// Karras 2012 split search (gprt_lbvh.hlsl:64-103, :287-338). One thread per inner node.
__global__ void __launch_bounds__(256) k_karras(const uint32_t *__restrict__ codes, const uint32_t *__restrict__ ids,
                                                uint32_t n, BVHNode *__restrict__ nodes,
                                                uint32_t *__restrict__ parentLeaf,
                                                uint32_t *__restrict__ parentInner) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = n;
  if (i >= N - 1) return;
  const uint64_t ki = key_at(codes, ids, i);
  const int d = delta64(codes, ids, ki, i + 1, N) >= delta64(codes, ids, ki, i - 1, N) ? 1 : -1;
  const int dmin = delta64(codes, ids, ki, i - d, N);
  int64_t lmax = 2;
  while (delta64(codes, ids, ki, i + lmax * d, N) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
    if (delta64(codes, ids, ki, i + (l + t) * d, N) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta64(codes, ids, ki, j, N);
  int64_t s = 0;
  int64_t t = l;
  do {
    t = (t + 1) >> 1;
    if (delta64(codes, ids, ki, i + (s + t) * d, N) > dnode) s += t;
  } while (t > 1);
  const int64_t split = i + s * d + (d < 0 ? -1 : 0);
  const int64_t first = d > 0 ? i : j;
  const int64_t last = d > 0 ? j : i;
  const int64_t left = split, right = split + 1;
  int c0, c1;
  if (left == first) {
    c0 = ~(int)left;
    parentLeaf[left] = ((uint32_t)i << 1) | 0u;
  } else {
    c0 = (int)left;
    parentInner[left] = ((uint32_t)i << 1) | 0u;
  }
  if (right == last) {
    c1 = ~(int)right;
    parentLeaf[right] = ((uint32_t)i << 1) | 1u;
  } else {
    c1 = (int)right;
    parentInner[right] = ((uint32_t)i << 1) | 1u;
  }
  nodes[i].d = make_int4(c0, c1, 0, 0);
  if (i == 0) parentInner[0] = 0xFFFFFFFFu;
}

__device__ __forceinline__ void store_child_box(BVHNode *nodes, uint32_t node, uint32_t side, const float b[6]) {
  float2 *p = (float2 *)((float *)&nodes[node] + side * 6);
  p[0] = make_float2(b[0], b[1]);
  p[1] = make_float2(b[2], b[3]);
  p[2] = make_float2(b[4], b[5]);
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

// Leaf records + bottom-up refit with arrival counters. One thread per sorted leaf position.
__global__ void __launch_bounds__(256) k_refit(GeomView g, const uint32_t *__restrict__ ids, uint32_t n,
                                               BVHNode *__restrict__ nodes, TriRecord *__restrict__ tris,
                                               const uint32_t *__restrict__ parentLeaf,
                                               const uint32_t *__restrict__ parentInner,
                                               uint32_t *__restrict__ counters, float *__restrict__ header) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t prim = ids[k];
  const uint3 t = load_tri(g, prim);
  const float3 a = load_vertex(g, t.x), b = load_vertex(g, t.y), c = load_vertex(g, t.z);
  TriRecord rec;
  rec.v0 = make_float4(a.x, a.y, a.z, __uint_as_float(prim));
  rec.v1 = make_float4(b.x, b.y, b.z, 0.0f);
  rec.v2 = make_float4(c.x, c.y, c.z, 0.0f);
  tris[k] = rec;
  float box[6];
  box[0] = fmin_ord(a.x, fmin_ord(b.x, c.x));
  box[1] = fmin_ord(a.y, fmin_ord(b.y, c.y));
  box[2] = fmin_ord(a.z, fmin_ord(b.z, c.z));
  box[3] = fmax_ord(a.x, fmax_ord(b.x, c.x));
  box[4] = fmax_ord(a.y, fmax_ord(b.y, c.y));
  box[5] = fmax_ord(a.z, fmax_ord(b.z, c.z));
  if (n == 1) {
    for (int q = 0; q < 6; ++q) header[8 + q] = box[q];
    return;
  }
  uint32_t link = parentLeaf[k];
  // A radix tree over distinct 64-bit keys is at most 65 levels deep; the bound and the index
  // checks only guard against a malformed topology.
  for (int level = 0; level < 72; ++level) {
    const uint32_t node = link >> 1, side = link & 1u;
    if (node >= n - 1) return;
    store_child_box(nodes, node, side, box);
    const uint32_t old =
        __hip_atomic_fetch_add(&counters[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old != 1u) return;  // first arrival: the sibling's thread finishes this node
    float sib[6];
    load_child_box(nodes, node, side ^ 1u, sib);
    box[0] = fmin_ord(box[0], sib[0]);
    box[1] = fmin_ord(box[1], sib[1]);
    box[2] = fmin_ord(box[2], sib[2]);
    box[3] = fmax_ord(box[3], sib[3]);
    box[4] = fmax_ord(box[4], sib[4]);
    box[5] = fmax_ord(box[5], sib[5]);
    link = parentInner[node];
    if (link == 0xFFFFFFFFu) {
      for (int q = 0; q < 6; ++q) header[8 + q] = box[q];
      return;
    }
  }
}

// Reference-layout export (gprt/development/gprt_lbvh.h:28-55).
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
    outKeys[k] = ((uint64_t)codes[k] << 32) | ids[k];
  }
  for (int q = 0; q < 6; ++q) outBoxes[(size_t)x * 6 + q] = b[q];
}

GeomView geom_view(Geom *g) {
  GeomView v;
  v.verts = (const char *)g->vbuf->d + g->voffset;
  v.vstride = g->vstride;
  v.idx = (const char *)g->ibuf->d + g->ioffset;
  v.istride = g->istride;
  v.nverts = g->vcount;
  return v;
}

unsigned grid_for(size_t n, unsigned block, unsigned cap) {
  size_t b = (n + block - 1) / block;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

bool lbvh_reserve(Context *ctx, LBVHStorage &st, uint32_t n) {
  const SortConfig cfg = ctx->sortCfg;
  const size_t nn0 = n ? n : 1;
  if (n <= st.capacity && st.block && sort_scratch_bytes(nn0, 32, 32, 0, 30, cfg) <= st.sortScratchBytes) {
    st.n = n;
    return true;
  }
  lbvh_release(st);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t nn = n ? n : 1;
  const size_t inner = nn > 1 ? nn - 1 : 1;
  const size_t sortBytes = sort_scratch_bytes(nn, 32, 32, 0, 30, cfg);
  size_t off = 0;
  const size_t oCodes = off;
  off += al(nn * 4);
  const size_t oIds = off;
  off += al(nn * 4);
  const size_t oNodes = off;
  off += al(inner * sizeof(BVHNode));
  const size_t oTris = off;
  off += al(nn * sizeof(TriRecord));
  const size_t oPL = off;
  off += al(nn * 4);
  const size_t oPI = off;
  off += al(inner * 4);
  const size_t oCnt = off;
  off += al(inner * 4);
  const size_t oHdr = off;
  off += al(kHeaderFloats * 4);
  const size_t oKT = off;
  off += al(nn * 4);
  const size_t oIT = off;
  off += al(nn * 4);
  const size_t oSS = off;
  off += al(sortBytes);
  if (!hip_ok(hipMalloc(&st.block, off), "hipMalloc(lbvh)")) return false;
  char *b = (char *)st.block;
  st.blockBytes = off;
  st.capacity = (uint32_t)nn;
  st.mcodes = (uint32_t *)(b + oCodes);
  st.ids = (uint32_t *)(b + oIds);
  st.nodes = b + oNodes;
  st.tris = b + oTris;
  st.parentLeaf = (uint32_t *)(b + oPL);
  st.parentInner = (uint32_t *)(b + oPI);
  st.counters = (uint32_t *)(b + oCnt);
  st.header = (float *)(b + oHdr);
  st.keysTmp = (uint32_t *)(b + oKT);
  st.idsTmp = (uint32_t *)(b + oIT);
  st.sortScratch = b + oSS;
  st.sortScratchBytes = sortBytes;
  st.n = n;
  return true;
}

void lbvh_release(LBVHStorage &st) {
  if (st.block) hipFree(st.block);
  st = LBVHStorage();
}

bool lbvh_build(Context *ctx, hipStream_t s, Accel *a) {
  Geom *g = a->geom;
  const uint32_t n = g->icount;
  if (!lbvh_reserve(ctx, a->bvh, n)) return false;
  LBVHStorage &st = a->bvh;
  const GeomView gv = geom_view(g);
  const uint32_t inner = n > 1 ? n - 1 : 0;
  launch(ctx, s, "lbvh_init", k_build_init, dim3(grid_for(inner > 16 ? inner : 16, 256, 2048)), dim3(256), 0,
         st.counters, inner, (uint32_t *)st.header);
  launch(ctx, s, "lbvh_bounds", k_bounds, dim3(grid_for(n, 256 * 8, (unsigned)ctx->cuCount * 8)), dim3(256), 0, gv,
         n, (uint32_t *)st.header);
  launch(ctx, s, "lbvh_keys", k_keys, dim3(grid_for(n, 256, 1u << 20)), dim3(256), 0, gv, n,
         (const uint32_t *)st.header, st.mcodes, st.ids);
  if (n > 1) {
    if (!sort_run(ctx, s, st.mcodes, st.ids, st.keysTmp, st.idsTmp, n, 32, 32, 0, 30, st.sortScratch,
                  ctx->sortCfg))
      return false;
    launch(ctx, s, "lbvh_karras", k_karras, dim3(grid_for(inner, 256, 1u << 22)), dim3(256), 0,
           (const uint32_t *)st.mcodes, (const uint32_t *)st.ids, n, (BVHNode *)st.nodes, st.parentLeaf,
           st.parentInner);
  }
  launch(ctx, s, "lbvh_refit", k_refit, dim3(grid_for(n, 256, 1u << 22)), dim3(256), 0, gv,
         (const uint32_t *)st.ids, n, (BVHNode *)st.nodes, (TriRecord *)st.tris, (const uint32_t *)st.parentLeaf,
         (const uint32_t *)st.parentInner, st.counters, st.header);
  return true;
}

bool lbvh_refit(Context *ctx, hipStream_t s, Accel *a) {
  LBVHStorage &st = a->bvh;
  const uint32_t n = st.n;
  const GeomView gv = geom_view(a->geom);
  const uint32_t inner = n > 1 ? n - 1 : 0;
  if (inner) fill_u32(ctx, s, st.counters, 0u, inner);
  launch(ctx, s, "lbvh_refit", k_refit, dim3(grid_for(n, 256, 1u << 22)), dim3(256), 0, gv,
         (const uint32_t *)st.ids, n, (BVHNode *)st.nodes, (TriRecord *)st.tris, (const uint32_t *)st.parentLeaf,
         (const uint32_t *)st.parentInner, st.counters, st.header);
  return true;
}

bool lbvh_export(Context *ctx, hipStream_t s, Accel *a, int32_t *nodes, float *aabbs, uint64_t *keys) {
  LBVHStorage &st = a->bvh;
  const uint32_t n = st.n;
  launch(ctx, s, "lbvh_export", k_export, dim3(grid_for(2 * (size_t)n - 1, 256, 1u << 22)), dim3(256), 0,
         (const BVHNode *)st.nodes, (const uint32_t *)st.mcodes, (const uint32_t *)st.ids,
         (const uint32_t *)st.parentLeaf, (const uint32_t *)st.parentInner, (const float *)st.header, n,
         (int4 *)nodes, aabbs, keys);
  return true;
}

}  // namespace gprt
