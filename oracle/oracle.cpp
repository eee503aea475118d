// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of GPRT's software acceleration-structure path, used by tests/ as the checker
// and by bench.py as the timed CPU baseline. Nothing in the product (gprt_amd/) links, loads or
// calls this file. Built with -O2 -ffp-contract=off so its FP32 arithmetic rounds exactly as
// written.
//
// Each function cites the reference behaviour it restates (paths relative to the GPRT tree):
//   sort      tests/t02-bufferSort/hostCode.cpp:59,115 (std::sort) and the stable in-place LSD
//             semantics of gprt/gprt.cpp:8821-9073 -> std::stable_sort by key, payload follows.
//   scan      tests/t03-bufferScans/hostCode.cpp:53-58 (exclusive sum), :114-137 (partition),
//             :194-211 (select); gprt/gprt_scan.slang:84-92 (forward / reverse addresses).
//   morton    gprt/development/gprt_lbvh.hlsl:25-38 (separate_bits, morton_encode3D),
//             :121-146 (triangle bounds / centroid), :262-277 (quantisation).
//   karras    gprt/development/gprt_lbvh.hlsl:40-103 (delta, determine_range incl. the float
//             `divf` split loop) and :287-338 (child / parent links), keyed on 64-bit keys.
//   refit     gprt/development/gprt_lbvh.hlsl:400-428 (every leaf folds its box into every
//             ancestor; order-free min/max).
//   layout    gprt/development/gprt_lbvh.h:28-55 (int4 nodes, float3 pairs).
//   points / edges  gprt/development/gprt_lbvh.hlsl:105-119, :131-137, :188-213, :228-260 (their
//             boxes and key points), built by the same stages as triangles.
//   trace     closest hit with the open interval tmin < t (gprt/gprt_fallbacks.slang:108,139);
//             triangle test = watertight ray/triangle intersection (Woop, Benthin, Wald, JCGT
//             2013), restated here independently of the device code.
//   boxes     AABB geometry (gprt/gprt_host.h:614, :1282; the reference hands the boxes to the
//             Vulkan implementation's builder, gprt/gprt.cpp:3961-4003): the LBVH stages above with a box's corners as its leaf box
//             and its centre (lo + hi) * 0.5 as its key point; the closest hit is the box surface
//             (entry distance, or the exit distance from inside), standing in for a user
//             intersection program.
//
// Pinning (see tests/test_oracle.py): Morton known-answer values, closed forms of the t03
// patterns, brute-force sort/scan on small inputs, an independent top-down radix-tree builder
// against the Karras restatement, brute-force per-node boxes against the refit, and an FP64
// ray/plane evaluation plus textbook hits against the triangle test.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

// Total order on float bit patterns (-0 < +0): min/max results independent of fold order.
inline uint32_t key_of(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (u >> 31) ? ~u : (u | 0x80000000u);
}
inline float omin(float a, float b) { return key_of(b) < key_of(a) ? b : a; }
inline float omax(float a, float b) { return key_of(b) > key_of(a) ? b : a; }

inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

struct V3 {
  float x, y, z;
};

inline V3 vert(const float *verts, uint32_t i) { return {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]}; }

// std::stable_sort of (key, payload) pairs by key.
template <typename K, typename V>
static void stable_kv(K *keys, V *vals, size_t n) {
  std::vector<size_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return keys[a] < keys[b]; });
  std::vector<K> k2(n);
  for (size_t i = 0; i < n; ++i) k2[i] = keys[idx[i]];
  if (vals) {
    std::vector<V> v2(n);
    for (size_t i = 0; i < n; ++i) v2[i] = vals[idx[i]];
    std::memcpy(vals, v2.data(), n * sizeof(V));
  }
  std::memcpy(keys, k2.data(), n * sizeof(K));
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------------------------------------
// Sort
// ------------------------------------------------------------------------------------------------

// keyBits 32|64, valBits 0|32|64; in place.
void orc_stable_sort(void *keys, void *vals, size_t n, int keyBits, int valBits) {
  if (keyBits == 32) {
    if (valBits == 64)
      stable_kv((uint32_t *)keys, (uint64_t *)vals, n);
    else
      stable_kv((uint32_t *)keys, (uint32_t *)(valBits ? vals : nullptr), n);
  } else {
    if (valBits == 32)
      stable_kv((uint64_t *)keys, (uint32_t *)vals, n);
    else
      stable_kv((uint64_t *)keys, (uint64_t *)(valBits ? vals : nullptr), n);
  }
}

// ------------------------------------------------------------------------------------------------
// Scan family
// ------------------------------------------------------------------------------------------------
uint32_t orc_exclusive_sum(const uint32_t *in, uint32_t *out, size_t n) {
  uint32_t total = 0;
  for (size_t i = 0; i < n; ++i) total += in[i];
  std::exclusive_scan(in, in + n, out, 0u);
  return total;
}

static inline bool selected(int32_t v, int selectPositive) { return selectPositive ? v >= 0 : v < 0; }

uint32_t orc_partition(const int32_t *in, int32_t *out, size_t n, int selectPositive) {
  uint32_t rank = 0;
  for (size_t i = 0; i < n; ++i) {
    if (selected(in[i], selectPositive)) {
      out[rank] = in[i];
      ++rank;
    } else {
      out[(n - 1) - (i - rank)] = in[i];
    }
  }
  return rank;
}

uint32_t orc_select(const int32_t *in, int32_t *out, size_t n, int selectPositive) {
  uint32_t rank = 0;
  for (size_t i = 0; i < n; ++i)
    if (selected(in[i], selectPositive)) out[rank++] = in[i];
  return rank;
}

// ------------------------------------------------------------------------------------------------
// Morton keys
// ------------------------------------------------------------------------------------------------
uint32_t orc_separate_bits(uint32_t n) {
  n &= 0x000003FF;
  n = (n ^ (n << 16)) & 0xFF0000FF;
  n = (n ^ (n << 8)) & 0x0300F00F;
  n = (n ^ (n << 4)) & 0x030C30C3;
  n = (n ^ (n << 2)) & 0x09249249;
  return n;
}

uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) {
  return orc_separate_bits(x) | (orc_separate_bits(y) << 1) | (orc_separate_bits(z) << 2);
}

// verts: float3[nv]; idx: uint3[ntri]; out: {lo.xyz, hi.xyz}
void orc_scene_bounds(const float *verts, const uint32_t *idx, size_t ntri, float *out) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t t = 0; t < ntri; ++t) {
    for (int c = 0; c < 3; ++c) {
      const V3 p = vert(verts, idx[3 * t + c]);
      lo[0] = omin(lo[0], p.x);
      lo[1] = omin(lo[1], p.y);
      lo[2] = omin(lo[2], p.z);
      hi[0] = omax(hi[0], p.x);
      hi[1] = omax(hi[1], p.y);
      hi[2] = omax(hi[2], p.z);
    }
  }
  for (int a = 0; a < 3; ++a) {
    out[a] = lo[a];
    out[3 + a] = hi[a];
  }
}

// keys[i] = morton30(centroid_i) << 32 | i (unsorted, in primitive order).
void orc_morton_keys(const float *verts, const uint32_t *idx, size_t ntri, const float *bounds, uint64_t *keys) {
  for (size_t i = 0; i < ntri; ++i) {
    const V3 p1 = vert(verts, idx[3 * i]), p2 = vert(verts, idx[3 * i + 1]), p3 = vert(verts, idx[3 * i + 2]);
    float c[3] = {(p1.x + p2.x + p3.x) / 3.f, (p1.y + p2.y + p3.y) / 3.f, (p1.z + p2.z + p3.z) / 3.f};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      float pt = (c[a] - bounds[a]) / (bounds[3 + a] - bounds[a]);
      pt = std::fmin(std::fmax(pt * 1024.f, 0.f), 1023.f);
      q[a] = (uint32_t)pt;
    }
    keys[i] = ((uint64_t)orc_morton3(q[0], q[1], q[2]) << 32) | (uint64_t)i;
  }
}

// ------------------------------------------------------------------------------------------------
// Karras radix tree over sorted, distinct 64-bit keys. nodes: int4[2n-1] {left, right, parent,
// leaf}; inner nodes [0, n-2], leaf i at n-1+i.
// ------------------------------------------------------------------------------------------------
static int delta(const uint64_t *k, long n, long i, long j) {
  if (j < 0 || j >= n) return -1;
  return clz64(k[i] ^ k[j]);
}

void orc_karras(const uint64_t *keys, size_t n_, int32_t *nodes) {
  const long n = (long)n_;
  for (long x = 0; x < 2 * n - 1; ++x) {
    nodes[4 * x + 0] = nodes[4 * x + 1] = nodes[4 * x + 2] = nodes[4 * x + 3] = -1;
  }
  for (long i = 0; i < n; ++i) nodes[4 * (n - 1 + i) + 3] = (int32_t)(uint32_t)(keys[i] & 0xFFFFFFFFu);
  for (long i = 0; i < n - 1; ++i) {
    const int d = delta(keys, n, i, i + 1) >= delta(keys, n, i, i - 1) ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    long lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    long l = 0;
    for (long t = lmax >> 1; t >= 1; t >>= 1)
      if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const long j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    // Split search as written in gprt_lbvh.hlsl:84-95, including its float divisor loop.
    long s = 0;
    float divf = 2.f;
    long t = (long)std::ceil(l / divf);
    for (; t >= 1; divf *= 2.f, t = (long)std::ceil(l / divf))
      if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    const long split = i + s * d + std::min(d, 0);
    const long first = std::min(i, j), last = std::max(i, j);
    const long left = split, right = split + 1;
    if (left == first) {
      nodes[4 * i + 0] = (int32_t)(n - 1 + left);
      nodes[4 * (n - 1 + left) + 2] = (int32_t)i;
    } else {
      nodes[4 * i + 0] = (int32_t)left;
      nodes[4 * left + 2] = (int32_t)i;
    }
    if (right == last) {
      nodes[4 * i + 1] = (int32_t)(n - 1 + right);
      nodes[4 * (n - 1 + right) + 2] = (int32_t)i;
    } else {
      nodes[4 * i + 1] = (int32_t)right;
      nodes[4 * right + 2] = (int32_t)i;
    }
  }
}

// Independent builder: recursive split of [first, last] at the first key that differs from
// keys[first] in the highest bit where keys[first] and keys[last] differ. A left child covering
// [first, s] is numbered s, a right child covering [s+1, last] is numbered s+1; leaves n-1+i.
void orc_topdown(const uint64_t *keys, size_t n_, int32_t *nodes) {
  const long n = (long)n_;
  for (long x = 0; x < 2 * n - 1; ++x) {
    nodes[4 * x + 0] = nodes[4 * x + 1] = nodes[4 * x + 2] = nodes[4 * x + 3] = -1;
  }
  for (long i = 0; i < n; ++i) nodes[4 * (n - 1 + i) + 3] = (int32_t)(uint32_t)(keys[i] & 0xFFFFFFFFu);
  if (n < 2) return;
  struct Item {
    long node, first, last;
  };
  std::vector<Item> stack{{0, 0, n - 1}};
  while (!stack.empty()) {
    const Item it = stack.back();
    stack.pop_back();
    const int bit = 63 - clz64(keys[it.first] ^ keys[it.last]);  // highest differing bit
    // keys in [first, last] are sorted: those with `bit` clear come first.
    long lo = it.first, hi = it.last;  // find last index with bit clear
    while (lo < hi) {
      const long mid = (lo + hi + 1) / 2;
      if ((keys[mid] >> bit) & 1)
        hi = mid - 1;
      else
        lo = mid;
    }
    const long s = lo;
    const long lnode = (s == it.first) ? n - 1 + s : s;
    const long rnode = (s + 1 == it.last) ? n - 1 + s + 1 : s + 1;
    nodes[4 * it.node + 0] = (int32_t)lnode;
    nodes[4 * it.node + 1] = (int32_t)rnode;
    nodes[4 * lnode + 2] = (int32_t)it.node;
    nodes[4 * rnode + 2] = (int32_t)it.node;
    if (s > it.first) stack.push_back({s, it.first, s});
    if (s + 1 < it.last) stack.push_back({s + 1, s + 1, it.last});
  }
}

static void tri_box(const float *verts, const uint32_t *idx, uint32_t prim, float *b) {
  const V3 p1 = vert(verts, idx[3 * prim]), p2 = vert(verts, idx[3 * prim + 1]), p3 = vert(verts, idx[3 * prim + 2]);
  b[0] = omin(p1.x, omin(p2.x, p3.x));
  b[1] = omin(p1.y, omin(p2.y, p3.y));
  b[2] = omin(p1.z, omin(p2.z, p3.z));
  b[3] = omax(p1.x, omax(p2.x, p3.x));
  b[4] = omax(p1.y, omax(p2.y, p3.y));
  b[5] = omax(p1.z, omax(p2.z, p3.z));
}

// Refit as gprt_lbvh.hlsl:400-428: every leaf writes its box and folds it into all ancestors.
// aabbs: float[6 * (2n-1)]; leafBox(prim, box) gives a primitive's box.
extern "C++" {
template <class LeafBox>
static void refit_with(const int32_t *nodes, size_t n_, float *aabbs, LeafBox leafBox) {
  const long n = (long)n_;
  for (long x = 0; x < 2 * n - 1; ++x) {
    for (int a = 0; a < 3; ++a) {
      aabbs[6 * x + a] = 1e38f;
      aabbs[6 * x + 3 + a] = -1e38f;
    }
  }
  for (long i = 0; i < n; ++i) {
    const long leaf = n - 1 + i;
    float b[6];
    leafBox((uint32_t)nodes[4 * leaf + 3], b);
    std::memcpy(&aabbs[6 * leaf], b, sizeof(b));
    for (long p = nodes[4 * leaf + 2]; p >= 0; p = nodes[4 * p + 2]) {
      for (int a = 0; a < 3; ++a) {
        aabbs[6 * p + a] = omin(aabbs[6 * p + a], b[a]);
        aabbs[6 * p + 3 + a] = omax(aabbs[6 * p + 3 + a], b[3 + a]);
      }
    }
  }
}

}  // extern "C++"

void orc_refit(const float *verts, const uint32_t *idx, const int32_t *nodes, size_t n_, float *aabbs) {
  refit_with(nodes, n_, aabbs, [&](uint32_t prim, float *b) { tri_box(verts, idx, prim, b); });
}

// Independent box check: each node's box = min/max over the triangles of its leaf range, with
// the range found by walking to the leftmost and rightmost leaf.
void orc_refit_bruteforce(const float *verts, const uint32_t *idx, const int32_t *nodes, size_t n_, float *aabbs) {
  const long n = (long)n_;
  for (long x = 0; x < 2 * n - 1; ++x) {
    long a = x, b = x;
    while (a < n - 1) a = nodes[4 * a + 0];
    while (b < n - 1) b = nodes[4 * b + 1];
    float box[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (long leaf = a; leaf <= b; ++leaf) {
      float t[6];
      tri_box(verts, idx, (uint32_t)nodes[4 * leaf + 3], t);
      for (int c = 0; c < 3; ++c) {
        box[c] = omin(box[c], t[c]);
        box[3 + c] = omax(box[3 + c], t[3 + c]);
      }
    }
    std::memcpy(&aabbs[6 * x], box, sizeof(box));
  }
}

// Whole pipeline in the reference's stage order (gprt/development/gprt_lbvh.h:170-195).
// keysOut: sorted uint64[n]; nodes int4[2n-1]; aabbs float[6(2n-1)].
void orc_lbvh_build(const float *verts, const uint32_t *idx, size_t n, uint64_t *keysOut, int32_t *nodes,
                    float *aabbs) {
  float bounds[6];
  orc_scene_bounds(verts, idx, n, bounds);
  orc_morton_keys(verts, idx, n, bounds, keysOut);
  orc_stable_sort(keysOut, nullptr, n, 64, 0);
  orc_karras(keysOut, n, nodes);
  orc_refit(verts, idx, nodes, n, aabbs);
}

// ------------------------------------------------------------------------------------------------
// Point and edge LBVHs (gprt/development/gprt_lbvh.h:58-62, :163-195 with GPRT_LBVH_POINTS /
// GPRT_LBVH_EDGES). A point's box is the point and its key point the point itself
// (gprt_lbvh.hlsl:105-110, :228-243); an edge (uint2 into the vertices) has the box of its two end
// points and the key point (p1 + p2) / 2 (:112-119, :131-137, :245-260). Scene box over every
// primitive box (:188-213), quantisation as orc_morton_keys, then the stages of orc_lbvh_build.
// ------------------------------------------------------------------------------------------------
static uint64_t key_of_point(const float c[3], const float *bounds, uint32_t prim) {
  uint32_t q[3];
  for (int a = 0; a < 3; ++a) {
    float pt = (c[a] - bounds[a]) / (bounds[3 + a] - bounds[a]);
    pt = std::fmin(std::fmax(pt * 1024.f, 0.f), 1023.f);
    q[a] = (uint32_t)pt;
  }
  return ((uint64_t)orc_morton3(q[0], q[1], q[2]) << 32) | (uint64_t)prim;
}

static void fold_point(float *b, const float *p) {
  for (int a = 0; a < 3; ++a) {
    b[a] = omin(b[a], p[a]);
    b[3 + a] = omax(b[3 + a], p[a]);
  }
}

// edges == nullptr: points (primitive i is vertex i).
void orc_lbvh_build_pe(const float *verts, const uint32_t *edges, size_t n, uint64_t *keys, int32_t *nodes,
                       float *aabbs) {
  auto primBox = [&](uint32_t prim, float *b) {
    for (int a = 0; a < 3; ++a) {
      b[a] = INFINITY;
      b[3 + a] = -INFINITY;
    }
    if (edges) {
      fold_point(b, verts + 3 * (size_t)edges[2 * prim]);
      fold_point(b, verts + 3 * (size_t)edges[2 * prim + 1]);
    } else {
      fold_point(b, verts + 3 * (size_t)prim);
    }
  };
  float bounds[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; ++i) {
    float b[6];
    primBox((uint32_t)i, b);
    for (int a = 0; a < 3; ++a) {
      bounds[a] = omin(bounds[a], b[a]);
      bounds[3 + a] = omax(bounds[3 + a], b[3 + a]);
    }
  }
  for (size_t i = 0; i < n; ++i) {
    float c[3];
    if (edges) {
      const float *p1 = verts + 3 * (size_t)edges[2 * i], *p2 = verts + 3 * (size_t)edges[2 * i + 1];
      for (int a = 0; a < 3; ++a) c[a] = (p1[a] + p2[a]) / 2.f;
    } else {
      for (int a = 0; a < 3; ++a) c[a] = verts[3 * i + a];
    }
    keys[i] = key_of_point(c, bounds, (uint32_t)i);
  }
  orc_stable_sort(keys, nullptr, n, 64, 0);
  orc_karras(keys, n, nodes);
  refit_with(nodes, n, aabbs, primBox);
}

// ------------------------------------------------------------------------------------------------
// Ray / triangle (watertight, FP32) and closest hit
// ------------------------------------------------------------------------------------------------
struct Pre {
  float o[3], d[3], tmin, tmax;
  int kx, ky, kz;
  float sx, sy, sz;
};

static Pre pre_of(const float *ray) {
  Pre p;
  p.o[0] = ray[0];
  p.o[1] = ray[1];
  p.o[2] = ray[2];
  p.tmin = ray[3];
  p.d[0] = ray[4];
  p.d[1] = ray[5];
  p.d[2] = ray[6];
  p.tmax = ray[7];
  const float ax = std::fabs(p.d[0]), ay = std::fabs(p.d[1]), az = std::fabs(p.d[2]);
  p.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  p.kx = (p.kz + 1) % 3;
  p.ky = (p.kx + 1) % 3;
  if (p.d[p.kz] < 0.f) std::swap(p.kx, p.ky);
  p.sx = p.d[p.kx] / p.d[p.kz];
  p.sy = p.d[p.ky] / p.d[p.kz];
  p.sz = 1.0f / p.d[p.kz];
  return p;
}

// Returns 1 on a hit with tmin < t (t, u = weight of v1, v = weight of v2).
static int hit_tri(const Pre &r, const V3 &v0, const V3 &v1, const V3 &v2, float *t, float *u, float *v) {
  const float A[3] = {v0.x - r.o[0], v0.y - r.o[1], v0.z - r.o[2]};
  const float B[3] = {v1.x - r.o[0], v1.y - r.o[1], v1.z - r.o[2]};
  const float C[3] = {v2.x - r.o[0], v2.y - r.o[1], v2.z - r.o[2]};
  const float Ax = A[r.kx] - r.sx * A[r.kz];
  const float Ay = A[r.ky] - r.sy * A[r.kz];
  const float Bx = B[r.kx] - r.sx * B[r.kz];
  const float By = B[r.ky] - r.sy * B[r.kz];
  const float Cx = C[r.kx] - r.sx * C[r.kz];
  const float Cy = C[r.ky] - r.sy * C[r.kz];
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return 0;
  const float det = U + V + W;
  if (det == 0.f) return 0;
  const float Az = r.sz * A[r.kz];
  const float Bz = r.sz * B[r.kz];
  const float Cz = r.sz * C[r.kz];
  const float T = U * Az + V * Bz + W * Cz;
  const float tt = T / det;
  if (!(tt > r.tmin)) return 0;
  *t = tt;
  *u = V / det;
  *v = W / det;
  return 1;
}

int orc_intersect(const float *ray, const float *v0, const float *v1, const float *v2, float *t, float *u, float *v) {
  const Pre r = pre_of(ray);
  return hit_tri(r, {v0[0], v0[1], v0[2]}, {v1[0], v1[1], v1[2]}, {v2[0], v2[1], v2[2]}, t, u, v);
}

// FP64 distance along the ray to the triangle's plane: (v0-o).n / d.n, n = (v1-v0)x(v2-v0).
double orc_plane_t_fp64(const float *ray, const float *v0, const float *v1, const float *v2) {
  const double o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
  const double e1[3] = {(double)v1[0] - v0[0], (double)v1[1] - v0[1], (double)v1[2] - v0[2]};
  const double e2[3] = {(double)v2[0] - v0[0], (double)v2[1] - v0[1], (double)v2[2] - v0[2]};
  const double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const double w[3] = {v0[0] - o[0], v0[1] - o[1], v0[2] - o[2]};
  return (w[0] * nn[0] + w[1] * nn[1] + w[2] * nn[2]) / (d[0] * nn[0] + d[1] * nn[1] + d[2] * nn[2]);
}

static inline void consider(const Pre &r, const float *verts, const uint32_t *idx, uint32_t prim, float &tbest,
                            int32_t &best, float &bu, float &bv) {
  float t, u, v;
  if (hit_tri(r, vert(verts, idx[3 * prim]), vert(verts, idx[3 * prim + 1]), vert(verts, idx[3 * prim + 2]), &t, &u,
              &v)) {
    if (t < tbest || (t == tbest && (uint32_t)prim < (uint32_t)best)) {
      tbest = t;
      best = (int32_t)prim;
      bu = u;
      bv = v;
    }
  }
}

// hits: {float t; int32 primID; float u, v} per ray; t = tmax and primID = -1 on a miss.
void orc_trace_bruteforce(const float *rays, size_t nrays, const float *verts, const uint32_t *idx, size_t ntri,
                          float *hits, int threads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
  for (long i = 0; i < (long)nrays; ++i) {
    const Pre r = pre_of(rays + 8 * i);
    float tbest = r.tmax, bu = 0.f, bv = 0.f;
    int32_t best = -1;
    for (size_t p = 0; p < ntri; ++p) consider(r, verts, idx, (uint32_t)p, tbest, best, bu, bv);
    hits[4 * i + 0] = tbest;
    std::memcpy(&hits[4 * i + 1], &best, 4);
    hits[4 * i + 2] = bu;
    hits[4 * i + 3] = bv;
  }
}

// Slab test with the same conservative widening rule as the device traversal: the far distance
// is scaled by 1 + 2*gamma(3), and both ends get a relative slack of 2^-14 of |tnear| + |tfar|.
static bool slab(const Pre &r, const float *inv, const float *b, float tbest, float *tn_out) {
  float tn = -INFINITY, tf = INFINITY;
  for (int a = 0; a < 3; ++a) {
    const float t0 = (b[a] - r.o[a]) * inv[a], t1 = (b[3 + a] - r.o[a]) * inv[a];
    tn = std::fmax(tn, std::fmin(t0, t1));
    tf = std::fmin(tf, std::fmax(t0, t1));
  }
  const float g3 = (3.0f * 5.9604645e-08f) / (1.0f - 3.0f * 5.9604645e-08f);
  const float slack = (1.0f / 16384.0f) * (std::fabs(tn) + std::fabs(tf));
  *tn_out = tn;
  return std::fmax(tn - slack, r.tmin) <= std::fmin(tf * (1.0f + 2.0f * g3) + slack, tbest);
}

// Closest hit of one ray (8 floats) over the reference-layout tree; adds the inner nodes visited
// and the triangles tested to the counters.
static void trace_ray(const float *ray, const float *verts, const uint32_t *idx, const int32_t *nodes,
                      const float *aabbs, long n, float *hit, uint64_t &visited, uint64_t &tested) {
  const Pre r = pre_of(ray);
  float inv[3];
  for (int a = 0; a < 3; ++a) {
    const float d = std::fabs(r.d[a]) < 1e-30f ? std::copysign(1e-30f, r.d[a]) : r.d[a];
    inv[a] = 1.0f / d;
  }
  float tbest = r.tmax, bu = 0.f, bv = 0.f;
  int32_t best = -1;
  if (n == 1) {
    consider(r, verts, idx, (uint32_t)nodes[3], tbest, best, bu, bv);
    ++tested;
  } else {
    long stack[128];
    int sp = 0;
    long node = 0;
    while (true) {
      ++visited;
      const long c[2] = {nodes[4 * node + 0], nodes[4 * node + 1]};
      float tn[2];
      bool h[2];
      for (int s = 0; s < 2; ++s) h[s] = slab(r, inv, &aabbs[6 * c[s]], tbest, &tn[s]);
      for (int s = 0; s < 2; ++s) {
        if (h[s] && c[s] >= n - 1) {
          consider(r, verts, idx, (uint32_t)nodes[4 * c[s] + 3], tbest, best, bu, bv);
          ++tested;
          h[s] = false;
        }
      }
      if (h[0] && h[1]) {
        const bool nearFirst = tn[0] <= tn[1];
        stack[sp++] = nearFirst ? c[1] : c[0];
        node = nearFirst ? c[0] : c[1];
      } else if (h[0]) {
        node = c[0];
      } else if (h[1]) {
        node = c[1];
      } else {
        if (sp == 0) break;
        node = stack[--sp];
      }
    }
  }
  hit[0] = tbest;
  std::memcpy(&hit[1], &best, 4);
  hit[2] = bu;
  hit[3] = bv;
}

// Serial (threads=1) or per-ray parallel BVH traversal over the reference-layout tree.
// counters (optional): [0] inner nodes visited, [1] triangles tested, summed over rays.
void orc_trace_bvh(const float *rays, size_t nrays, const float *verts, const uint32_t *idx, const int32_t *nodes,
                   const float *aabbs, size_t n_, float *hits, uint64_t *counters, int threads) {
  const long n = (long)n_;
  uint64_t visited = 0, tested = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1) reduction(+ : visited, tested)
  for (long i = 0; i < (long)nrays; ++i) trace_ray(rays + 8 * i, verts, idx, nodes, aabbs, n, hits + 4 * i, visited, tested);
  if (counters) {
    counters[0] = visited;
    counters[1] = tested;
  }
}

// One ray, no threading (for callers that parallelise over rays themselves).
void orc_trace_bvh_ray(const float *ray, const float *verts, const uint32_t *idx, const int32_t *nodes,
                       const float *aabbs, size_t n, float *hit) {
  uint64_t visited = 0, tested = 0;
  trace_ray(ray, verts, idx, nodes, aabbs, (long)n, hit, visited, tested);
}

// ------------------------------------------------------------------------------------------------
// Box primitives. boxes: float[6 * n] {lo.xyz, hi.xyz} as given (not necessarily ordered).
// ------------------------------------------------------------------------------------------------
static void box_of(const float *boxes, uint32_t prim, float *b) {
  const float *p = boxes + 6 * (size_t)prim;
  for (int a = 0; a < 3; ++a) {
    b[a] = omin(p[a], p[3 + a]);
    b[3 + a] = omax(p[a], p[3 + a]);
  }
}

// Scene box over both corners of every box.
void orc_box_bounds(const float *boxes, size_t n, float *out) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; ++i) {
    const float *p = boxes + 6 * i;
    for (int c = 0; c < 2; ++c)
      for (int a = 0; a < 3; ++a) {
        lo[a] = omin(lo[a], p[3 * c + a]);
        hi[a] = omax(hi[a], p[3 * c + a]);
      }
  }
  for (int a = 0; a < 3; ++a) {
    out[a] = lo[a];
    out[3 + a] = hi[a];
  }
}

// keys[i] = morton30(centre_i) << 32 | i, centre = (lo + hi) * 0.5, quantised as orc_morton_keys.
void orc_box_keys(const float *boxes, size_t n, const float *bounds, uint64_t *keys) {
  for (size_t i = 0; i < n; ++i) {
    const float *p = boxes + 6 * i;
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      const float c = (p[a] + p[3 + a]) * 0.5f;
      float pt = (c - bounds[a]) / (bounds[3 + a] - bounds[a]);
      pt = std::fmin(std::fmax(pt * 1024.f, 0.f), 1023.f);
      q[a] = (uint32_t)pt;
    }
    keys[i] = ((uint64_t)orc_morton3(q[0], q[1], q[2]) << 32) | (uint64_t)i;
  }
}

void orc_box_refit(const float *boxes, const int32_t *nodes, size_t n, float *aabbs) {
  refit_with(nodes, n, aabbs, [&](uint32_t prim, float *b) { box_of(boxes, prim, b); });
}

void orc_lbvh_build_boxes(const float *boxes, size_t n, uint64_t *keysOut, int32_t *nodes, float *aabbs) {
  float bounds[6];
  orc_box_bounds(boxes, n, bounds);
  orc_box_keys(boxes, n, bounds, keysOut);
  orc_stable_sort(keysOut, nullptr, n, 64, 0);
  orc_karras(keysOut, n, nodes);
  orc_box_refit(boxes, nodes, n, aabbs);
}

// Box surface hit: slab distances with the total-order min/max; the entry distance when the ray
// enters the box after tmin, else the exit distance when it starts inside. inv = 1 / d with
// |d| clamped to 1e-30 (the traversal's reciprocal).
static int hit_box(const Pre &r, const float *inv, const float *b, float *t) {
  float t0[3], t1[3];
  for (int a = 0; a < 3; ++a) {
    t0[a] = (b[a] - r.o[a]) * inv[a];
    t1[a] = (b[3 + a] - r.o[a]) * inv[a];
  }
  const float tn = omax(omax(omin(t0[0], t1[0]), omin(t0[1], t1[1])), omin(t0[2], t1[2]));
  const float tf = omin(omin(omax(t0[0], t1[0]), omax(t0[1], t1[1])), omax(t0[2], t1[2]));
  if (!(tn <= tf)) return 0;
  if (tn > r.tmin) {
    *t = tn;
    return 1;
  }
  if (tf > r.tmin) {
    *t = tf;
    return 1;
  }
  return 0;
}

int orc_box_hit(const float *ray, const float *box, float *t) {
  const Pre r = pre_of(ray);
  float inv[3], b[6];
  for (int a = 0; a < 3; ++a) {
    const float d = std::fabs(r.d[a]) < 1e-30f ? std::copysign(1e-30f, r.d[a]) : r.d[a];
    inv[a] = 1.0f / d;
    b[a] = omin(box[a], box[3 + a]);
    b[3 + a] = omax(box[a], box[3 + a]);
  }
  return hit_box(r, inv, b, t);
}

// hits as orc_trace_bruteforce (u = v = 0): every box tested, the smallest (t, primID) kept.
void orc_trace_boxes_bruteforce(const float *rays, size_t nrays, const float *boxes, size_t n, float *hits,
                                int threads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
  for (long i = 0; i < (long)nrays; ++i) {
    const Pre r = pre_of(rays + 8 * i);
    float inv[3];
    for (int a = 0; a < 3; ++a) {
      const float d = std::fabs(r.d[a]) < 1e-30f ? std::copysign(1e-30f, r.d[a]) : r.d[a];
      inv[a] = 1.0f / d;
    }
    float tbest = r.tmax;
    int32_t best = -1;
    for (size_t p = 0; p < n; ++p) {
      float b[6], t;
      box_of(boxes, (uint32_t)p, b);
      if (hit_box(r, inv, b, &t) && (t < tbest || (t == tbest && (uint32_t)p < (uint32_t)best))) {
        tbest = t;
        best = (int32_t)p;
      }
    }
    hits[4 * i + 0] = tbest;
    std::memcpy(&hits[4 * i + 1], &best, 4);
    hits[4 * i + 2] = 0.f;
    hits[4 * i + 3] = 0.f;
  }
}

int orc_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
