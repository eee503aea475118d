// ORACLE — TEST INFRASTRUCTURE ONLY: the all-cores CPU baseline.
//
// OpenMP forms of the oracle's functions (oracle.cpp), timed by bench.py beside the GPU path as
// the multi-threaded CPU baseline. Each produces exactly what its serial counterpart in
// oracle.cpp produces (tests/test_oracle_omp.py checks that), so the baseline does the same work
// the GPU does. Nothing in the product (gprt_amd/) links, loads or calls this file.
//
//   orc_par_radix_sort      LSD radix sort, 8-bit digits, per-thread digit histograms over
//                           contiguous blocks, stable scatter; the reference's sort semantics
//                           (stable, payload follows its key; gprt/gprt.cpp:8821-9073).
//   orc_par_exclusive_sum   two-pass blocked scan: block sums, their prefix, block scans
//                           (tests/t03-bufferScans/hostCode.cpp:53-58 semantics, mod 2^32).
//   orc_par_lbvh_build      the reference's stage order (gprt/development/gprt_lbvh.h:170-195):
//                           scene box (per-thread min/max), Morton keys (parallel over
//                           triangles), the radix sort of the 30-bit codes with primID payload
//                           (equal to sorting morton << 32 | primID), Karras per inner node
//                           (gprt_lbvh.hlsl:40-103, :287-338), and a bottom-up refit in which
//                           the second child to arrive at a node (an atomic flag per node) merges
//                           both children's boxes (the order-free result of gprt_lbvh.hlsl:400-428).
//   orc_par_refit           that refit alone over a kept topology (an update).
//   orc_par_lbvh_build_many one serial oracle build per mesh, meshes spread over the threads.
//   orc_trace_bvh_multi     per-ray parallel closest hit over several meshes' BVHs (identity
//                           instances): the minimum of (t, mesh, primID).
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {
void orc_scene_bounds(const float *verts, const uint32_t *idx, size_t ntri, float *out);
void orc_karras(const uint64_t *keys, size_t n_, int32_t *nodes);
void orc_lbvh_build(const float *verts, const uint32_t *idx, size_t n, uint64_t *keysOut, int32_t *nodes,
                    float *aabbs);
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z);
void orc_trace_bvh_ray(const float *ray, const float *verts, const uint32_t *idx, const int32_t *nodes,
                       const float *aabbs, size_t n, float *hit);
}

namespace {

inline int nthreads(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

// Total order on float bit patterns, as oracle.cpp (-0 < +0).
inline uint32_t key_of(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (u >> 31) ? ~u : (u | 0x80000000u);
}
inline float omin(float a, float b) { return key_of(b) < key_of(a) ? b : a; }
inline float omax(float a, float b) { return key_of(b) > key_of(a) ? b : a; }

// Stable LSD radix sort over bits [firstBit, lastBit) of the keys, 8 bits per pass.
template <class K, class V>
void par_radix(K *keys, V *vals, size_t n, int firstBit, int lastBit, int T) {
  if (n < 2) return;
  std::vector<K> kt(n);
  std::vector<V> vt(vals ? n : 0);
  K *ks = keys, *kd = kt.data();
  V *vs = vals, *vd = vals ? vt.data() : nullptr;
  std::vector<size_t> hist((size_t)T * 256);
  for (int shift = firstBit; shift < lastBit; shift += 8) {
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num(), nt = omp_get_num_threads();
      const size_t lo = n * (size_t)t / nt, hi = n * (size_t)(t + 1) / nt;
      size_t *h = &hist[(size_t)t * 256];
      std::fill(h, h + 256, (size_t)0);
      for (size_t i = lo; i < hi; ++i) ++h[(ks[i] >> shift) & 255u];
#pragma omp barrier
#pragma omp single
      {
        // Digit-major, thread-minor offsets: equal digits keep their input order (stable).
        size_t off = 0;
        for (int d = 0; d < 256; ++d)
          for (int u = 0; u < nt; ++u) {
            const size_t c = hist[(size_t)u * 256 + d];
            hist[(size_t)u * 256 + d] = off;
            off += c;
          }
      }
      for (size_t i = lo; i < hi; ++i) {
        const size_t o = h[(ks[i] >> shift) & 255u]++;
        kd[o] = ks[i];
        if (vs) vd[o] = vs[i];
      }
    }
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  if (ks != keys) {
#pragma omp parallel for num_threads(T) schedule(static)
    for (long i = 0; i < (long)n; ++i) {
      keys[i] = ks[i];
      if (vals) vals[i] = vs[i];
    }
  }
}

template <class K>
void radix_kv(void *keys, void *vals, size_t n, int valBits, int T) {
  const int bits = (int)sizeof(K) * 8;
  if (valBits == 32)
    par_radix((K *)keys, (uint32_t *)vals, n, 0, bits, T);
  else if (valBits == 64)
    par_radix((K *)keys, (uint64_t *)vals, n, 0, bits, T);
  else
    par_radix((K *)keys, (uint32_t *)nullptr, n, 0, bits, T);
}

}  // namespace

extern "C" {

int orc_num_procs(void) { return omp_get_num_procs(); }

void orc_par_radix_sort(void *keys, void *vals, size_t n, int keyBits, int valBits, int threads) {
  const int T = nthreads(threads);
  if (keyBits == 32)
    radix_kv<uint32_t>(keys, vals, n, valBits, T);
  else
    radix_kv<uint64_t>(keys, vals, n, valBits, T);
}

uint32_t orc_par_exclusive_sum(const uint32_t *in, uint32_t *out, size_t n, int threads) {
  const int T = nthreads(threads);
  std::vector<uint32_t> part((size_t)T + 1, 0u);
  int used = 1;
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const size_t lo = n * (size_t)t / nt, hi = n * (size_t)(t + 1) / nt;
    uint32_t s = 0;
    for (size_t i = lo; i < hi; ++i) s += in[i];
    part[(size_t)t + 1] = s;
#pragma omp barrier
#pragma omp single
    {
      used = nt;
      for (int u = 0; u < nt; ++u) part[(size_t)u + 1] += part[u];
    }
    uint32_t run = part[t];
    for (size_t i = lo; i < hi; ++i) {
      const uint32_t v = in[i];  // read before the write: in and out may alias
      out[i] = run;
      run += v;
    }
  }
  return part[used];
}

}  // extern "C"

// Bottom-up refit over a reference-layout tree: a leaf writes its primitive's box and climbs; at
// each inner node the first child to arrive stops (an atomic flag per node) and the second merges
// both children's boxes and goes on. The result is order-free (min / max in the total order).
template <class LeafBox>
static void par_refit_with(const int32_t *nodes, size_t n, float *aabbs, int T, LeafBox leafBox) {
  const long N = (long)n;
  std::vector<std::atomic<int>> flag(N > 1 ? (size_t)(N - 1) : 1);
#pragma omp parallel for num_threads(T) schedule(static)
  for (long i = 0; i < N - 1; ++i) flag[i].store(0, std::memory_order_relaxed);
#pragma omp parallel for num_threads(T) schedule(static)
  for (long i = 0; i < N; ++i) {
    const long leaf = N - 1 + i;
    leafBox((uint32_t)nodes[4 * leaf + 3], &aabbs[6 * leaf]);
    for (long p = nodes[4 * leaf + 2]; p >= 0; p = nodes[4 * p + 2]) {
      if (flag[p].fetch_add(1, std::memory_order_acq_rel) == 0) break;
      const float *c0 = &aabbs[6 * (long)nodes[4 * p + 0]], *c1 = &aabbs[6 * (long)nodes[4 * p + 1]];
      float *o = &aabbs[6 * p];
      for (int a = 0; a < 3; ++a) {
        o[a] = omin(c0[a], c1[a]);
        o[3 + a] = omax(c0[3 + a], c1[3 + a]);
      }
    }
  }
}

extern "C" {

void orc_par_refit(const float *verts, const uint32_t *idx, const int32_t *nodes, size_t n, float *aabbs,
                   int threads) {
  par_refit_with(nodes, n, aabbs, nthreads(threads), [&](uint32_t prim, float *b) {
    const float *p1 = verts + 3 * (size_t)idx[3 * prim], *p2 = verts + 3 * (size_t)idx[3 * prim + 1],
                *p3 = verts + 3 * (size_t)idx[3 * prim + 2];
    for (int a = 0; a < 3; ++a) {
      b[a] = omin(p1[a], omin(p2[a], p3[a]));
      b[3 + a] = omax(p1[a], omax(p2[a], p3[a]));
    }
  });
}

}  // extern "C"

// Scene box over `count` points (point(i, p) gives point i): per-thread min / max in the total
// order (exact, order-free), then combined.
template <class Point>
static void par_bounds(long count, int T, float bounds[6], Point point) {
  std::vector<float> part((size_t)T * 6);
  int used = 1;
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
#pragma omp for schedule(static)
    for (long i = 0; i < count; ++i) {
      const float *p = point(i);
      for (int a = 0; a < 3; ++a) {
        b[a] = omin(b[a], p[a]);
        b[3 + a] = omax(b[3 + a], p[a]);
      }
    }
    std::memcpy(&part[(size_t)t * 6], b, sizeof(b));
#pragma omp single
    used = nt;
  }
  for (int a = 0; a < 3; ++a) {
    bounds[a] = INFINITY;
    bounds[3 + a] = -INFINITY;
  }
  for (int t = 0; t < used; ++t)
    for (int a = 0; a < 3; ++a) {
      bounds[a] = omin(bounds[a], part[(size_t)t * 6 + a]);
      bounds[3 + a] = omax(bounds[3 + a], part[(size_t)t * 6 + 3 + a]);
    }
}

// Morton codes of key points (keyPoint(i, c) writes point i's key point, the arithmetic of
// orc_morton_keys), sorted with primID payloads into keysOut (morton << 32 | primID), then the
// Karras tree: every inner node independently (each node is the only writer of its own fields
// and of its children's parent fields).
template <class KeyPoint>
static void par_keys_karras(size_t n, const float bounds[6], int T, uint64_t *keysOut, int32_t *nodes,
                            KeyPoint keyPoint) {
  const long N = (long)n;
  std::vector<uint32_t> codes(n), ids(n);
#pragma omp parallel for num_threads(T) schedule(static)
  for (long i = 0; i < N; ++i) {
    float c[3];
    keyPoint(i, c);
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      float pt = (c[a] - bounds[a]) / (bounds[3 + a] - bounds[a]);
      pt = std::fmin(std::fmax(pt * 1024.f, 0.f), 1023.f);
      q[a] = (uint32_t)pt;
    }
    codes[i] = orc_morton3(q[0], q[1], q[2]);
    ids[i] = (uint32_t)i;
  }
  // Codes are 30 bits: four 8-bit passes; primIDs enter in increasing order, so the stable sort
  // yields the order of morton << 32 | primID.
  par_radix(codes.data(), ids.data(), n, 0, 32, T);
#pragma omp parallel for num_threads(T) schedule(static)
  for (long i = 0; i < N; ++i) keysOut[i] = ((uint64_t)codes[i] << 32) | ids[i];
  auto delta = [&](long i, long j) -> int {
    if (j < 0 || j >= N) return -1;
    const uint64_t x = keysOut[i] ^ keysOut[j];
    return x ? __builtin_clzll(x) : 64;
  };
#pragma omp parallel for num_threads(T) schedule(static)
  for (long x = 0; x < 2 * N - 1; ++x) {
    nodes[4 * x + 0] = nodes[4 * x + 1] = nodes[4 * x + 2] = -1;
    nodes[4 * x + 3] = x >= N - 1 ? (int32_t)ids[x - (N - 1)] : -1;
  }
#pragma omp parallel for num_threads(T) schedule(dynamic, 4096)
  for (long i = 0; i < N - 1; ++i) {
    const int d = delta(i, i + 1) >= delta(i, i - 1) ? 1 : -1;
    const int dmin = delta(i, i - d);
    long lmax = 2;
    while (delta(i, i + lmax * d) > dmin) lmax *= 2;
    long l = 0;
    for (long t = lmax >> 1; t >= 1; t >>= 1)
      if (delta(i, i + (l + t) * d) > dmin) l += t;
    const long j = i + l * d;
    const int dnode = delta(i, j);
    long s = 0;
    float divf = 2.f;
    for (long t = (long)std::ceil(l / divf); t >= 1; divf *= 2.f, t = (long)std::ceil(l / divf))
      if (delta(i, i + (s + t) * d) > dnode) s += t;
    const long split = i + s * d + std::min(d, 0);
    const long first = std::min(i, j), last = std::max(i, j);
    const long left = split == first ? N - 1 + split : split;
    const long right = split + 1 == last ? N - 1 + split + 1 : split + 1;
    nodes[4 * i + 0] = (int32_t)left;
    nodes[4 * i + 1] = (int32_t)right;
    nodes[4 * left + 2] = (int32_t)i;
    nodes[4 * right + 2] = (int32_t)i;
  }
}

extern "C" {

void orc_par_lbvh_build(const float *verts, const uint32_t *idx, size_t n, uint64_t *keysOut, int32_t *nodes,
                        float *aabbs, int threads) {
  const int T = nthreads(threads);
  float bounds[6];
  par_bounds(3 * (long)n, T, bounds, [&](long i) { return verts + 3 * (size_t)idx[i]; });
  par_keys_karras(n, bounds, T, keysOut, nodes, [&](long i, float *c) {
    const float *p1 = verts + 3 * (size_t)idx[3 * i], *p2 = verts + 3 * (size_t)idx[3 * i + 1],
                *p3 = verts + 3 * (size_t)idx[3 * i + 2];
    for (int a = 0; a < 3; ++a) c[a] = (p1[a] + p2[a] + p3[a]) / 3.f;
  });
  orc_par_refit(verts, idx, nodes, n, aabbs, threads);
}

// The LBVH over box primitives {lo.xyz, hi.xyz} (orc_lbvh_build_boxes): scene box over both
// corners, key point the centre (lo + hi) * 0.5, leaf box the ordered corners.
void orc_par_lbvh_build_boxes(const float *boxes, size_t n, uint64_t *keysOut, int32_t *nodes, float *aabbs,
                              int threads) {
  const int T = nthreads(threads);
  float bounds[6];
  par_bounds(2 * (long)n, T, bounds, [&](long i) { return boxes + 3 * (size_t)i; });
  par_keys_karras(n, bounds, T, keysOut, nodes, [&](long i, float *c) {
    const float *p = boxes + 6 * (size_t)i;
    for (int a = 0; a < 3; ++a) c[a] = (p[a] + p[3 + a]) * 0.5f;
  });
  par_refit_with(nodes, n, aabbs, T, [&](uint32_t prim, float *b) {
    const float *p = boxes + 6 * (size_t)prim;
    for (int a = 0; a < 3; ++a) {
      b[a] = omin(p[a], p[3 + a]);
      b[3 + a] = omax(p[a], p[3 + a]);
    }
  });
}

// Sphere bounds (the compute sample's ComputeBounds): box i = {centre - r, centre + r}.
void orc_par_sphere_bounds(const float *centres, const float *radii, size_t n, float *boxes, int threads) {
#pragma omp parallel for num_threads(nthreads(threads)) schedule(static)
  for (long i = 0; i < (long)n; ++i) {
    const float *c = centres + 3 * (size_t)i;
    float *b = boxes + 6 * (size_t)i;
    for (int a = 0; a < 3; ++a) {
      b[a] = c[a] - radii[i];
      b[3 + a] = c[a] + radii[i];
    }
  }
}

// sizes[m] triangles per mesh; mesh m's vertices start at vertex 3 * (its first triangle) and
// its indices are local to it (the blob-mesh layout). Outputs are concatenated per mesh: keys at
// the mesh's first triangle, nodes and boxes at 2 * first - m.
void orc_par_lbvh_build_many(const float *verts, const uint32_t *idx, const int64_t *sizes, size_t meshes,
                             uint64_t *keys, int32_t *nodes, float *aabbs, int threads) {
  std::vector<int64_t> first(meshes + 1, 0);
  for (size_t m = 0; m < meshes; ++m) first[m + 1] = first[m] + sizes[m];
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 1)
  for (long m = 0; m < (long)meshes; ++m) {
    const int64_t f = first[m], nodeOff = 2 * f - m;
    orc_lbvh_build(verts + 9 * (size_t)f, idx + 3 * (size_t)f, (size_t)sizes[m], keys + f, nodes + 4 * nodeOff,
                   aabbs + 6 * nodeOff);
  }
}

// Closest hit over several meshes' BVHs (identity instances, layout of orc_par_lbvh_build_many):
// per ray, the minimum of (t, mesh, primID); inst[i] = the mesh (-1 on a miss).
void orc_trace_bvh_multi(const float *rays, size_t nrays, const float *verts, const uint32_t *idx,
                         const int64_t *sizes, size_t meshes, const int32_t *nodes, const float *aabbs, float *hits,
                         int32_t *inst, int threads) {
  std::vector<int64_t> first(meshes + 1, 0);
  for (size_t m = 0; m < meshes; ++m) first[m + 1] = first[m] + sizes[m];
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 16)
  for (long i = 0; i < (long)nrays; ++i) {
    float r[8], best[4] = {0.f, 0.f, 0.f, 0.f};
    std::memcpy(r, rays + 8 * i, sizeof(r));
    best[0] = r[7];
    int32_t bp = -1, bm = -1;
    for (long m = 0; m < (long)meshes; ++m) {
      const int64_t f = first[m], nodeOff = 2 * f - m;
      float h[4];
      orc_trace_bvh_ray(r, verts + 9 * (size_t)f, idx + 3 * (size_t)f, nodes + 4 * nodeOff, aabbs + 6 * nodeOff,
                        (size_t)sizes[m], h);
      int32_t p;
      std::memcpy(&p, &h[1], 4);
      // Meshes come in increasing order, so an equal t from a later mesh never wins; a hit
      // tightens the interval for the meshes after it.
      if (p >= 0 && (bp < 0 || h[0] < best[0])) {
        std::memcpy(best, h, sizeof(h));
        bp = p;
        bm = (int32_t)m;
        r[7] = h[0];
      }
    }
    std::memcpy(hits + 4 * i, best, sizeof(best));
    std::memcpy(hits + 4 * i + 1, &bp, 4);
    inst[i] = bm;
  }
}

}  // extern "C"
