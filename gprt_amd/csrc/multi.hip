// Multi-instance closest hit over several triangle accels (each with its own LBVH, identity
// transforms) and device-side bounce-ray generation, for scenes made of many independently
// rebuilt meshes. The reference traces instances through a hardware TLAS
// (gprtInstanceAccelCreate, gprt/gprt_host.h:1418); here a ray tests every instance's root box,
// visits the hit instances nearest-first and keeps the lexicographic minimum of
// (t, instance, primID), so the result does not depend on the visiting order.
#include <cstring>
#include <vector>

#include "bvh.hpp"
#include "device_common.hpp"
#include "internal.hpp"
#include "traverse.hpp"

namespace gprt {
namespace {

struct InstanceDesc {
  const BVHNode *nodes;
  const TriRecord *tris;
  const float *header;  // root box at [8..13]
  const char *verts;
  const char *idx;
  uint32_t n, vstride, istride, nverts;
};

__device__ __forceinline__ bool root_hit(const RayPre &r, const float *h, float tbest, float &tn) {
  return box_hit(r, h[8], h[9], h[10], h[11], h[12], h[13], tbest, tn);
}

__global__ void __launch_bounds__(kTraceBlock)
    k_trace_multi(const InstanceDesc *__restrict__ inst, uint32_t ninst, const GPRTRay *__restrict__ rays,
                  GPRTHit *__restrict__ hits, int32_t *__restrict__ instOut, uint32_t count,
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
      HitState best;
      best.t = ray.tmax;
      best.u = best.v = 0.f;
      best.prim = -1;
      int bestInst = -1;
      // Instances whose root box the ray meets, visited in order of entry distance (at most 64
      // tracked in a mask; beyond that, index order).
      uint64_t pending = 0;
      for (uint32_t j = 0; j < ninst && j < 64; ++j) {
        float tn;
        if (root_hit(r, inst[j].header, best.t, tn)) pending |= 1ull << j;
      }
      for (uint32_t it = 0; it < ninst; ++it) {
        int j = -1;
        if (it < 64 && pending) {
          float tmin = INFINITY;
          for (uint64_t m = pending; m; m &= m - 1) {
            const int q = __builtin_ctzll(m);
            float tn;
            if (!root_hit(r, inst[q].header, best.t, tn)) {
              pending &= ~(1ull << q);
              continue;
            }
            if (tn < tmin) {
              tmin = tn;
              j = q;
            }
          }
          if (j < 0) {
            if (ninst <= 64) break;
            continue;
          }
          pending &= ~(1ull << j);
        } else if (it >= 64) {
          j = (int)it;
          float tn;
          if (!root_hit(r, inst[j].header, best.t, tn)) continue;
        } else {
          if (ninst <= 64) break;
          continue;
        }
        const InstanceDesc d = inst[j];
        HitState h;
        h.t = best.t;
        h.u = h.v = 0.f;
        h.prim = -1;
        traverse(r, d.nodes, d.tris, d.n, d.n == 1 ? ~0 : 0, h, s_stack, tid, mySpill);
        if (h.prim >= 0 && (h.t < best.t || (h.t == best.t && (bestInst < 0 || j < bestInst)))) {
          best = h;
          bestInst = j;
        }
      }
      GPRTHit o;
      o.t = best.prim >= 0 ? best.t : ray.tmax;
      o.primID = best.prim;
      o.u = best.u;
      o.v = best.v;
      hits[ri] = o;
      if (instOut) instOut[ri] = bestInst;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Bounce rays: compact the hit rays (flags -> exclusive scan -> scatter), then ray j starts at the
// hit point of hit (j mod hits), cosine-weighted about the geometric normal facing the incoming
// ray (Malley's method, basis of Duff et al. 2017), FP64 then one rounding, as scenes.bounce_rays.
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

__global__ void __launch_bounds__(256) k_bounce(const InstanceDesc *__restrict__ inst, uint32_t ninst,
                                                const GPRTRay *__restrict__ prim, const GPRTHit *__restrict__ hits,
                                                const int32_t *__restrict__ instIds, const uint32_t *__restrict__ list,
                                                const uint32_t *__restrict__ numHits, uint32_t count, uint64_t seed,
                                                float tmin, GPRTRay *__restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const uint32_t nh = *numHits;
  GPRTRay o;
  if (nh == 0) {
    o = GPRTRay{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 1.f, -1.f};  // empty interval: traces as a miss
    out[j] = o;
    return;
  }
  const uint32_t src = list[j % nh];
  const GPRTRay r = prim[src];
  const GPRTHit h = hits[src];
  const int32_t ii = instIds ? instIds[src] : 0;
  const InstanceDesc d = inst[(uint32_t)ii < ninst ? ii : 0];
  const uint32_t p = (uint32_t)h.primID < d.n ? (uint32_t)h.primID : d.n - 1;
  const uint32_t *tri = (const uint32_t *)(d.idx + (size_t)p * d.istride);
  double v[3][3];
  for (int k = 0; k < 3; ++k) {
    const uint32_t vi = tri[k] < d.nverts ? tri[k] : d.nverts - 1;
    const float *pv = (const float *)(d.verts + (size_t)vi * d.vstride);
    v[k][0] = pv[0];
    v[k][1] = pv[1];
    v[k][2] = pv[2];
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
  const double b = n[0] * n[1] * a;
  const double t1[3] = {1.0 + sgn * n[0] * n[0] * a, sgn * b, -sgn * n[0]};
  const double t2[3] = {b, sgn + n[1] * n[1] * a, -n[1]};
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

struct InstanceCache {
  std::vector<InstanceDesc> host;
  InstanceDesc *dev = nullptr;
  size_t cap = 0;
};

InstanceCache &instance_cache(Context *ctx) {
  static thread_local std::vector<std::pair<Context *, InstanceCache>> caches;
  for (auto &c : caches)
    if (c.first == ctx) return c.second;
  caches.emplace_back(ctx, InstanceCache());
  return caches.back().second;
}

// Uploads (when changed) the instance table of `accels` and returns its device address.
const InstanceDesc *instance_table(Context *ctx, hipStream_t s, Accel **accels, uint32_t count) {
  std::vector<InstanceDesc> t(count);
  for (uint32_t i = 0; i < count; ++i) {
    Accel *a = accels[i];
    Geom *g = a->geom;
    std::memset(&t[i], 0, sizeof(InstanceDesc));
    t[i].nodes = (const BVHNode *)a->nodes;
    t[i].tris = (const TriRecord *)a->tris;
    t[i].header = a->header;
    t[i].verts = (const char *)g->vbuf->d + g->voffset;
    t[i].idx = (const char *)g->ibuf->d + g->ioffset;
    t[i].n = a->n;
    t[i].vstride = g->vstride;
    t[i].istride = g->istride;
    t[i].nverts = g->vcount;
  }
  InstanceCache &c = instance_cache(ctx);
  if (c.host.size() == t.size() && std::memcmp(c.host.data(), t.data(), t.size() * sizeof(InstanceDesc)) == 0)
    return c.dev;
  if (count > c.cap) {
    if (c.dev) {
      hipStreamSynchronize(s);
      hipFree(c.dev);
    }
    c.dev = nullptr;
    c.cap = 0;
    if (!hip_ok(hipMalloc((void **)&c.dev, count * sizeof(InstanceDesc)), "hipMalloc(instances)")) return nullptr;
    c.cap = count;
  }
  c.host = t;
  if (!hip_ok(hipMemcpyAsync(c.dev, c.host.data(), count * sizeof(InstanceDesc), hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(instances)"))
    return nullptr;
  return c.dev;
}

bool collect(GPRTAccel *_accels, uint32_t count, std::vector<Accel *> &out, const char *fn) {
  out.resize(count);
  for (uint32_t i = 0; i < count; ++i) {
    out[i] = (Accel *)_accels[i];
    if (!out[i] || !out[i]->built) {
      fail("%s: accel %u is null or not built", fn, i);
      return false;
    }
  }
  return true;
}

}  // namespace
}  // namespace gprt

using namespace gprt;

GPRT_API void gprtAccelTraceClosestHitMulti(GPRTContext _ctx, GPRTAccel *_accels, uint32_t count, GPRTBuffer _rays,
                                            GPRTBuffer _hits, uint32_t n, GPRTBuffer _instanceIDs) {
  Context *ctx = (Context *)_ctx;
  Buffer *rays = (Buffer *)_rays, *hits = (Buffer *)_hits, *ids = (Buffer *)_instanceIDs;
  if (!ctx || !_accels || !rays || !hits || count == 0) {
    fail("gprtAccelTraceClosestHitMulti: null argument or no accels");
    return;
  }
  std::vector<Accel *> accels;
  if (!collect(_accels, count, accels, "gprtAccelTraceClosestHitMulti")) return;
  if ((size_t)n * sizeof(GPRTRay) > rays->size || (size_t)n * sizeof(GPRTHit) > hits->size ||
      (ids && (size_t)n * 4 > ids->size)) {
    fail("gprtAccelTraceClosestHitMulti: %u rays exceed a buffer", n);
    return;
  }
  if (n == 0) return;
  hipStream_t s = ctx->stream;
  const InstanceDesc *table = instance_table(ctx, s, accels.data(), count);
  if (!table) return;
  const unsigned blocks = (unsigned)ctx->cuCount * kBlocksPerCU;
  const size_t need = (size_t)blocks * kTraceBlock * (kMaxStack - kLdsStack) * 4 + 256;
  if (need > ctx->traceBytes) {
    if (ctx->traceScratch) {
      hipStreamSynchronize(s);
      hipFree(ctx->traceScratch);
      ctx->traceScratch = nullptr;
      ctx->traceBytes = 0;
    }
    if (!hip_ok(hipMalloc(&ctx->traceScratch, need), "hipMalloc(trace scratch)")) return;
    ctx->traceBytes = need;
  }
  uint32_t *work = (uint32_t *)ctx->traceScratch;
  fill_u32(ctx, s, work, 0u, 1);
  launch(ctx, s, "trace_closest_multi", k_trace_multi, dim3(blocks), dim3(kTraceBlock), 0, table, count,
         (const GPRTRay *)rays->d, (GPRTHit *)hits->d, ids ? (int32_t *)ids->d : (int32_t *)nullptr, n, work,
         (uint32_t *)((char *)ctx->traceScratch + 256));
}

GPRT_API void gprtSynthBounceRays(GPRTContext _ctx, GPRTAccel *_accels, uint32_t count, GPRTBuffer _rays,
                                  GPRTBuffer _hits, GPRTBuffer _instanceIDs, uint32_t n, GPRTBuffer _out,
                                  uint32_t outCount, uint64_t seed, float tmin) {
  Context *ctx = (Context *)_ctx;
  Buffer *rays = (Buffer *)_rays, *hits = (Buffer *)_hits, *ids = (Buffer *)_instanceIDs, *out = (Buffer *)_out;
  if (!ctx || !_accels || !rays || !hits || !out || count == 0) {
    fail("gprtSynthBounceRays: null argument or no accels");
    return;
  }
  std::vector<Accel *> accels;
  if (!collect(_accels, count, accels, "gprtSynthBounceRays")) return;
  if ((size_t)n * sizeof(GPRTRay) > rays->size || (size_t)n * sizeof(GPRTHit) > hits->size ||
      (ids && (size_t)n * 4 > ids->size) || (size_t)outCount * sizeof(GPRTRay) > out->size) {
    fail("gprtSynthBounceRays: sizes exceed a buffer");
    return;
  }
  if (n == 0 || outCount == 0) return;
  hipStream_t s = ctx->stream;
  const InstanceDesc *table = instance_table(ctx, s, accels.data(), count);
  if (!table) return;
  const size_t scanBytes = scan_scratch_bytes(n);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t need = al(scanBytes) + 3 * al((size_t)n * 4);
  void *sp = ctx->ensure_arena(need);
  if (!sp) return;
  char *b = (char *)sp;
  uint32_t *state = (uint32_t *)b;
  uint32_t *flags = (uint32_t *)(b + al(scanBytes));
  uint32_t *offs = (uint32_t *)(b + al(scanBytes) + al((size_t)n * 4));
  uint32_t *list = (uint32_t *)(b + al(scanBytes) + 2 * al((size_t)n * 4));
  const unsigned g = (n + 255) / 256;
  launch(ctx, s, "synth_hit_flags", k_hit_flags, dim3(g), dim3(256), 0, (const GPRTHit *)hits->d, n, flags);
  scan_run(ctx, s, flags, offs, n, 0, false, state, nullptr);
  launch(ctx, s, "synth_hit_list", k_hit_list, dim3(g), dim3(256), 0, (const uint32_t *)flags,
         (const uint32_t *)offs, n, list);
  launch(ctx, s, "synth_bounce", k_bounce, dim3((outCount + 255) / 256), dim3(256), 0, table, count,
         (const GPRTRay *)rays->d, (const GPRTHit *)hits->d, ids ? (const int32_t *)ids->d : (const int32_t *)nullptr,
         (const uint32_t *)list, (const uint32_t *)state, outCount, seed, tmin, (GPRTRay *)out->d);
}
