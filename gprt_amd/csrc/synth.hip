// Synthetic workload generators on the device (benchmarks and tests): the blob meshes and the
// per-frame vertex perturbation of the dynamic scene. They follow gprt_amd/scenes.py exactly:
// counter-based SplitMix64 streams keyed by (seed, stream), 24-bit uniforms, FP64 Box-Muller,
// one rounding to FP32 at the end. The device's FP64 log / cos may differ from the host's in the
// last bit, so a rare coordinate can differ by one FP32 ulp from the numpy generator; parity tests
// therefore read the generated geometry back and give the oracle exactly what the GPU built from.
#include <vector>

#include "internal.hpp"

namespace gprt {
namespace {

constexpr uint64_t kGold = 0x9E3779B97F4A7C15ull;

__host__ __device__ __forceinline__ uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix(seed * kGold + stream * 0xD1B54A32D192ED03ull + 1ull);
}

__device__ __forceinline__ double rand01(uint64_t key, uint64_t ctr) {
  return (double)(mix(key + (ctr + 1ull) * kGold) >> 40) * (1.0 / 16777216.0);
}

// Blob parameters per (mesh, blob): centre xyz and sigma, in FP64.
__global__ void __launch_bounds__(256) k_blob_params(uint64_t seed0, uint32_t meshes, int blobs,
                                                     double *__restrict__ params) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= meshes * (uint32_t)blobs) return;
  const uint32_t m = i / blobs, b = i % blobs;
  const uint64_t seed = seed0 + m;
  const uint64_t k10 = stream_key(seed, 10), k11 = stream_key(seed, 11);
  params[4 * i + 0] = rand01(k10, 3 * b);
  params[4 * i + 1] = rand01(k10, 3 * b + 1);
  params[4 * i + 2] = rand01(k10, 3 * b + 2);
  params[4 * i + 3] = exp(log(0.02) + rand01(k11, b) * (log(0.1) - log(0.02)));
}

__global__ void __launch_bounds__(256) k_blob_tris(const int64_t *__restrict__ offs, uint32_t meshes, uint64_t seed0,
                                                   int blobs, double offset, const double *__restrict__ params,
                                                   uint64_t total, float *__restrict__ verts,
                                                   uint32_t *__restrict__ idx) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  uint32_t lo = 0, hi = meshes;  // largest m with offs[m] <= t
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((uint64_t)offs[mid] <= t) lo = mid; else hi = mid;
  }
  const uint32_t m = lo;
  const uint64_t j = t - (uint64_t)offs[m];
  const uint64_t n = (uint64_t)(offs[m + 1] - offs[m]);
  const uint32_t b = (uint32_t)((j * (uint64_t)blobs) / n);
  const uint64_t seed = seed0 + m;
  const uint64_t k12 = stream_key(seed, 12), k13 = stream_key(seed, 13), k14 = stream_key(seed, 14);
  const double *p = params + 4 * ((size_t)m * blobs + b);
  double cent[3];
  for (int q = 0; q < 3; ++q) {
    const double u1 = 1.0 - rand01(k12, 3 * j + q);
    const double u2 = rand01(k13, 3 * j + q);
    cent[q] = p[q] + p[3] * (sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2));
  }
  for (int v = 0; v < 3; ++v)
    for (int q = 0; q < 3; ++q) {
      const double off = rand01(k14, 9 * j + 3 * v + q) * (2.0 * offset) - offset;
      verts[(3 * t + v) * 3 + q] = (float)(cent[q] + off);
    }
  for (int v = 0; v < 3; ++v) idx[3 * t + v] = (uint32_t)(3 * j + v);
}

// dst = src + U(-amp, amp) per coordinate (FP64 then one rounding), seeded per call.
__global__ void __launch_bounds__(256) k_perturb(const float *__restrict__ src, float *__restrict__ dst, uint64_t n,
                                                 uint64_t seed, double amp) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = stream_key(seed, 60);
  dst[i] = (float)((double)src[i] + (rand01(key, i) * 2.0 - 1.0) * amp);
}

}  // namespace
}  // namespace gprt

using namespace gprt;

GPRT_API void gprtSynthBlobMeshes(GPRTContext _ctx, const int64_t *sizes, uint32_t meshes, uint64_t seed0,
                                  float offset, GPRTBuffer _verts, GPRTBuffer _idx) {
  Context *ctx = (Context *)_ctx;
  Buffer *vb = (Buffer *)_verts, *ib = (Buffer *)_idx;
  if (!ctx || !sizes || !vb || !ib || meshes == 0) {
    fail("gprtSynthBlobMeshes: null argument or no meshes");
    return;
  }
  std::vector<int64_t> offs(meshes + 1, 0);
  for (uint32_t m = 0; m < meshes; ++m) {
    if (sizes[m] <= 0) {
      fail("gprtSynthBlobMeshes: mesh %u has %lld triangles", m, (long long)sizes[m]);
      return;
    }
    offs[m + 1] = offs[m] + sizes[m];
  }
  const uint64_t total = (uint64_t)offs[meshes];
  if (vb->size < total * 36 && !buffer_realloc(vb, total * 36, false)) return;
  if (ib->size < total * 12 && !buffer_realloc(ib, total * 12, false)) return;
  const int blobs = 64;
  void *tmp = nullptr;
  const size_t paramBytes = (size_t)meshes * blobs * 4 * sizeof(double);
  if (!hip_ok(hipMalloc(&tmp, paramBytes + (meshes + 1) * 8), "hipMalloc(synth)")) return;
  double *params = (double *)tmp;
  int64_t *dOffs = (int64_t *)((char *)tmp + paramBytes);
  hipStream_t s = ctx->stream;
  hip_ok(hipMemcpyAsync(dOffs, offs.data(), (meshes + 1) * 8, hipMemcpyHostToDevice, s), "hipMemcpyAsync(offs)");
  launch(ctx, s, "synth_blob_params", k_blob_params, dim3((meshes * blobs + 255) / 256), dim3(256), 0, seed0, meshes,
         blobs, params);
  launch(ctx, s, "synth_blob_tris", k_blob_tris, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
         (const int64_t *)dOffs, meshes, seed0, blobs, (double)offset, (const double *)params, total,
         (float *)vb->d, (uint32_t *)ib->d);
  hipStreamSynchronize(s);
  hipFree(tmp);
}

GPRT_API void gprtSynthPerturb(GPRTContext _ctx, GPRTBuffer _src, GPRTBuffer _dst, uint64_t floats, uint64_t seed,
                               float amplitude) {
  Context *ctx = (Context *)_ctx;
  Buffer *src = (Buffer *)_src, *dst = (Buffer *)_dst;
  if (!ctx || !src || !dst) {
    fail("gprtSynthPerturb: null argument");
    return;
  }
  if (floats * 4 > src->size || floats * 4 > dst->size) {
    fail("gprtSynthPerturb: %llu floats exceed a buffer", (unsigned long long)floats);
    return;
  }
  if (!floats) return;
  launch(ctx, ctx->stream, "synth_perturb", k_perturb, dim3((unsigned)((floats + 255) / 256)), dim3(256), 0,
         (const float *)src->d, (float *)dst->d, floats, seed, (double)amplitude);
}
