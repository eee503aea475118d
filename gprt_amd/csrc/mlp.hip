// Learned treelet proposal for gfx950: a 102 -> 128 -> 128 -> 128 -> 4096 MLP in bf16 on MFMA
// with FP32 accumulation, trained on the device by hand-written forward / backward / Adam
// kernels, and a proposal-first closest-hit traversal whose result equals the plain one.
//
// The reference has no counterpart (its traversal is the Vulkan driver's); the closest relative
// is the hand-written Adam of samples/s9-differentiable/hostCode.cpp:66-69, :375-397. Standard
// Adam (bias-corrected first and second moments, beta1 0.9, beta2 0.999) is used here.
//
// Treelets: the LBVH is cut at depth 12; every node at depth 12, and every leaf shallower than
// that, roots one treelet (at most 4096). A ray's label is the treelet holding the primitive the
// plain traversal hit. At inference the top-1 treelet is traversed first to obtain a hit that
// tightens tmax, then the exact traversal runs from the root with that (t, primID) as its
// starting best. Because the candidate is a real hit and the closest-hit rule is the
// lexicographic minimum of (t, primID), the result cannot differ from the plain traversal.
//
// Layouts (all weights transposed, Wt[out][in], so a K-slice of one output is contiguous):
//   fp32 parameters, flat: W1t[128][128] b1[128] W2t[128][128] b2[128] W3t[128][128] b3[128]
//                          W4t[4096][128] b4[4096]   (input columns 102..127 of W1t stay 0)
//   bf16 copies of W1t..W4t for the MFMA kernels; Adam moments m, v in the fp32 layout.
//   Input features per ray (102): o.xyz, d.xyz, then for each of those 6 values v and each band
//   k = 0..7: sin(pi * 2^k * v), cos(pi * 2^k * v); padded with zeros to 128.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh.hpp"
#include "device_common.hpp"
#include "internal.hpp"
#include "mlp.hpp"

namespace gprt {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

union Frag {
  bf16x8_t v;
  uint4 u;
  uint16_t h[8];
};

__device__ __forceinline__ f32x4_t mfma16(const Frag &a, const Frag &b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
}

// FP32 -> bf16 with round-to-nearest-even (NaN stays NaN).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  union {
    float f;
    uint32_t u;
  } x;
  x.f = f;
  if ((x.u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((x.u >> 16) | 0x40u);
  const uint32_t r = 0x7FFFu + ((x.u >> 16) & 1u);
  return (uint16_t)((x.u + r) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ------------------------------------------------------------------------------------------------
// Input encoding, produced directly as 64 packed words (two bf16 per word, low half first):
// words 0..2 = (o.x,o.y) (o.z,d.x) (d.y,d.z); word 3 + 8c + k = (sin, cos)(2^k pi v_c); rest 0.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pack2(float lo, float hi) { return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16); }

__device__ __forceinline__ void encode_ray(const GPRTRay &r, uint32_t (&wds)[kMlpInPad / 2]) {
  const float v[6] = {r.ox, r.oy, r.oz, r.dx, r.dy, r.dz};
  wds[0] = pack2(v[0], v[1]);
  wds[1] = pack2(v[2], v[3]);
  wds[2] = pack2(v[4], v[5]);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    float x = v[c];  // 2^k * v is exact, so sin(pi * 2^k * v) needs no large-argument reduction
#pragma unroll
    for (int k = 0; k < kMlpBands; ++k) {
      float sn, cs;
      sincospif(x, &sn, &cs);
      wds[3 + c * kMlpBands + k] = pack2(sn, cs);
      x *= 2.0f;
    }
  }
#pragma unroll
  for (int i = 3 + 6 * kMlpBands; i < kMlpInPad / 2; ++i) wds[i] = 0u;
}

__device__ __forceinline__ void store_features(const uint32_t (&wds)[kMlpInPad / 2], uint16_t *dst) {
#pragma unroll
  for (int q = 0; q < kMlpInPad / 8; ++q)
    *(uint4 *)(dst + 8 * q) = make_uint4(wds[4 * q], wds[4 * q + 1], wds[4 * q + 2], wds[4 * q + 3]);
}

__global__ void __launch_bounds__(256) k_encode(const GPRTRay *__restrict__ rays, uint32_t count,
                                                uint16_t *__restrict__ X) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t wds[kMlpInPad / 2];
  encode_ray(rays[i], wds);
  store_features(wds, X + (size_t)i * kMlpInPad);
}

// ------------------------------------------------------------------------------------------------
// Generic bf16 GEMM on MFMA 16x16x32 with FP32 accumulation:
//   C[m][n] = alpha * sum_k A(m,k) B(k,n) (+ bias[n]) (* (mask(m,n) > 0)) (relu) (+= C)
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]: any transposition through strides.
// Workgroup tile 64x64 (4 waves of 32x32 = 2x2 MFMA tiles), K step 64 staged in LDS. Each operand
// is read 16 bytes at a time along whichever of its strides is 1 (K-contiguous rows are copied,
// M/N-contiguous columns are transposed into LDS); other layouts, unaligned operands and tile
// edges use element loads. Split K (split > 1): workgroup z covers the K range
// [z*kPer, (z+1)*kPer) and writes its partial sums to ws[z][M][N]; k_gemm_reduce adds the splits
// in order and applies the epilogue, so the result does not depend on scheduling.
// ------------------------------------------------------------------------------------------------
struct GemmArgs {
  const uint16_t *A;
  long long sam, sak;
  const uint16_t *B;
  long long sbk, sbn;
  float *Cf;
  uint16_t *Cb;
  long long scm, scn;
  int M, N, K;
  const float *bias;
  const uint16_t *mask;
  long long smm, smn;
  int relu;
  int accumulate;
  float alpha;
  int aMode, bMode;  // 0: element loads; 1: K-contiguous 16-byte loads; 2: M/N-contiguous 16-byte loads
  int split, kPer;
  float *ws;
};

constexpr int kGK = 64;    // K step
constexpr int kGPad = 72;  // LDS row length in elements (16-byte aligned rows)

// S[row][k] = P(r0 + row, k0 + k) for a 64 x 64 block, where P(r, k) = P[r*srow + k*sk], zero
// outside [0, rows) x [k0, kEnd).
__device__ __forceinline__ void gemm_stage(const uint16_t *__restrict__ P, long long srow, long long sk, int mode,
                                           int rows, int r0, int k0, int kEnd, uint16_t (*S)[kGPad], int tid) {
  if (mode == 1) {
#pragma unroll
    for (int c = tid; c < 512; c += 256) {
      const int row = c >> 3, kc = (c & 7) * 8;
      const int r = r0 + row, k = k0 + kc;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (r < rows && k + 8 <= kEnd) {
        v = *(const uint4 *)(P + (long long)r * srow + k);
      } else if (r < rows) {
        uint16_t e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) e[j] = k + j < kEnd ? P[(long long)r * srow + k + j] : (uint16_t)0;
        v = make_uint4(e[0] | (uint32_t)e[1] << 16, e[2] | (uint32_t)e[3] << 16, e[4] | (uint32_t)e[5] << 16,
                       e[6] | (uint32_t)e[7] << 16);
      }
      *(uint4 *)&S[row][kc] = v;
    }
  } else if (mode == 2) {
#pragma unroll
    for (int c = tid; c < 512; c += 256) {
      const int kk = c >> 3, rc = (c & 7) * 8;
      const int r = r0 + rc, k = k0 + kk;
      uint16_t e[8];
      if (k < kEnd && r + 8 <= rows) {
        const uint4 v = *(const uint4 *)(P + r + (long long)k * sk);
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          e[2 * j] = (uint16_t)(w4[j] & 0xFFFFu);
          e[2 * j + 1] = (uint16_t)(w4[j] >> 16);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) e[j] = (k < kEnd && r + j < rows) ? P[r + j + (long long)k * sk] : (uint16_t)0;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) S[rc + j][kk] = e[j];
    }
  } else {
    const int row = tid >> 2, kq = (tid & 3) * 16;
    const int r = r0 + row;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int k = k0 + kq + j;
      S[row][kq + j] = (r < rows && k < kEnd) ? P[(long long)r * srow + (long long)k * sk] : (uint16_t)0;
    }
  }
}

__device__ __forceinline__ void gemm_epilogue(const GemmArgs &g, int m, int n, float sum) {
  float v = g.alpha * sum;
  if (g.bias) v += g.bias[n];
  if (g.mask && !(bf2f(g.mask[(long long)m * g.smm + (long long)n * g.smn]) > 0.f)) v = 0.f;
  if (g.relu) v = fmaxf(v, 0.f);
  const long long o = (long long)m * g.scm + (long long)n * g.scn;
  if (g.Cf) {
    if (g.accumulate)
      g.Cf[o] += v;
    else
      g.Cf[o] = v;
  }
  if (g.Cb) g.Cb[o] = f2bf(v);
}

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64][kGPad];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64][kGPad];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int kBeg = blockIdx.z * g.kPer, kEnd = min(g.K, kBeg + g.kPer);
  f32x4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kBeg; k0 < kEnd; k0 += kGK) {
    gemm_stage(g.A, g.sam, g.sak, g.aMode, g.M, m0, k0, kEnd, As, tid);
    gemm_stage(g.B, g.sbn, g.sbk, g.bMode, g.N, n0, k0, kEnd, Bs, tid);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGK; ks += 32) {
      Frag a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i].u = *(const uint4 *)&As[wm + i * 16 + (l & 15)][ks + 8 * (l >> 4)];
        b[i].u = *(const uint4 *)&Bs[wn + i * 16 + (l & 15)][ks + 8 * (l >> 4)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + i * 16 + 4 * (l >> 4) + q;
        const int n = n0 + wn + j * 16 + (l & 15);
        if (m >= g.M || n >= g.N) continue;
        if (g.split > 1)
          g.ws[((long long)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][q];
        else
          gemm_epilogue(g, m, n, acc[i][j][q]);
      }
}

// Sums the split-K partials in split order, then applies the epilogue.
__global__ void __launch_bounds__(256) k_gemm_reduce(GemmArgs g) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long mn = (long long)g.M * g.N;
  if (i >= mn) return;
  float sum = 0.f;
  for (int z = 0; z < g.split; ++z) sum += g.ws[z * mn + i];
  gemm_epilogue(g, (int)(i / g.N), (int)(i % g.N), sum);
}

// ------------------------------------------------------------------------------------------------
// Softmax cross-entropy gradient: dZ = (softmax(z) - onehot(y)) * scale; rows with y < 0 give 0.
// One workgroup per row. loss += -log softmax(z)[y] * scale.
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_softmax_ce(const float *__restrict__ Z, const int32_t *__restrict__ labels,
                                                    uint16_t *__restrict__ dZ, float scale, float *__restrict__ loss,
                                                    uint32_t *__restrict__ correct) {
  static_assert(kMlpOut == 4096, "one row is 16 logits per thread");
  const uint32_t row = blockIdx.x;
  const int tid = threadIdx.x;
  const float *z = Z + (size_t)row * kMlpOut;
  uint16_t *dz = dZ + (size_t)row * kMlpOut;
  const int y = labels[row];
  __shared__ float s_red[4];
  __shared__ int s_idx[4];
  // The row is read once: thread tid holds logits j*1024 + 4*tid + q (increasing with j, q).
  float v[16];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float4 f = *(const float4 *)(z + j * 1024 + 4 * tid);
    v[4 * j] = f.x;
    v[4 * j + 1] = f.y;
    v[4 * j + 2] = f.z;
    v[4 * j + 3] = f.w;
  }
  float mx = -INFINITY;
  int amx = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    if (v[i] > mx) {
      mx = v[i];
      amx = (i >> 2) * 1024 + 4 * tid + (i & 3);
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const float om = __shfl_xor(mx, o, 64);
    const int oi = __shfl_xor(amx, o, 64);
    if (om > mx || (om == mx && oi < amx)) {
      mx = om;
      amx = oi;
    }
  }
  if ((tid & 63) == 0) {
    s_red[tid >> 6] = mx;
    s_idx[tid >> 6] = amx;
  }
  __syncthreads();
  mx = s_red[0];
  amx = s_idx[0];
  for (int w = 1; w < 4; ++w)
    if (s_red[w] > mx || (s_red[w] == mx && s_idx[w] < amx)) {
      mx = s_red[w];
      amx = s_idx[w];
    }
  __syncthreads();
  float e[16], sum = 0.f;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    e[i] = expf(v[i] - mx);
    sum += e[i];
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if ((tid & 63) == 0) s_red[tid >> 6] = sum;
  __syncthreads();
  sum = s_red[0] + s_red[1] + s_red[2] + s_red[3];
  const float inv = 1.0f / sum;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint16_t o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = j * 1024 + 4 * tid + q;
      const float g = y >= 0 ? (e[4 * j + q] * inv - (c == y ? 1.f : 0.f)) * scale : 0.f;
      o[q] = f2bf(g);
    }
    *(uint2 *)(dz + j * 1024 + 4 * tid) = make_uint2(o[0] | (uint32_t)o[1] << 16, o[2] | (uint32_t)o[3] << 16);
  }
  if (tid == 0 && y >= 0) {
    atomicAdd(loss, (logf(sum) + mx - z[y]) * scale);
    if (amx == y) atomicAdd(correct, 1u);
  }
}

// Column sums (bias gradients), deterministic in two passes: k_colsum_part sums a chunk of rows
// for 64 columns per workgroup (8 columns per thread with 16-byte loads when `vec`; 32 row lanes
// reduced in LDS in a fixed order) into part[chunk][n]; k_colsum_fin adds the chunks in order.
__global__ void __launch_bounds__(256) k_colsum_part(const uint16_t *__restrict__ A, int M, int N, int lda,
                                                     int rowsPerChunk, int vec, float *__restrict__ part) {
  __shared__ float s[32][65];
  const int cg = threadIdx.x & 7, rl = threadIdx.x >> 3;
  const int n0 = blockIdx.x * 64 + cg * 8;
  const int mBeg = blockIdx.y * rowsPerChunk, mEnd = min(M, mBeg + rowsPerChunk);
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int m = mBeg + rl; m < mEnd; m += 32) {
    const uint16_t *row = A + (size_t)m * lda;
    if (vec && n0 + 8 <= N) {
      const uint4 u = *(const uint4 *)(row + n0);
      const uint32_t w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[2 * j] += bf2f((uint16_t)(w4[j] & 0xFFFFu));
        acc[2 * j + 1] += bf2f((uint16_t)(w4[j] >> 16));
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (n0 + j < N) acc[j] += bf2f(row[n0 + j]);
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) s[rl][cg * 8 + j] = acc[j];
  __syncthreads();
  if (threadIdx.x < 64) {
    float t = 0.f;
    for (int r = 0; r < 32; ++r) t += s[r][threadIdx.x];
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n < N) part[(size_t)blockIdx.y * N + n] = t;
  }
}

__global__ void __launch_bounds__(256) k_colsum_fin(const float *__restrict__ part, int chunks, int N,
                                                    float *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float t = 0.f;
  for (int c = 0; c < chunks; ++c) t += part[(size_t)c * N + n];
  out[n] = t;
}

// Adam (bias-corrected) on the fp32 parameters; refreshes the bf16 weight copies.
__global__ void __launch_bounds__(256) k_adam(float *__restrict__ w, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, uint32_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2,
                                              const uint32_t *__restrict__ bfMap, uint16_t *__restrict__ wbf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  const float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
  w[i] = wi;
  const uint32_t b = bfMap[i];
  if (b != 0xFFFFFFFFu) wbf[b] = f2bf(wi);
}

__global__ void __launch_bounds__(256) k_to_bf16(const float *__restrict__ w, const uint32_t *__restrict__ bfMap,
                                                 uint16_t *__restrict__ wbf, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = bfMap[i];
  if (b != 0xFFFFFFFFu) wbf[b] = f2bf(w[i]);
}

// ------------------------------------------------------------------------------------------------
// Inference in one kernel, k_infer_top1, 256 rays per workgroup (64 per wave). Each wave encodes
// its rays and runs the three 128x128 hidden layers over a 64 x 128 activation tile in LDS (weight
// fragments read from L2); the layer-3 activations then move to registers and the 128x4096 output
// layer runs with a running argmax, W4t streamed through a two-chunk LDS ring that reuses the
// activation tiles' space. Neither the activations nor the logits are written to memory: per ray
// the kernel reads its 32-byte ray and writes a class (and optionally the winning logit). 80 KB of
// LDS and at most 256 registers let two workgroups (two waves per SIMD) share a CU and hide each
// other's stalls.
// ------------------------------------------------------------------------------------------------
constexpr int kInferRaysPerWave = 64;
constexpr int kInferRays = 4 * kInferRaysPerWave;
constexpr int kActStride = 136;  // bf16 per activation row in LDS (128 + 8 pad)

// One wave: encodes its 64 rays (rays [ray0, ray0 + 64), zero features past `count`) into its
// 64 x 128 activation tile `act` in LDS and runs the three 128x128 hidden layers over it; the
// bf16 layer-3 activations are left in `act`. Weight fragments are read from L2.
__device__ __forceinline__ void hidden_layers(const GPRTRay *__restrict__ rays, uint32_t count, uint32_t ray0,
                                              uint16_t *act, const uint16_t *__restrict__ W1t,
                                              const uint16_t *__restrict__ W2t, const uint16_t *__restrict__ W3t,
                                              const float *__restrict__ bias, int l) {
  const uint32_t ri = ray0 + l;
  {
    uint32_t wds[kMlpInPad / 2];
    if (ri < count) {
      encode_ray(load_ray_stream(rays + ri), wds);
    } else {
#pragma unroll
      for (int i = 0; i < kMlpInPad / 2; ++i) wds[i] = 0u;
    }
    store_features(wds, &act[l * kActStride]);
  }
  __builtin_amdgcn_wave_barrier();
  const uint16_t *Ws[3] = {W1t, W2t, W3t};
  for (int layer = 0; layer < 3; ++layer) {
    const uint16_t *W = Ws[layer];
    const float *b = bias + layer * kMlpHid;
    // The wave's whole 64 x 128 input tile goes to registers first, so each 16-column block of
    // outputs can be written back over it as soon as it is done (16 accumulators live at a time).
    Frag a[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        a[i][ks].u = *(const uint4 *)&act[(i * 16 + (l & 15)) * kActStride + ks * 32 + 8 * (l >> 4)];
    __builtin_amdgcn_wave_barrier();
#pragma unroll 2
    for (int j = 0; j < 8; ++j) {
      f32x4_t acc[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        Frag bf;
        bf.u = *(const uint4 *)&W[(j * 16 + (l & 15)) * 128 + ks * 32 + 8 * (l >> 4)];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = mfma16(a[i][ks], bf, acc[i]);
      }
      const int col = j * 16 + (l & 15);
      const float bc = b[col];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = i * 16 + 4 * (l >> 4) + q;
          act[row * kActStride + col] = __builtin_bit_cast(uint16_t, (__bf16)fmaxf(acc[i][q] + bc, 0.f));
        }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// The hidden layers run first (each wave over its 64 rays, its activation tile in the LDS that the
// W4t ring takes afterwards), so the layer-3 activations go from LDS straight to registers.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    k_infer_top1(uint32_t count, const uint16_t *__restrict__ W4t, const float *__restrict__ bias,
                 int32_t *__restrict__ outClass, float *__restrict__ outLogit, uint32_t numClasses, float gate,
                 const GPRTRay *__restrict__ rays, const uint16_t *__restrict__ W1t,
                 const uint16_t *__restrict__ W2t, const uint16_t *__restrict__ W3t) {
  extern __shared__ __attribute__((aligned(16))) uint16_t smem[];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const uint32_t ray0 = blockIdx.x * kInferRays;
  // A fragments (layer-3 activations of this wave's 64 rays) stay in registers.
  Frag a[4][4];
  {
    static_assert(4 * kInferRaysPerWave * kActStride * 2 <= 2 * 128 * 256 + kMlpOut * 4,
                  "the activation tiles fit in the ring and bias area");
    uint16_t *act = smem + (size_t)w * kInferRaysPerWave * kActStride;
    hidden_layers(rays, count, ray0 + w * kInferRaysPerWave, act, W1t, W2t, W3t, bias, l);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        a[i][ks].u = *(const uint4 *)&act[(i * 16 + (l & 15)) * kActStride + ks * 32 + 8 * (l >> 4)];
    __syncthreads();  // every wave has its fragments: the ring and the bias take the LDS
  }
  float best[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) best[i][q] = -FLT_MAX;
  // W4t streams through a two-chunk LDS ring by direct global-to-LDS loads: chunk c + 1 is in flight while chunk c is
  // multiplied, and the copies use no registers. A chunk is 128 rows x 256 bytes, unpadded, with
  // the 16-byte units of row r stored at position q ^ (r & 15), so the 16 rows read by one
  // B-fragment load fall into distinct units. The output bias is staged in LDS after the ring,
  // so the multiply loop issues no vector-memory loads of its own and the explicit vmcnt waits
  // below only count the ring copies.
  constexpr int kOutChunk = 128;
  constexpr int kChunks = kMlpOut / kOutChunk;
  constexpr int kChunkBytes = kOutChunk * 256;
  // Classes beyond numClasses get a bias of -FLT_MAX, so they never win and the multiply loop needs
  // no per-element check (a finite value: -inf would encode to a signalling NaN below, which
  // v_max_f32 turns into a quiet NaN that poisons the running maximum); the accumulators start at
  // the bias, so no add is needed either.
  float *sBias = (float *)((char *)smem + 2 * 128 * 256);
  for (int i = tid; i < kMlpOut; i += 256) sBias[i] = (uint32_t)i < numClasses ? bias[3 * kMlpHid + i] : -FLT_MAX;
  const uint32_t ringBase = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) void *)smem;
  // Wave w copies 1 KB blocks 8w .. 8w + 7 of a chunk (4 rows each).
  auto issue = [&](int chunk, int buf) {
#pragma unroll
    for (int sblk = 0; sblk < 8; ++sblk) {
      const int blk = w * 8 + sblk;
      const int r = 4 * blk + (l >> 4);
      const int q = (l & 15) ^ (r & 15);
      const uint16_t *g = W4t + ((size_t)(chunk * kOutChunk + r) * 128 + q * 8);
      const uint32_t dst = __builtin_amdgcn_readfirstlane(ringBase + (uint32_t)(buf * kChunkBytes + blk * 1024));
      asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(dst), "v"(g) : "memory", "m0");
    }
  };
  // Only the chunks that hold bound classes are multiplied: later classes could never win.
  const int nChunks = (int)min((uint32_t)kChunks, max(1u, (numClasses + kOutChunk - 1) / kOutChunk));
  issue(0, 0);
  for (int c = 0; c < nChunks; ++c) {
    if (c + 1 < nChunks) {
      issue(c + 1, (c + 1) & 1);
      __builtin_amdgcn_s_waitcnt(0x0F78);  // vmcnt(8): this wave's copies of chunk c have landed
    } else {
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    }
    __syncthreads();  // every wave's copies of chunk c are visible
    const uint16_t *cur = (const uint16_t *)((const char *)smem + (c & 1) * kChunkBytes);
    const int c0 = c * kOutChunk;
    // Column block j (16 classes) of this chunk: 16 MFMAs into accumulators seeded with the bias.
    auto block = [&](int j, f32x4_t (&acc)[4]) __attribute__((always_inline)) {
      const int r = j * 16 + (l & 15);
      const float b = sBias[c0 + r];
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[i] = f32x4_t{b, b, b, b};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        Frag bf;
        const int q = (ks * 4 + (l >> 4)) ^ (r & 15);
        bf.u = *(const uint4 *)&cur[r * 128 + q * 8];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = mfma16(a[i][ks], bf, acc[i]);
      }
    };
    // Running maximum with the class carried in the value: the low 12 mantissa bits of each
    // logit are replaced by 4095 - class (2 VALU operations per logit). Among logits equal after
    // the truncation (relative difference below 2^-11) the lower class wins when positive.
    auto keep = [&](int j, const f32x4_t (&acc)[4]) __attribute__((always_inline)) {
      const uint32_t code = 4095u - (uint32_t)(c0 + j * 16 + (l & 15));
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          best[i][q] = fmaxf(best[i][q], __uint_as_float((__float_as_uint(acc[i][q]) & 0xFFFFF000u) | code));
    };
    // Two accumulator sets: the MFMAs of block j + 1 run while block j's maxima are taken.
    f32x4_t accA[4], accB[4];
    block(0, accA);
#pragma unroll
    for (int j = 0; j < kOutChunk / 16; j += 2) {
      block(j + 1, accB);
      keep(j, accA);
      if (j + 2 < kOutChunk / 16) block(j + 2, accA);
      keep(j + 1, accB);
    }
    __syncthreads();  // chunk c's buffer is refilled in iteration c + 1
  }
  // Reduce over the 16 lanes that share rows (lane bits 0..3 index the column).
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float v = best[i][q];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
      int c = 4095 - (int)(__float_as_uint(v) & 0xFFFu);
      if ((uint32_t)c >= numClasses) c = 0;  // only without any finite logit
      if (v < gate) c = -1;                  // below the confidence gate: no proposal
      const uint32_t ri = ray0 + w * kInferRaysPerWave + i * 16 + 4 * (l >> 4) + q;
      if ((l & 15) == 0 && ri < count) {
        outClass[ri] = c;
        if (outLogit) outLogit[ri] = v;
      }
    }}

// ------------------------------------------------------------------------------------------------
// Treelets
// ------------------------------------------------------------------------------------------------
// One workgroup: breadth-first over the top `depth` levels (depth <= 12, so a level holds at most
// 4096 nodes). classRoot[c] = node code (inner index >= 0, or ~leaf); innerClass[node] = c for
// inner roots; leafClass[leaf] = c for leaves that are themselves roots (shallower than `depth`, or
// at it). Classes are numbered breadth-first, left to right within a level: each level's frontier
// is compacted by a workgroup prefix sum, never by atomics, so the numbering depends on the tree
// alone and a network trained against it stays valid when the proposal is bound again (in another
// process, after deserialisation, or after a rebuild that yields the same tree).
__global__ void __launch_bounds__(1024) k_treelet_cut(const BVHNode *__restrict__ nodes, uint32_t nprims, int depth,
                                                      int32_t *__restrict__ classRoot, int32_t *__restrict__ innerClass,
                                                      int32_t *__restrict__ leafClass, uint32_t *__restrict__ numClasses) {
  constexpr int kPer = 4;  // frontier entries per thread: 1024 x 4 = 4096
  __shared__ int32_t s_front[2][4096];
  __shared__ uint32_t s_wsum[1024 / 64];
  __shared__ uint32_t s_n[2];
  __shared__ uint32_t s_cls;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) {
    s_n[0] = 1;
    s_cls = 0;
    s_front[0][0] = nprims == 1 ? ~0 : 0;
  }
  __syncthreads();
  for (int d = 0; d <= depth; ++d) {
    const int cur = d & 1, nxt = cur ^ 1;
    const uint32_t n = s_n[cur], clsBase = s_cls;
    // Thread t takes entries [4t, 4t + 4). An entry is a class (a leaf, or a node at the cut
    // depth) or passes its two children to the next level; packed counts: classes in the high 16
    // bits, children in the low 16 (each total is at most 4096).
    int32_t code[kPer];
    uint32_t mine = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const uint32_t i = (uint32_t)(kPer * tid + j);
      code[j] = i < n ? s_front[cur][i] : 0;
      if (i < n) mine += (code[j] < 0 || d == depth) ? (1u << 16) : 2u;
    }
    const uint32_t incl = wave_incl_sum(mine);
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t off = incl - mine, total = 0;
    for (int w = 0; w < 1024 / 64; ++w) {
      const uint32_t x = s_wsum[w];
      if (w < wave) off += x;
      total += x;
    }
    uint32_t c = clsBase + (off >> 16), o = off & 0xFFFFu;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if ((uint32_t)(kPer * tid + j) >= n) break;
      if (code[j] < 0 || d == depth) {
        classRoot[c] = code[j];
        if (code[j] < 0)
          leafClass[~code[j]] = (int32_t)c;
        else
          innerClass[code[j]] = (int32_t)c;
        ++c;
      } else {
        const int4 dd = nodes[code[j]].d;
        s_front[nxt][o] = dd.x;
        s_front[nxt][o + 1] = dd.y;
        o += 2;
      }
    }
    __syncthreads();  // this level's counts and frontier are read before they change
    if (tid == 0) {
      s_n[nxt] = total & 0xFFFFu;
      s_cls = clsBase + (total >> 16);
    }
    __syncthreads();
  }
  if (tid == 0) *numClasses = s_cls;
}

// Every leaf below the cut inherits the class of its ancestor at the cut depth.
__global__ void __launch_bounds__(256) k_leaf_class(const uint32_t *__restrict__ parentLeaf,
                                                    const uint32_t *__restrict__ parentInner,
                                                    const int32_t *__restrict__ innerClass,
                                                    int32_t *__restrict__ leafClass, const uint32_t *__restrict__ ids,
                                                    int32_t *__restrict__ primClass, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int32_t c = leafClass[k];
  if (c < 0 && n > 1) {
    uint32_t link = parentLeaf[k];
    for (int s = 0; s < 72 && c < 0; ++s) {
      const uint32_t node = link >> 1;
      if (node >= n - 1) break;
      c = innerClass[node];
      link = parentInner[node];
      if (link == 0xFFFFFFFFu) break;
    }
    leafClass[k] = c;
  }
  const uint32_t prim = ids[k];
  if (prim < n) primClass[prim] = c;
}

__global__ void __launch_bounds__(256) k_labels(const GPRTHit *__restrict__ hits, uint32_t count,
                                                const int32_t *__restrict__ primClass, uint32_t nprims,
                                                int32_t *__restrict__ labels) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int32_t p = hits[i].primID;
  labels[i] = (p >= 0 && (uint32_t)p < nprims) ? primClass[p] : -1;
}

__global__ void __launch_bounds__(256) k_labels_of_leaves(const int32_t *__restrict__ leaves, uint32_t count,
                                                          const int32_t *__restrict__ leafClass, uint32_t nleaves,
                                                          int32_t *__restrict__ labels) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int32_t k = leaves[i];
  labels[i] = (k >= 0 && (uint32_t)k < nleaves) ? leafClass[k] : -1;
}

unsigned blocks_for(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// out[n] = sum over the M rows of A (bf16, row stride lda), through the GEMM workspace.
void colsum(Context *ctx, hipStream_t s, const uint16_t *A, int M, int N, int lda, float *out, const Proposal *p) {
  int rows = 256;
  while ((size_t)((M + rows - 1) / rows) * N * 4 > p->gemmWsBytes) rows *= 2;
  const int chunks = std::max(1, (M + rows - 1) / rows);
  const int vec = ((uintptr_t)A & 15u) == 0 && lda % 8 == 0;
  launch(ctx, s, "mlp_colsum", k_colsum_part, dim3((N + 63) / 64, chunks), dim3(256), 0, A, M, N, lda, rows, vec,
         p->gemmWs);
  launch(ctx, s, "mlp_colsum", k_colsum_fin, dim3(blocks_for(N, 256)), dim3(256), 0, (const float *)p->gemmWs, chunks,
         N, out);
}

// Picks each operand's load mode and the K split, then launches (and reduces the splits).
void gemm(Context *ctx, hipStream_t s, GemmArgs g, const char *name, const Proposal *p) {
  auto vec_ok = [](const void *ptr, long long otherStride) {
    return ((uintptr_t)ptr & 15u) == 0 && otherStride % 8 == 0;
  };
  g.aMode = (g.sak == 1 && vec_ok(g.A, g.sam)) ? 1 : (g.sam == 1 && vec_ok(g.A, g.sak)) ? 2 : 0;
  g.bMode = (g.sbk == 1 && vec_ok(g.B, g.sbn)) ? 1 : (g.sbn == 1 && vec_ok(g.B, g.sbk)) ? 2 : 0;
  const long long tiles = (long long)((g.N + 63) / 64) * ((g.M + 63) / 64);
  // Few output tiles over a long K (weight gradients): split K so ~512 workgroups run, each over
  // at least 256 of K, as far as the workspace holds the partials.
  int split = 1;
  if (tiles < 256 && g.K >= 512 && p->gemmWs) {
    split = (int)std::min<long long>(std::max<long long>(512 / tiles, 1), g.K / 256);
    while (split > 1 && (size_t)split * g.M * g.N * 4 > p->gemmWsBytes) --split;
  }
  g.kPer = ((g.K + split - 1) / split + kGK - 1) / kGK * kGK;
  g.split = (g.K + g.kPer - 1) / g.kPer;
  if (g.split < 1) g.split = 1;
  g.ws = p->gemmWs;
  dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, g.split);
  launch(ctx, s, name, k_gemm, grid, dim3(256), 0, g);
  if (g.split > 1)
    launch(ctx, s, "mlp_gemm_reduce", k_gemm_reduce, dim3((unsigned)(((long long)g.M * g.N + 255) / 256)), dim3(256),
           0, g);
}

GemmArgs gemm_args(const uint16_t *A, long long sam, long long sak, const uint16_t *B, long long sbk, long long sbn,
                   int M, int N, int K) {
  GemmArgs g;
  std::memset(&g, 0, sizeof(g));
  g.A = A;
  g.sam = sam;
  g.sak = sak;
  g.B = B;
  g.sbk = sbk;
  g.sbn = sbn;
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = 1.f;
  return g;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------------
size_t mlp_param_count() { return kMlpParams; }

bool mlp_alloc(Context *ctx, Proposal *p, uint32_t maxBatch) {
  (void)ctx;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t P = kMlpParams;
  size_t off = 0;
  const size_t oW = off;
  off += al(P * 4);
  const size_t oG = off;
  off += al(P * 4);
  const size_t oM = off;
  off += al(P * 4);
  const size_t oV = off;
  off += al(P * 4);
  const size_t oMap = off;
  off += al(P * 4);
  const size_t oBf = off;
  off += al(kMlpWeights * 2);
  const size_t oBias = off;
  off += al(kMlpBiases * 4);
  const size_t oStat = off;
  off += al(64);
  const size_t B = maxBatch;
  const size_t oX = off;
  off += al(B * kMlpInPad * 2);
  const size_t oH = off;
  off += al(3 * B * kMlpHid * 2);
  const size_t oD = off;
  off += al(3 * B * kMlpHid * 2);
  const size_t oZ = off;
  off += al(B * kMlpOut * 4);
  const size_t oDZ = off;
  off += al(B * kMlpOut * 2);
  const size_t oL = off;
  off += al(B * 4);
  const size_t oWs = off;  // split-K partial sums of the training GEMMs
  const size_t wsBytes = (size_t)16 << 20;
  off += al(wsBytes);
  if (!hip_ok(hipMalloc(&p->block, off), "hipMalloc(proposal)")) return false;
  p->blockBytes = off;
  char *b = (char *)p->block;
  p->w = (float *)(b + oW);
  p->g = (float *)(b + oG);
  p->m = (float *)(b + oM);
  p->v = (float *)(b + oV);
  p->bfMap = (uint32_t *)(b + oMap);
  p->wbf = (uint16_t *)(b + oBf);
  p->bias = (float *)(b + oBias);
  p->stats = (float *)(b + oStat);
  p->X = (uint16_t *)(b + oX);
  p->H = (uint16_t *)(b + oH);
  p->dA = (uint16_t *)(b + oD);
  p->Z = (float *)(b + oZ);
  p->dZ = (uint16_t *)(b + oDZ);
  p->labels = (int32_t *)(b + oL);
  p->gemmWs = (float *)(b + oWs);
  p->gemmWsBytes = wsBytes;
  p->maxBatch = maxBatch;
  // Map from flat fp32 parameter index to its bf16 copy (weights only).
  std::vector<uint32_t> map(P, 0xFFFFFFFFu);
  for (int layer = 0; layer < 4; ++layer) {
    const size_t wOff = mlp_weight_offset(layer), n = mlp_weight_size(layer), bfOff = mlp_bf16_offset(layer);
    for (size_t i = 0; i < n; ++i) map[wOff + i] = (uint32_t)(bfOff + i);
  }
  if (!hip_ok(hipMemcpy(p->bfMap, map.data(), P * 4, hipMemcpyHostToDevice), "hipMemcpy(bfMap)")) return false;
  return true;
}

void mlp_release(Proposal *p) {
  if (p->block) hipFree(p->block);
  if (p->treeBlock) hipFree(p->treeBlock);
  p->block = p->treeBlock = nullptr;
}

// Copies fp32 parameters to the device and refreshes the bf16 weights and the bias vector.
bool mlp_set_params(Context *ctx, hipStream_t s, Proposal *p, const float *host) {
  if (!hip_ok(hipMemcpyAsync(p->w, host, kMlpParams * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync(params)"))
    return false;
  launch(ctx, s, "mlp_to_bf16", k_to_bf16, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0, (const float *)p->w,
         (const uint32_t *)p->bfMap, p->wbf, (uint32_t)kMlpParams);
  fill_u32(ctx, s, p->m, 0u, kMlpParams);
  fill_u32(ctx, s, p->v, 0u, kMlpParams);
  p->step = 0;
  return mlp_refresh_bias(ctx, s, p);
}

bool mlp_refresh_bias(Context *ctx, hipStream_t s, Proposal *p) {
  (void)ctx;
  for (int layer = 0; layer < 4; ++layer) {
    const size_t n = layer < 3 ? kMlpHid : kMlpOut;
    if (!hip_ok(hipMemcpyAsync(p->bias + layer * kMlpHid, p->w + mlp_bias_offset(layer), n * 4,
                               hipMemcpyDeviceToDevice, s),
                "hipMemcpyAsync(bias)"))
      return false;
  }
  return true;
}

const uint16_t *mlp_wbf(const Proposal *p, int layer) { return p->wbf + mlp_bf16_offset(layer); }

// Forward of a batch with stored activations (training / logits). Z gets the fp32 logits.
void mlp_forward_batch(Context *ctx, hipStream_t s, Proposal *p, const void *rays, uint32_t n) {
  launch(ctx, s, "mlp_encode", k_encode, dim3(blocks_for(n, 256)), dim3(256), 0, (const GPRTRay *)rays, n, p->X);
  const uint16_t *in = p->X;
  int inStride = kMlpInPad;
  for (int layer = 0; layer < 3; ++layer) {
    uint16_t *out = p->H + (size_t)layer * p->maxBatch * kMlpHid;
    GemmArgs g = gemm_args(in, inStride, 1, mlp_wbf(p, layer), 1, kMlpInPad, (int)n, kMlpHid, kMlpInPad);
    g.Cb = out;
    g.scm = kMlpHid;
    g.scn = 1;
    g.bias = p->w + mlp_bias_offset(layer);
    g.relu = 1;
    gemm(ctx, s, g, "mlp_gemm_fwd", p);
    in = out;
    inStride = kMlpHid;
  }
  GemmArgs g = gemm_args(in, kMlpHid, 1, mlp_wbf(p, 3), 1, kMlpHid, (int)n, kMlpOut, kMlpHid);
  g.Cf = p->Z;
  g.scm = kMlpOut;
  g.scn = 1;
  g.bias = p->w + mlp_bias_offset(3);
  gemm(ctx, s, g, "mlp_gemm_fwd", p);
}

// Gradients of the mean cross-entropy over the labelled rays of one batch (into p->g).
void mlp_backward_batch(Context *ctx, hipStream_t s, Proposal *p, uint32_t n, uint32_t labelled) {
  const float scale = labelled ? 1.0f / (float)labelled : 0.f;
  fill_u32(ctx, s, p->g, 0u, kMlpParams);  // each gradient entry is then written exactly once
  launch(ctx, s, "mlp_softmax_ce", k_softmax_ce, dim3(n), dim3(256), 0, (const float *)p->Z,
         (const int32_t *)p->labels, p->dZ, scale, p->stats, (uint32_t *)(p->stats + 1));
  const size_t B = p->maxBatch;
  const uint16_t *H[3] = {p->H, p->H + B * kMlpHid, p->H + 2 * B * kMlpHid};
  // Output layer: dW4t = dZ^T H3 ; db4 = colsum(dZ) ; dA3 = (dZ W4t) * (H3 > 0)
  {
    GemmArgs g = gemm_args(p->dZ, 1, kMlpOut, H[2], kMlpHid, 1, kMlpOut, kMlpHid, (int)n);
    g.Cf = p->g + mlp_weight_offset(3);
    g.scm = kMlpHid;
    g.scn = 1;
    gemm(ctx, s, g, "mlp_gemm_bwd", p);
    colsum(ctx, s, p->dZ, (int)n, kMlpOut, kMlpOut, p->g + mlp_bias_offset(3), p);
    GemmArgs h = gemm_args(p->dZ, kMlpOut, 1, mlp_wbf(p, 3), kMlpHid, 1, (int)n, kMlpHid, kMlpOut);
    h.Cb = p->dA + 2 * B * kMlpHid;
    h.scm = kMlpHid;
    h.scn = 1;
    h.mask = H[2];
    h.smm = kMlpHid;
    h.smn = 1;
    gemm(ctx, s, h, "mlp_gemm_bwd", p);
  }
  for (int layer = 2; layer >= 0; --layer) {
    const uint16_t *dA = p->dA + (size_t)layer * B * kMlpHid;
    const uint16_t *in = layer == 0 ? p->X : H[layer - 1];
    const int K = kMlpInPad;  // all hidden-layer inputs are 128 wide
    GemmArgs g = gemm_args(dA, 1, kMlpHid, in, K, 1, kMlpHid, K, (int)n);
    g.Cf = p->g + mlp_weight_offset(layer);
    g.scm = K;
    g.scn = 1;
    gemm(ctx, s, g, "mlp_gemm_bwd", p);
    colsum(ctx, s, dA, (int)n, kMlpHid, kMlpHid, p->g + mlp_bias_offset(layer), p);
    if (layer > 0) {
      GemmArgs h = gemm_args(dA, kMlpHid, 1, mlp_wbf(p, layer), kMlpInPad, 1, (int)n, kMlpHid, kMlpHid);
      h.Cb = p->dA + (size_t)(layer - 1) * B * kMlpHid;
      h.scm = kMlpHid;
      h.scn = 1;
      h.mask = H[layer - 1];
      h.smm = kMlpHid;
      h.smn = 1;
      gemm(ctx, s, h, "mlp_gemm_bwd", p);
    }
  }
}

void mlp_adam(Context *ctx, hipStream_t s, Proposal *p, float lr) {
  p->step += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, (float)p->step), c2 = 1.f - std::pow(b2, (float)p->step);
  launch(ctx, s, "mlp_adam", k_adam, dim3(blocks_for(kMlpParams, 256)), dim3(256), 0, p->w, (const float *)p->g, p->m,
         p->v, (uint32_t)kMlpParams, lr, b1, b2, eps, c1, c2, (const uint32_t *)p->bfMap, p->wbf);
  mlp_refresh_bias(ctx, s, p);
}

void mlp_labels_of_leaves(Context *ctx, hipStream_t s, Proposal *p, const int32_t *leaves, uint32_t n,
                          int32_t *labels) {
  launch(ctx, s, "mlp_labels", k_labels_of_leaves, dim3(blocks_for(n, 256)), dim3(256), 0, leaves, n,
         tlas_leaf_class(p->accel), p->accel->n, labels);
}

void mlp_labels(Context *ctx, hipStream_t s, Proposal *p, const void *hits, uint32_t n, int32_t *labels) {
  launch(ctx, s, "mlp_labels", k_labels, dim3(blocks_for(n, 256)), dim3(256), 0, (const GPRTHit *)hits, n,
         (const int32_t *)p->primClass, p->treePrims, labels);
}

void mlp_infer(Context *ctx, hipStream_t s, Proposal *p, const void *rays, uint32_t n, int32_t *cls, float *logit,
               float gate) {
  const uint32_t classes = p->numClasses ? p->numClasses : (uint32_t)kMlpOut;
  const size_t ldsOut = 2 * 128 * 256 + (size_t)kMlpOut * 4;  // W4t ring + output bias
  launch(ctx, s, "mlp_infer_fused", k_infer_top1, dim3(blocks_for(n, kInferRays)), dim3(256), ldsOut, n,
         mlp_wbf(p, 3), (const float *)p->bias, cls, logit, classes, gate,
         (const GPRTRay *)rays, mlp_wbf(p, 0), mlp_wbf(p, 1), mlp_wbf(p, 2));
}

bool mlp_treelets(Context *ctx, hipStream_t s, Proposal *p, Accel *a, int depth) {
  const uint32_t n = a->n;
  if (p->treeBlock && p->treeCap < n) {
    hipStreamSynchronize(s);
    hipFree(p->treeBlock);
    p->treeBlock = nullptr;
  }
  if (!p->treeBlock) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t bytes = al(4096 * 4) + al((size_t)n * 4) * 3 + al(64);
    if (!hip_ok(hipMalloc(&p->treeBlock, bytes), "hipMalloc(treelets)")) return false;
    char *b = (char *)p->treeBlock;
    p->classRoot = (int32_t *)b;
    p->innerClass = (int32_t *)(b + al(4096 * 4));
    p->leafClass = (int32_t *)(b + al(4096 * 4) + al((size_t)n * 4));
    p->primClass = (int32_t *)(b + al(4096 * 4) + 2 * al((size_t)n * 4));
    p->numClassesDev = (uint32_t *)(b + al(4096 * 4) + 3 * al((size_t)n * 4));
    p->treeCap = n;
  }
  BVHStore *st = a->store;
  const MeshDesc &d = st->desc[a->mesh];
  fill_u32(ctx, s, p->innerClass, 0xFFFFFFFFu, n);
  fill_u32(ctx, s, p->leafClass, 0xFFFFFFFFu, n);
  launch(ctx, s, "mlp_treelet_cut", k_treelet_cut, dim3(1), dim3(1024), 0, (const BVHNode *)a->nodes, n, depth,
         p->classRoot, p->innerClass, p->leafClass, p->numClassesDev);
  lbvh_parents(ctx, s, a);
  launch(ctx, s, "mlp_leaf_class", k_leaf_class, dim3(blocks_for(n, 256)), dim3(256), 0,
         (const uint32_t *)(st->parentLeaf + d.primOffset), (const uint32_t *)(st->parentInner + d.innerOffset),
         (const int32_t *)p->innerClass, p->leafClass, (const uint32_t *)(st->ids + d.primOffset), p->primClass, n);
  uint32_t nc = 0;
  if (!hip_ok(hipMemcpyAsync(&nc, p->numClassesDev, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(classes)"))
    return false;
  if (!hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize(treelets)")) return false;
  p->numClasses = nc;
  p->treePrims = n;
  p->accel = a;
  return true;
}

}  // namespace gprt
