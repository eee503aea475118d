// Stable LSD radix sort (keys 32/64-bit, optional 32/64-bit payload) for gfx950.
//
// Semantics follow GPRT's buffer sort (gprt/gprt.cpp:8821-9073, gprt/gprt_sort.slang): ascending,
// stable, in place, payload follows its key. The reference runs 16 passes of 4-bit digits, each
// as five dispatches (count, count-reduce, scan, scan-add, scatter; gprt/gprt_sort.slang:109-637).
// This implementation is a one-sweep design instead:
//  * k_hist reads the keys once and builds the digit histograms of every pass (LDS atomics);
//  * k_hist_scan turns them into global digit offsets and marks passes whose digit is constant
//    over all keys as inactive (they would be identity permutations);
//  * one k_onesweep launch per pass: a tile ticket, wave64 multisplit ranking (BITS ballots per
//    key give each lane the mask of peers with the same digit), per-digit decoupled look-back
//    across tiles (four predecessors per step), a stable local reorder in LDS and a global write
//    in digit runs;
//  * ping-pong buffers are picked on the device from the active-pass mask, and k_final_copy
//    moves the result back when an odd number of passes ran — no host round trip, so the whole
//    sort can be captured in a hipGraph.
#include "device_common.hpp"
#include "internal.hpp"

namespace gprt {

namespace {

constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;

// State layout (uint32 words):
//   [0, P*R)            digit counts, then exclusive digit starts, per pass
//   [P*R]               active-pass mask
//   [P*R + 1, +P)       tile tickets per pass
//   [P*R + 32, ...)     look-back status [P][tiles][R]
struct SortState {
  uint32_t *hist;
  uint32_t *activeMask;
  uint32_t *tickets;
  uint32_t *status;
};

__host__ __device__ inline SortState state_view(void *base, int passes, int radix) {
  uint32_t *p = (uint32_t *)base;
  SortState s;
  s.hist = p;
  s.activeMask = p + (size_t)passes * radix;
  s.tickets = s.activeMask + 1;
  s.status = p + (size_t)passes * radix + 32;
  return s;
}

template <typename KT>
__device__ __forceinline__ uint32_t digit_of(KT k, int shift, uint32_t mask) {
  return (uint32_t)(k >> shift) & mask;
}

// ------------------------------------------------------------------------------------------------
// Histogram of all passes.
// ------------------------------------------------------------------------------------------------
template <typename KT, int BITS>
__global__ void __launch_bounds__(256)
    k_hist(const KT *__restrict__ keys, size_t n, int loBit, int hiBit, int passes, uint32_t *__restrict__ hist) {
  constexpr int R = 1 << BITS;
  extern __shared__ uint32_t s_h[];  // [passes][R]
  for (int i = threadIdx.x; i < passes * R; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  // kHistBatch keys per thread at a time, each load coalesced over the grid: the batch's loads are
  // all in flight before the first key's counts (a load per thread at a time leaves HBM idle).
  constexpr int kHistBatch = 8;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += stride * kHistBatch) {
    KT k[kHistBatch];
#pragma unroll
    for (int u = 0; u < kHistBatch; ++u) {
      const size_t i = i0 + (size_t)u * stride;
      k[u] = i < n ? keys[i] : (KT)0;
    }
#pragma unroll
    for (int u = 0; u < kHistBatch; ++u) {
      if (i0 + (size_t)u * stride >= n) break;  // the rest of the batch is past the end too
      for (int p = 0; p < passes; ++p) {
        const int shift = loBit + p * BITS;
        const int bits = (hiBit - shift) < BITS ? (hiBit - shift) : BITS;
        const uint32_t d = digit_of(k[u], shift, (1u << bits) - 1u);
        atomicAdd(&s_h[p * R + d], 1u);
      }
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < passes * R; j += blockDim.x) {
    const uint32_t c = s_h[j];
    if (c) atomicAdd(&hist[j], c);
  }
}

// Exclusive scan of every pass's histogram; a pass is active unless one digit holds every key.
template <int BITS>
__global__ void __launch_bounds__(256) k_hist_scan(uint32_t *__restrict__ hist, uint32_t *__restrict__ activeMask,
                                                   int passes, uint32_t n) {
  constexpr int R = 1 << BITS;
  constexpr int PER = (R + 255) / 256;
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_any;
  const int tid = threadIdx.x;
  uint32_t mask = 0;
  for (int p = 0; p < passes; ++p) {
    uint32_t c[PER];
    uint32_t sum = 0;
    bool full = false;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int d = tid * PER + j;
      c[j] = d < R ? hist[p * R + d] : 0u;
      full |= (c[j] == n);
      sum += c[j];
    }
    if (tid == 0) s_any = 0;
    __syncthreads();
    if (full) s_any = 1;
    const uint32_t incl = wave_incl_sum(sum);
    if ((tid & 63) == 63) s_w[tid >> 6] = incl;
    __syncthreads();
    uint32_t off = incl - sum;
    for (int w = 0; w < (tid >> 6); ++w) off += s_w[w];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int d = tid * PER + j;
      if (d < R) hist[p * R + d] = off;
      off += c[j];
    }
    __syncthreads();
    if (!s_any) mask |= 1u << p;
    __syncthreads();
  }
  if (tid == 0) *activeMask = mask;
}

// ------------------------------------------------------------------------------------------------
// One pass: ticketed tiles, wave multisplit ranking, per-digit look-back, LDS reorder, scatter.
// ------------------------------------------------------------------------------------------------
template <int VB>
struct ValType;
template <>
struct ValType<0> {
  using T = uint32_t;
};
template <>
struct ValType<32> {
  using T = uint32_t;
};
template <>
struct ValType<64> {
  using T = uint64_t;
};

template <typename KT, int VB, int BITS, int THREADS, int ITEMS>
__global__ void __launch_bounds__(THREADS)
    k_onesweep(KT *__restrict__ keys0, KT *__restrict__ keys1, void *__restrict__ vals0, void *__restrict__ vals1,
               uint32_t n, int pass, int shift, int hiBit, void *__restrict__ stateBase, int passes,
               uint32_t numTiles) {
  using VT = typename ValType<VB>::T;
  constexpr int R = 1 << BITS;
  constexpr int W = THREADS / 64;
  constexpr int TILE = THREADS * ITEMS;
  constexpr int DPT = (R + THREADS - 1) / THREADS;  // digits per thread
  constexpr size_t KBYTES = sizeof(KT) * TILE;
  constexpr size_t VBYTES = VB ? sizeof(VT) * TILE : 0;
  constexpr size_t BUFBYTES = KBYTES > VBYTES ? KBYTES : VBYTES;

  __shared__ __attribute__((aligned(16))) unsigned char s_buf[BUFBYTES];
  __shared__ uint32_t s_whist[W][R];
  __shared__ uint32_t s_tdigit[R];
  __shared__ uint32_t s_gbase[R];
  __shared__ uint32_t s_wtot[W];
  __shared__ uint32_t s_tile;

  SortState st = state_view(stateBase, passes, R);
  const uint32_t active = *st.activeMask;
  if (!((active >> pass) & 1u)) return;
  const int before = __builtin_popcount(active & ((1u << pass) - 1u));
  const KT *__restrict__ src = (before & 1) ? keys1 : keys0;
  KT *__restrict__ dst = (before & 1) ? keys0 : keys1;
  const VT *__restrict__ vsrc = (const VT *)((before & 1) ? vals1 : vals0);
  VT *__restrict__ vdst = (VT *)((before & 1) ? vals0 : vals1);

  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int l = tid & 63;
  const int bits = (hiBit - shift) < BITS ? (hiBit - shift) : BITS;
  const uint32_t dmask = (1u << bits) - 1u;

  if (tid == 0) s_tile = ticket(&st.tickets[pass]);
  for (int i = tid; i < W * R; i += THREADS) (&s_whist[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const size_t tileBase = (size_t)tile * TILE;
  const size_t waveBase = tileBase + (size_t)w * (ITEMS * 64);
  const uint32_t valid = (n - tileBase) < (size_t)TILE ? (uint32_t)(n - tileBase) : (uint32_t)TILE;

  KT k[ITEMS];
  uint32_t d[ITEMS];
  uint32_t pos[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const size_t idx = waveBase + (size_t)j * 64 + l;
    k[j] = idx < n ? src[idx] : ~(KT)0;
    d[j] = digit_of(k[j], shift, dmask);
  }
  // 32-bit payloads of 32-bit keys are loaded with the keys, so their loads overlap the ranking and
  // the look-back (the registers fit without lowering occupancy); wider ones after the key write.
  constexpr bool kEarlyVals = VB == 32 && sizeof(KT) == 4;
  VT ve[kEarlyVals ? ITEMS : 1];
  if constexpr (kEarlyVals) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const size_t idx = waveBase + (size_t)j * 64 + l;
      ve[j] = idx < n ? vsrc[idx] : (VT)0;
    }
  }

  // Wave multisplit: peers with the same digit, rank among them, wave-private running counts.
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint64_t m = match_digit<BITS>(d[j], ~0ull);
    const uint32_t r = mask_rank(m);
    const uint32_t c = (uint32_t)__popcll(m);
    const uint32_t pre = s_whist[w][d[j]];
    if (r == 0) s_whist[w][d[j]] = pre + c;
    pos[j] = pre + r;
  }
  __syncthreads();

  // Tile digit counts; wave offsets become exclusive; publish aggregates as early as possible.
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
      uint32_t *slot = &st.status[((size_t)pass * numTiles + tile) * R + dd];
      st_agent(slot, (tile == 0 ? kFlagInc : kFlagAgg) | sum);
    }
  }

  // Local exclusive scan of the tile's digit counts (thread-contiguous digits).
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

  // Per-digit look-back over preceding tiles.
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int dd = tid * DPT + q;
    if (dd >= R) continue;
    uint32_t excl = 0;
    if (tile > 0) {
      // Four predecessors per step (four loads in flight): aggregates are summed up to the first
      // inclusive value; at a tile that has not published yet the walk waits and resumes there.
      int64_t t = (int64_t)tile - 1;
      while (t >= 0) {
        const int nb = (int)(t + 1 < 4 ? t + 1 : 4);
        uint32_t sv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (i < nb) sv[i] = ld_agent(&st.status[((size_t)pass * numTiles + (t - i)) * R + dd]);
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
        t -= used;
        if (used < nb) __builtin_amdgcn_s_sleep(1);
      }
      st_agent(&st.status[((size_t)pass * numTiles + tile) * R + dd], kFlagInc | (excl + cnt[q]));
    }
    s_gbase[dd] = st.hist[pass * R + dd] + excl;
  }
  __syncthreads();
  // s_gbase[d] - s_tdigit[d] maps a local sorted position to its global position.
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const int dd = tid * DPT + q;
    if (dd < R) s_gbase[dd] -= s_tdigit[dd];
  }

  // Stable local reorder of the keys in LDS.
  KT *s_keys = (KT *)s_buf;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    pos[j] = s_tdigit[d[j]] + s_whist[w][d[j]] + pos[j];
    s_keys[pos[j]] = k[j];
  }
  __syncthreads();

  // The sorted keys back from LDS (every position of the tile holds a key, the padding included)
  // and their global positions: all LDS reads are issued before the first write, so the wave waits
  // once instead of twice per key.
  KT sk[ITEMS];
  uint32_t gaddr[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) sk[j] = s_keys[j * THREADS + tid];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) gaddr[j] = s_gbase[digit_of(sk[j], shift, dmask)] + (uint32_t)(j * THREADS + tid);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    // gaddr < n always holds for the tile's own keys when histogram and look-back agree; it guards
    // the write.
    if ((uint32_t)(j * THREADS + tid) < valid && gaddr[j] < n) dst[gaddr[j]] = sk[j];
  }

  if (VB) {
    __syncthreads();
    VT *s_vals = (VT *)s_buf;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if constexpr (kEarlyVals) {
        s_vals[pos[j]] = ve[j];
      } else {
        const size_t idx = waveBase + (size_t)j * 64 + l;
        s_vals[pos[j]] = idx < n ? vsrc[idx] : (VT)0;
      }
    }
    __syncthreads();
    VT sv[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) sv[j] = s_vals[j * THREADS + tid];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
      if ((uint32_t)(j * THREADS + tid) < valid && gaddr[j] < n) vdst[gaddr[j]] = sv[j];
  }
}

// Moves the result back into the caller's buffers when an odd number of passes ran.
template <typename KT, int VB>
__global__ void __launch_bounds__(256) k_final_copy(KT *__restrict__ keys0, const KT *__restrict__ keys1,
                                                     void *__restrict__ vals0, const void *__restrict__ vals1,
                                                     size_t n, const uint32_t *__restrict__ activeMask) {
  using VT = typename ValType<VB>::T;
  if (!(__builtin_popcount(*activeMask) & 1)) return;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    keys0[i] = keys1[i];
    if (VB) ((VT *)vals0)[i] = ((const VT *)vals1)[i];
  }
}

// ------------------------------------------------------------------------------------------------
// Host dispatch
// ------------------------------------------------------------------------------------------------
inline int passes_for(int loBit, int hiBit, int bits) { return (hiBit - loBit + bits - 1) / bits; }

// Keys per tile of one sort: 2,048 when asked for; 4,096 for 64-bit keys or payloads (their
// larger tiles do not fit the LDS); automatically 16,384 (1,024 threads x 16) for at least 2^22
// narrow keys and 8,192 below (fewer, larger tiles cost fewer look-back steps once every CU has
// tiles to rank); otherwise the configured size.
constexpr size_t kAutoLargeTileKeys = (size_t)1 << 22;

int tile_keys_effective(const SortConfig &cfg, size_t n, int keyBits, int valBits) {
  const bool wide = keyBits == 64 || valBits == 64;
  if (cfg.tileKeys == 2048) return 2048;
  if (wide) return 4096;  // 256 threads x 16 keys of 64 bits fit the LDS; 4,096 to 16,384 asked: 4,096
  if (cfg.tileKeys == 0) return n >= kAutoLargeTileKeys ? 16384 : 8192;
  return cfg.tileKeys;
}

template <typename KT, int VB, int BITS, int THREADS, int ITEMS>
void run_typed(Context *ctx, hipStream_t s, void *keys, void *vals, void *keysAlt, void *valsAlt, size_t n, int loBit,
               int hiBit, void *scratch) {
  constexpr int R = 1 << BITS;
  constexpr int TILE = THREADS * ITEMS;
  const int passes = passes_for(loBit, hiBit, BITS);
  const uint32_t tiles = (uint32_t)((n + TILE - 1) / TILE);
  SortState st = state_view(scratch, passes, R);
  const size_t words = (size_t)passes * R + 32 + (size_t)passes * tiles * R;
  fill_u32(ctx, s, scratch, 0u, words);
  size_t hb = (n + 256 * 16 - 1) / (256 * 16);
  const size_t hbMax = (size_t)ctx->cuCount * 4;
  if (hb > hbMax) hb = hbMax;
  if (hb < 1) hb = 1;
  launch(ctx, s, "sort_hist", k_hist<KT, BITS>, dim3((unsigned)hb), dim3(256), (size_t)passes * R * 4,
         (const KT *)keys, n, loBit, hiBit, passes, st.hist);
  launch(ctx, s, "sort_hist_scan", k_hist_scan<BITS>, dim3(1), dim3(256), 0, st.hist, st.activeMask, passes,
         (uint32_t)n);
  for (int p = 0; p < passes; ++p) {
    launch(ctx, s, "sort_onesweep", k_onesweep<KT, VB, BITS, THREADS, ITEMS>, dim3(tiles), dim3(THREADS), 0,
           (KT *)keys, (KT *)keysAlt, vals, valsAlt, (uint32_t)n, p, loBit + p * BITS, hiBit, scratch, passes,
           tiles);
  }
  size_t cb = (n + 255) / 256;
  if (cb > 8192) cb = 8192;
  launch(ctx, s, "sort_final_copy", k_final_copy<KT, VB>, dim3((unsigned)cb), dim3(256), 0, (KT *)keys,
         (const KT *)keysAlt, vals, (const void *)valsAlt, n, (const uint32_t *)st.activeMask);
}

template <typename KT, int VB, int BITS>
bool run_bits(Context *ctx, hipStream_t s, void *keys, void *vals, void *keysAlt, void *valsAlt, size_t n, int loBit,
              int hiBit, void *scratch, const SortConfig &cfg) {
  constexpr bool wide = sizeof(KT) == 8 || VB == 64;
  switch (tile_keys_effective(cfg, n, (int)sizeof(KT) * 8, VB)) {
    case 2048: run_typed<KT, VB, BITS, 256, 8>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch); return true;
    case 4096: run_typed<KT, VB, BITS, 256, 16>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch); return true;
    case 8192:
      if constexpr (!wide) {
        run_typed<KT, VB, BITS, 512, 16>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch);
        return true;
      }
      break;
    case 16384:
      if constexpr (!wide) {
        run_typed<KT, VB, BITS, 1024, 16>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch);
        return true;
      }
      break;
  }
  fail("sort: tileKeys %d is not one of 0, 2048, 4096, 8192, 16384", cfg.tileKeys);
  return false;
}

template <typename KT, int VB>
bool run_kv(Context *ctx, hipStream_t s, void *keys, void *vals, void *keysAlt, void *valsAlt, size_t n, int loBit,
            int hiBit, void *scratch, const SortConfig &cfg) {
  switch (cfg.digitBits) {
    case 4: return run_bits<KT, VB, 4>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
    case 6: return run_bits<KT, VB, 6>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
    case 8: return run_bits<KT, VB, 8>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
    default: fail("sort: digitBits %d is not one of 4, 6, 8", cfg.digitBits); return false;
  }
}

}  // namespace

size_t sort_scratch_bytes(size_t n, int keyBits, int valBits, int loBit, int hiBit, const SortConfig &cfg) {
  const int R = 1 << cfg.digitBits;
  const int passes = passes_for(loBit, hiBit, cfg.digitBits);
  const size_t tile = (size_t)tile_keys_effective(cfg, n, keyBits, valBits);
  const size_t tiles = (n + tile - 1) / tile;
  return ((size_t)passes * R + 32 + (size_t)passes * tiles * R) * 4;
}

bool sort_run(Context *ctx, hipStream_t s, void *keys, void *vals, void *keysAlt, void *valsAlt, size_t n, int keyBits,
              int valBits, int loBit, int hiBit, void *scratch, const SortConfig &cfg) {
  if (n <= 1) return true;
  if (loBit < 0 || hiBit > keyBits || loBit >= hiBit) {
    fail("sort: bit range [%d, %d) invalid for %d-bit keys", loBit, hiBit, keyBits);
    return false;
  }
  if (keyBits == 32) {
    if (valBits == 0) return run_kv<uint32_t, 0>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
    if (valBits == 32) return run_kv<uint32_t, 32>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
    return run_kv<uint32_t, 64>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
  }
  if (valBits == 0) return run_kv<uint64_t, 0>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
  if (valBits == 32) return run_kv<uint64_t, 32>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
  return run_kv<uint64_t, 64>(ctx, s, keys, vals, keysAlt, valsAlt, n, loBit, hiBit, scratch, cfg);
}

}  // namespace gprt
