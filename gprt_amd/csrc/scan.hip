// Single-pass exclusive scan / partition / select over 32-bit items for gfx950.
//
// Behaviour follows GPRT's scan family (gprt/gprt_scan.slang:84-392; host side
// gprt/gprt.cpp:8739-8818, commented-out path): an atomic ticket assigns tiles in launch order so
// a tile only ever waits on tiles that are already resident (gprt/gprt_scan.slang:112-115), and
// each tile publishes an aggregate, looks back over its predecessors, then publishes its inclusive
// prefix. Partition writes selected items at their exclusive rank and unselected items at
// (n-1) - (i - rank) (gprt/gprt_scan.slang:89-92); select writes only the selected items.
//
// MI355X design:
//  * A tile is WAVES wave64s; each wave owns 64 lanes x K uint4 = 256K consecutive items, loaded
//    with 16-byte coalesced vector loads and scanned with wave-level prefix sums.
//  * The look-back is done by a whole wave: lane l inspects predecessor tile (p-1-l), so 64
//    predecessors are examined per round trip instead of one.
//  * Status words are 64-bit {flag:32 | value:32}: values wrap modulo 2^32 with no 2^30 limit
//    (the reference packs value<<2|flag into 32 bits, gprt/gprt_scan.slang:223-228).
//  * State layout (uint64 words): [0] total (low 32 bits), [1] tile ticket, [2 + t] tile t status.
#include "device_common.hpp"
#include "internal.hpp"

namespace gprt {

namespace {

constexpr uint64_t kAgg = 1ull << 32;
constexpr uint64_t kInc = 2ull << 32;

template <int WAVES, int K>
__global__ void __launch_bounds__(WAVES * 64)
    k_scan(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint32_t n, int mode, int selectPositive,
           uint64_t *__restrict__ state, uint32_t numTiles) {
  constexpr uint32_t WAVE_ITEMS = 64 * K * 4;
  constexpr uint32_t TILE = WAVES * WAVE_ITEMS;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wsum[WAVES];
  __shared__ uint32_t s_prefix;

  const uint32_t tid = threadIdx.x;
  const uint32_t w = tid >> 6;
  const uint32_t l = tid & 63;

  if (tid == 0) s_tile = ticket((uint32_t *)&state[1]);
  __syncthreads();
  const uint32_t tile = s_tile;
  const size_t waveBase = (size_t)tile * TILE + (size_t)w * WAVE_ITEMS;

  uint32_t v[K][4];
  uint32_t f[K][4];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const size_t idx = waveBase + ((size_t)k * 64 + l) * 4;
    if (idx + 3 < n) {
      uint4 q = *(const uint4 *)(in + idx);
      v[k][0] = q.x;
      v[k][1] = q.y;
      v[k][2] = q.z;
      v[k][3] = q.w;
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) v[k][c] = (idx + c < n) ? in[idx + c] : 0u;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (mode == 0) {
        f[k][c] = v[k][c];
      } else {
        const bool inRange = idx + c < n;
        const bool sel = selectPositive ? ((int32_t)v[k][c] >= 0) : ((int32_t)v[k][c] < 0);
        f[k][c] = (inRange && sel) ? 1u : 0u;
      }
    }
  }

  // Exclusive prefix of every item within the wave's chunk (order: k, lane, component).
  uint32_t pre[K][4];
  uint32_t carry = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t s = f[k][0] + f[k][1] + f[k][2] + f[k][3];
    const uint32_t incl = wave_incl_sum(s);
    const uint32_t excl = incl - s + carry;
    pre[k][0] = excl;
    pre[k][1] = excl + f[k][0];
    pre[k][2] = pre[k][1] + f[k][1];
    pre[k][3] = pre[k][2] + f[k][2];
    carry += wave_read(incl, 63);
  }
  if (l == 0) s_wsum[w] = carry;
  __syncthreads();

  uint32_t woff = 0, agg = 0;
#pragma unroll
  for (int i = 0; i < WAVES; ++i) {
    const uint32_t x = s_wsum[i];
    if ((uint32_t)i < w) woff += x;
    agg += x;
  }

  if (w == 0) {
    uint32_t excl = 0;
    if (tile == 0) {
      if (l == 0) st_agent(&state[2], kInc | agg);
    } else {
      if (l == 0) st_agent(&state[2 + tile], kAgg | agg);
      int64_t base = (int64_t)tile - 1;
      uint32_t acc = 0;
      while (true) {
        const int64_t t = base - (int64_t)l;
        const uint64_t word = (t >= 0) ? ld_agent(&state[2 + t]) : kInc;
        const uint32_t flag = (uint32_t)(word >> 32);
        const uint64_t incMask = wave_ballot(flag == 2u);
        const uint64_t readyMask = wave_ballot(flag != 0u);
        if (incMask) {
          const int first = __builtin_ctzll(incMask);
          const uint64_t need = (first == 63) ? ~0ull : ((1ull << (first + 1)) - 1ull);
          if ((readyMask & need) == need) {
            acc += wave_sum((int)l <= first ? (uint32_t)word : 0u);
            break;
          }
        } else if (readyMask == ~0ull) {
          acc += wave_sum((uint32_t)word);
          base -= 64;
          continue;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      excl = acc;
      if (l == 0) st_agent(&state[2 + tile], kInc | (uint32_t)(excl + agg));
    }
    if (l == 0) {
      s_prefix = excl;
      if (tile == numTiles - 1) st_agent(&state[0], (uint64_t)(uint32_t)(excl + agg));
    }
  }
  __syncthreads();
  const uint32_t prefix = s_prefix + woff;

#pragma unroll
  for (int k = 0; k < K; ++k) {
    const size_t idx = waveBase + ((size_t)k * 64 + l) * 4;
    if (mode == 0) {
      if (idx + 3 < n) {
        *(uint4 *)(out + idx) = make_uint4(prefix + pre[k][0], prefix + pre[k][1], prefix + pre[k][2],
                                           prefix + pre[k][3]);
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (idx + c < n) out[idx + c] = prefix + pre[k][c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const size_t i = idx + c;
        if (i >= n) continue;
        const uint32_t rank = prefix + pre[k][c];
        if (f[k][c]) {
          out[rank] = v[k][c];
        } else if (mode == 1) {
          out[(n - 1) - ((uint32_t)i - rank)] = v[k][c];
        }
      }
    }
  }
}

constexpr int kScanWaves = 4;
constexpr int kScanK = 4;
constexpr uint32_t kScanTile = kScanWaves * 64 * kScanK * 4;

}  // namespace

size_t scan_scratch_bytes(size_t n) {
  size_t tiles = (n + kScanTile - 1) / kScanTile;
  return (tiles + 2) * sizeof(uint64_t);
}

void scan_run(Context *ctx, hipStream_t s, const uint32_t *in, uint32_t *out, size_t n, int mode, bool selectPositive,
              void *scratch, uint32_t *totalDev) {
  (void)totalDev;  // state[0] is the total; the caller reads it from the scratch base
  const uint32_t tiles = (uint32_t)((n + kScanTile - 1) / kScanTile);
  fill_u32(ctx, s, scratch, 0u, ((size_t)tiles + 2) * 2);
  launch(ctx, s, "scan_lookback", k_scan<kScanWaves, kScanK>, dim3(tiles), dim3(kScanWaves * 64), 0, in, out,
         (uint32_t)n, mode, selectPositive ? 1 : 0, (uint64_t *)scratch, tiles);
}

}  // namespace gprt
