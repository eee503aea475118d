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
//    with 16-byte coalesced vector loads and scanned with DPP wave prefix sums. The sum streams
//    its input and output with nontemporal loads and stores.
//  * The sum's tiles are large (16 waves x 44 items per lane = 45,056 items, one workgroup per
//    CU): a tile waits in its look-back for predecessors that are still loading, and that wait
//    grows with the number of tiles in flight, so fewer, larger tiles spend less time waiting.
//  * The look-back is done by a whole wave: lane l inspects predecessor tile (p-1-l), so 64
//    predecessors are examined per round trip instead of one.
//  * Status words are 64-bit {flag:32 | value:32}: values wrap modulo 2^32 with no 2^30 limit
//    (the reference packs value<<2|flag into 32 bits, gprt/gprt_scan.slang:223-228).
//  * State layout (uint64 words): [0] total (low 32 bits), [1] tile ticket, [2 + t] tile t status.
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"

namespace gprt {

namespace {

constexpr uint64_t kAgg = 1ull << 32;
constexpr uint64_t kInc = 2ull << 32;

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// MODE 0: exclusive sum; 1: partition; 2: select.
template <int WAVES, int K, int MODE>
__global__ void __launch_bounds__(WAVES * 64)
    k_scan(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint32_t n, int selectPositive,
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

  // Only the items stay in registers; what an item counts (itself for the sum, its selection flag
  // for partition / select) is recomputed from it, and its prefix from its 4-item group's exclusive
  // offset, when the results are written.
  uint32_t v[K][4];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const size_t idx = waveBase + ((size_t)k * 64 + l) * 4;
    if (idx + 3 < n) {
      const u32x4_t q = MODE == 0 ? __builtin_nontemporal_load((const u32x4_t *)(in + idx)) : *(const u32x4_t *)(in + idx);
      v[k][0] = q.x;
      v[k][1] = q.y;
      v[k][2] = q.z;
      v[k][3] = q.w;
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) v[k][c] = (idx + c < n) ? in[idx + c] : 0u;
    }
  }
  auto counted = [&](int k, int c) -> uint32_t {
    if (MODE == 0) return v[k][c];
    const size_t idx = waveBase + ((size_t)k * 64 + l) * 4 + c;
    const bool sel = selectPositive ? ((int32_t)v[k][c] >= 0) : ((int32_t)v[k][c] < 0);
    return (idx < n && sel) ? 1u : 0u;
  };

  // Exclusive prefix of each lane's 4-item group within the wave's chunk (order: k, lane, item).
  uint32_t gex[K];
  uint32_t carry = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t s = counted(k, 0) + counted(k, 1) + counted(k, 2) + counted(k, 3);
    const uint32_t incl = wave_incl_sum(s);
    gex[k] = incl - s + carry;
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
    uint32_t r[4];
    r[0] = prefix + gex[k];
#pragma unroll
    for (int c = 1; c < 4; ++c) r[c] = r[c - 1] + counted(k, c - 1);
    if (MODE == 0) {
      if (idx + 3 < n) {
        __builtin_nontemporal_store(u32x4_t{r[0], r[1], r[2], r[3]}, (u32x4_t *)(out + idx));
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (idx + c < n) out[idx + c] = r[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const size_t i = idx + c;
        if (i >= n) continue;
        const uint32_t rank = r[c];
        if (counted(k, c)) {
          out[rank] = v[k][c];
        } else if (MODE == 1) {
          out[(n - 1) - ((uint32_t)i - rank)] = v[k][c];
        }
      }
    }
  }
}

// Tiles: 45,056 items for the sum (16 waves x 44 items per lane, one workgroup per CU; 113
// VGPRs, within the 128 that 16 waves on a CU allow), 4,096 for partition / select (4 waves x 16
// items per lane), whose flags and scattered writes need more registers.
constexpr int kScanWavesSum = 16, kScanWavesSel = 4;
constexpr int kScanKSum = 11, kScanKSel = 4;
constexpr uint32_t kScanTileSum = kScanWavesSum * 64 * kScanKSum * 4;
constexpr uint32_t kScanTileSel = kScanWavesSel * 64 * kScanKSel * 4;

// The look-back state of gprtBufferExclusiveSumNoAlloc: one device-resident array per GPU that
// the library runs on, held until the process ends, so that the call allocates nothing. It holds
// the state of one exclusive sum of up to kNoAllocMaxItems items: [0] total, [1] tile ticket,
// [2 + t] the status of tile t. Host code never passes `scratch` itself to a kernel (in host code
// the name stands for a host-side shadow); it uses the device address hipGetSymbolAddress returns
// for the current device, kept per device in g_noAllocState.
constexpr size_t kNoAllocMaxItems = 1ull << 31;
constexpr size_t kNoAllocWords = (1ull << 17) + 2;  // 1 MiB + 16 B
static_assert(kNoAllocWords >= (kNoAllocMaxItems + kScanTileSum - 1) / kScanTileSum + 2,
              "the static state holds every tile status of the largest allowed input");

std::mutex g_noAllocMutex;
std::vector<void *> g_noAllocState;  // device address of `scratch` per device ordinal

}  // namespace

static __device__ __align__(256) uint64_t scratch[kNoAllocWords];

bool scan_noalloc_init(int device) {
  std::lock_guard<std::mutex> lock(g_noAllocMutex);
  if (device < 0) return false;
  if ((size_t)device < g_noAllocState.size() && g_noAllocState[device]) return true;
  void *p = nullptr;
  if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scan state)") || !p) return false;
  if (g_noAllocState.size() <= (size_t)device) g_noAllocState.resize((size_t)device + 1, nullptr);
  g_noAllocState[device] = p;
  return true;
}

void *scan_noalloc_state(int device) {
  std::lock_guard<std::mutex> lock(g_noAllocMutex);
  return (device >= 0 && (size_t)device < g_noAllocState.size()) ? g_noAllocState[device] : nullptr;
}

size_t scan_noalloc_max_items() { return kNoAllocMaxItems; }

size_t scan_scratch_bytes(size_t n) {
  size_t tiles = (n + kScanTileSel - 1) / kScanTileSel;  // the smaller tile: enough for either
  return (tiles + 2) * sizeof(uint64_t);
}

void scan_run(Context *ctx, hipStream_t s, const uint32_t *in, uint32_t *out, size_t n, int mode, bool selectPositive,
              void *state, uint32_t *totalDev) {
  (void)totalDev;  // state[0] is the total; the caller reads it from the state's base
  const uint32_t tileItems = mode == 0 ? kScanTileSum : kScanTileSel;
  const uint32_t tiles = (uint32_t)((n + tileItems - 1) / tileItems);
  fill_u32(ctx, s, state, 0u, ((size_t)tiles + 2) * 2);
  const int sel = selectPositive ? 1 : 0;
  if (mode == 0)
    launch(ctx, s, "scan_lookback", k_scan<kScanWavesSum, kScanKSum, 0>, dim3(tiles), dim3(kScanWavesSum * 64), 0, in,
           out, (uint32_t)n, sel, (uint64_t *)state, tiles);
  else if (mode == 1)
    launch(ctx, s, "scan_lookback", k_scan<kScanWavesSel, kScanKSel, 1>, dim3(tiles), dim3(kScanWavesSel * 64), 0, in,
           out, (uint32_t)n, sel, (uint64_t *)state, tiles);
  else
    launch(ctx, s, "scan_lookback", k_scan<kScanWavesSel, kScanKSel, 2>, dim3(tiles), dim3(kScanWavesSel * 64), 0, in,
           out, (uint32_t)n, sel, (uint64_t *)state, tiles);
}

}  // namespace gprt
