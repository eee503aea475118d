// Host-side internals shared by the translation units of libgprt_hip.so.
// Context / Buffer / Geom / Accel objects behind the opaque GPRT handles, the error path, the
// kernel launch wrapper with optional per-kernel HIP-event timing, and the declarations of the
// per-stage launchers (scan, sort, LBVH, trace).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/gprt_hip.h"

namespace gprt {

// ------------------------------------------------------------------------------------------------
// Errors (gprt/gprt.cpp:205-251: print "#gprt error:" then raise(SIGINT)).
// ------------------------------------------------------------------------------------------------
void fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
bool hip_ok(hipError_t e, const char *what);

// ------------------------------------------------------------------------------------------------
// Context
// ------------------------------------------------------------------------------------------------
struct KernelTimer {
  bool enabled = false;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, uint32_t>> totals;
  hipEvent_t get();
  void drain();  // host wait + accumulate pending pairs
};

// Radix sort tuning knobs (the autotuner sweeps them).
struct SortConfig {
  int digitBits = 8;     // 4, 6 or 8
  int tileKeys = 4096;   // keys per tile: 256 threads x {8, 16} (x2 for 512 threads)
  int threads = 256;     // 256 or 512
};

struct Context {
  SortConfig sortCfg;  // used by every sort and LBVH build issued on this context
  int device = 0;
  int cuCount = 256;
  hipStream_t stream = nullptr;
  hipEvent_t profBegin = nullptr, profEnd = nullptr;
  bool profiling = false;
  KernelTimer timer;
  // Internal arena used when a caller passes a null scratch buffer.
  void *arena = nullptr;
  size_t arenaBytes = 0;
  void *ensure_arena(size_t bytes);
  // Trace work counter (first 256 B) and per-lane stack spill area.
  void *traceScratch = nullptr;
  size_t traceBytes = 0;
  // Small pinned host word for scan totals.
  uint32_t *pinnedWord = nullptr;
};

// ------------------------------------------------------------------------------------------------
// Buffers
// ------------------------------------------------------------------------------------------------
enum class BufKind { Host, Device, Shared };

struct Buffer {
  Context *ctx = nullptr;
  BufKind kind = BufKind::Device;
  size_t size = 0;    // bytes
  size_t align = 16;  // requested alignment (hipMalloc gives >= 256)
  void *d = nullptr;  // device address
  void *h = nullptr;  // host pointer: the pinned allocation (Host) or a staging copy (Device/Shared)
  bool mapped = false;
};

bool buffer_realloc(Buffer *b, size_t bytes, bool preserve);

// ------------------------------------------------------------------------------------------------
// Geometry and accel
// ------------------------------------------------------------------------------------------------
struct GeomType {
  Context *ctx;
  GPRTGeomKind kind;
  size_t recordSize;
};

struct Geom {
  Context *ctx;
  GeomType *type;
  Buffer *vbuf = nullptr;
  uint32_t vcount = 0, vstride = 12, voffset = 0;
  Buffer *ibuf = nullptr;
  uint32_t icount = 0, istride = 12, ioffset = 0;
};

// Device storage of one triangle LBVH (layout documented in bvh.hpp).
struct LBVHStorage {
  uint32_t n = 0;          // primitives the storage is sized for
  uint32_t capacity = 0;   // primitives allocated
  void *block = nullptr;   // one hipMalloc holding every array below
  size_t blockBytes = 0;
  uint32_t *mcodes = nullptr;     // [n] sorted morton codes
  uint32_t *ids = nullptr;        // [n] primIDs in sorted order
  void *nodes = nullptr;          // [n-1] BVHNode (64 B)
  void *tris = nullptr;           // [n] TriRecord (48 B), leaf order
  uint32_t *parentLeaf = nullptr; // [n]  parent<<1 | side
  uint32_t *parentInner = nullptr;// [n-1] parent<<1 | side (root: 0xFFFFFFFF)
  uint32_t *counters = nullptr;   // [n-1] refit arrival counters
  float *header = nullptr;        // [16] scene box (0..5), root box (8..13)
  uint32_t *keysTmp = nullptr;    // [n] sort ping-pong keys
  uint32_t *idsTmp = nullptr;     // [n] sort ping-pong ids
  void *sortScratch = nullptr;    // sort histograms / look-back state
  size_t sortScratchBytes = 0;
};

struct Accel {
  Context *ctx = nullptr;
  Geom *geom = nullptr;
  unsigned flags = 0;
  GPRTBuildMode mode = GPRT_BUILD_MODE_UNINITIALIZED;
  bool built = false;
  LBVHStorage bvh;
};

// ------------------------------------------------------------------------------------------------
// Kernel launch with optional per-kernel timing on the launch stream.
// ------------------------------------------------------------------------------------------------
template <typename K, typename... Args>
inline void launch(Context *ctx, hipStream_t s, const char *name, K kernel, dim3 grid, dim3 block, size_t shmem,
                   Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  const bool timed = ctx && ctx->timer.enabled;
  if (timed) {
    a = ctx->timer.get();
    b = ctx->timer.get();
    hipEventRecord(a, s);
  }
  hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
  if (timed) {
    hipEventRecord(b, s);
    ctx->timer.pending.push_back({name, a, b});
  }
}

// ------------------------------------------------------------------------------------------------
// Stage launchers
// ------------------------------------------------------------------------------------------------
// Scan family: mode 0 = exclusive sum, 1 = partition, 2 = select.
size_t scan_scratch_bytes(size_t n);
void scan_run(Context *ctx, hipStream_t s, const uint32_t *in, uint32_t *out, size_t n, int mode, bool selectPositive,
              void *scratch, uint32_t *totalDev);

size_t sort_scratch_bytes(size_t n, int keyBits, int valBits, int loBit, int hiBit, const SortConfig &cfg);
// Sorts n keys (keyBits 32|64) with optional payload (valBits 0|32|64) by key bits [loBit, hiBit).
// keysAlt/valsAlt are ping-pong buffers of the same size; the result lands in keys/vals.
bool sort_run(Context *ctx, hipStream_t s, void *keys, void *vals, void *keysAlt, void *valsAlt, size_t n, int keyBits,
              int valBits, int loBit, int hiBit, void *scratch, const SortConfig &cfg);

// Fill a device range with a 32-bit pattern (graph-safe; no hipMemset nodes).
void fill_u32(Context *ctx, hipStream_t s, void *dst, uint32_t value, size_t words);

// LBVH
bool lbvh_reserve(Context *ctx, LBVHStorage &st, uint32_t n);
void lbvh_release(LBVHStorage &st);
bool lbvh_build(Context *ctx, hipStream_t s, Accel *a);
bool lbvh_refit(Context *ctx, hipStream_t s, Accel *a);
bool lbvh_export(Context *ctx, hipStream_t s, Accel *a, int32_t *nodes, float *aabbs, uint64_t *keys);

// Trace
bool trace_closest(Context *ctx, hipStream_t s, Accel *a, const void *rays, void *hits, uint32_t count);

}  // namespace gprt
