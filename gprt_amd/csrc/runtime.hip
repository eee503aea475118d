// Runtime of libgprt_hip.so: context, buffers, errors, profiling and the C-ABI entry points of
// the buffer / scan / sort families. Replaces the Vulkan runtime core of GPRT
// (gprt/gprt.cpp:451-710 Context, :814-1196 Buffer, :6672-6840 queues) with one HIP device, one
// in-order stream and hipMalloc'd HBM.
#include <rocprofiler-sdk-roctx/roctx.h>

#include <chrono>
#include <csignal>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <thread>

#include "device_common.hpp"
#include "internal.hpp"

namespace gprt {

// ------------------------------------------------------------------------------------------------
// Errors
// ------------------------------------------------------------------------------------------------
static int g_errorMode = GPRT_ERROR_RAISE;
static thread_local std::string g_lastError;
static thread_local bool g_hasError = false;

void fail(const char *fmt, ...) {
  char msg[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof(msg), fmt, ap);
  va_end(ap);
  fprintf(stdout, "\033[0;31m#gprt error: %s\033[0m\n", msg);
  fflush(stdout);
  g_lastError = msg;
  g_hasError = true;
  if (g_errorMode == GPRT_ERROR_RAISE) raise(SIGINT);
}

bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  fail("%s failed: %s", what, hipGetErrorString(e));
  return false;
}

// CRC-32 with the reflected IEEE 802.3 polynomial 0xEDB88320, initial value and final xor ~0:
// the checksum of zlib's crc32 and of PNG / gzip (table-driven, one byte at a time).
bool stream_wait(Context *ctx, hipStream_t s, const char *what) {
  const double limit = ctx ? ctx->syncTimeout : 0.0;
  if (limit <= 0.0) return hip_ok(hipStreamSynchronize(s), what);
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t poll = 0;; ++poll) {
    const hipError_t e = hipStreamQuery(s);
    if (e == hipSuccess) return true;
    if (e != hipErrorNotReady) return hip_ok(e, what);
    const double waited = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (waited > limit) {
      fail("%s: the GPU did not finish the context's work within %.3g s (sync timeout, gprtSetSyncTimeout)", what,
           limit);
      return false;
    }
    // Short waits stay responsive; long ones stop competing for the CPU.
    if (poll < 4096)
      std::this_thread::yield();
    else
      std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

StageRange::StageRange(const char *name) { roctxRangePushA(name); }
StageRange::~StageRange() { roctxRangePop(); }

struct Crc32Table {
  uint32_t t[256];
  Crc32Table() {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      t[i] = c;
    }
  }
};

uint32_t crc32_update(uint32_t crc, const uint8_t *p, size_t n) {
  static const Crc32Table tab;  // initialised once, thread-safe
  const uint32_t *table = tab.t;
  crc = ~crc;
  for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
  return ~crc;
}


// ------------------------------------------------------------------------------------------------
// Kernel timer
// ------------------------------------------------------------------------------------------------
hipEvent_t KernelTimer::get() {
  if (pool.empty()) {
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
  }
  hipEvent_t e = pool.back();
  pool.pop_back();
  return e;
}

void KernelTimer::drain() {
  for (auto &p : pending) {
    hipEventSynchronize(p.b);
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.a, p.b);
    auto &t = totals[p.name];
    t.first += ms;
    t.second += 1;
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  pending.clear();
}

void *Context::ensure_arena(size_t bytes) {
  if (bytes <= arenaBytes) return arena;
  if (arena) {
    hipStreamSynchronize(stream);
    hipFree(arena);
    arena = nullptr;
    arenaBytes = 0;
  }
  size_t want = bytes + bytes / 8;
  if (!hip_ok(hipMalloc(&arena, want), "hipMalloc(arena)")) return nullptr;
  arenaBytes = want;
  return arena;
}

// ------------------------------------------------------------------------------------------------
// Fill kernel (graph-safe replacement for hipMemsetAsync)
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *__restrict__ dst, uint32_t value, size_t words) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < words; i += stride) dst[i] = value;
}

void fill_u32(Context *ctx, hipStream_t s, void *dst, uint32_t value, size_t words) {
  if (words == 0) return;
  size_t blocks = (words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  launch(ctx, s, "fill_u32", k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, (uint32_t *)dst, value, words);
}

// ------------------------------------------------------------------------------------------------
// Buffers
// ------------------------------------------------------------------------------------------------
static void buffer_free(Buffer *b) {
  if (b->kind == BufKind::Host) {
    if (b->h) hipHostFree(b->h);
  } else {
    if (b->d) hipFree(b->d);
    if (b->h) hipHostFree(b->h);
  }
  b->d = b->h = nullptr;
}

static bool buffer_alloc(Buffer *b, size_t bytes) {
  b->size = bytes;
  if (bytes == 0) return true;
  if (b->kind == BufKind::Host) {
    if (!hip_ok(hipHostMalloc(&b->h, bytes, hipHostMallocMapped), "hipHostMalloc")) return false;
    if (!hip_ok(hipHostGetDevicePointer(&b->d, b->h, 0), "hipHostGetDevicePointer")) return false;
  } else {
    if (!hip_ok(hipMalloc(&b->d, bytes), "hipMalloc")) return false;
  }
  return true;
}

bool buffer_realloc(Buffer *b, size_t bytes, bool preserve) {
  if (bytes == b->size) return true;
  Buffer nb = *b;
  nb.d = nb.h = nullptr;
  nb.mapped = false;
  if (!buffer_alloc(&nb, bytes)) return false;
  hipStreamSynchronize(b->ctx->stream);
  if (preserve && b->size && bytes) {
    size_t n = b->size < bytes ? b->size : bytes;
    if (b->kind == BufKind::Host)
      memcpy(nb.h, b->h, n);
    else if (!hip_ok(hipMemcpy(nb.d, b->d, n, hipMemcpyDeviceToDevice), "hipMemcpy(resize)"))
      return false;
  }
  buffer_free(b);
  b->d = nb.d;
  b->h = nb.h;
  b->size = bytes;
  b->mapped = false;
  return true;
}

static GPRTBuffer buffer_create(GPRTContext _ctx, BufKind kind, size_t size, size_t count, const void *init,
                                size_t alignment) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) {
    fail("buffer create: null context");
    return nullptr;
  }
  if (alignment == 0 || (alignment & (alignment - 1))) {
    fail("buffer create: alignment %zu is not a power of two", alignment);
    return nullptr;
  }
  Buffer *b = new Buffer();
  b->ctx = ctx;
  b->kind = kind;
  b->align = alignment;
  if (!buffer_alloc(b, size * count)) {
    delete b;
    return nullptr;
  }
  if (init && b->size) {
    if (kind == BufKind::Host)
      memcpy(b->h, init, b->size);
    else if (!hip_ok(hipMemcpy(b->d, init, b->size, hipMemcpyHostToDevice), "hipMemcpy(init)")) {
      buffer_free(b);
      delete b;
      return nullptr;
    }
  }
  return (GPRTBuffer)b;
}

}  // namespace gprt

using namespace gprt;

// ================================================================================================
// C-ABI
// ================================================================================================
GPRT_API const char *gprtVersion(void) { return "gprt-mi355x 0.1 (gfx950)"; }

GPRT_API void gprtSetErrorMode(int mode) { g_errorMode = mode; }
GPRT_API const char *gprtGetLastError(void) { return g_hasError ? g_lastError.c_str() : nullptr; }
GPRT_API void gprtClearLastError(void) {
  g_hasError = false;
  g_lastError.clear();
}

GPRT_API GPRTContext gprtContextCreate(int32_t *requestedDeviceIDs, int numDevices) {
  if (numDevices != 1) {
    fail("gprtContextCreate: numDevices=%d; this build drives exactly one GPU per context", numDevices);
    return nullptr;
  }
  int count = 0;
  if (!hip_ok(hipGetDeviceCount(&count), "hipGetDeviceCount")) return nullptr;
  int dev = requestedDeviceIDs ? requestedDeviceIDs[0] : 0;
  if (dev < 0 || dev >= count) {
    fail("gprtContextCreate: device %d out of range (%d devices)", dev, count);
    return nullptr;
  }
  if (!hip_ok(hipSetDevice(dev), "hipSetDevice")) return nullptr;
  if (!scan_noalloc_init(dev)) return nullptr;
  Context *ctx = new Context();
  ctx->device = dev;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ctx->cuCount = prop.multiProcessorCount;
  if (!hip_ok(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate")) {
    delete ctx;
    return nullptr;
  }
  hipEventCreate(&ctx->profBegin);
  hipEventCreate(&ctx->profEnd);
  hipHostMalloc((void **)&ctx->pinnedWord, 64, hipHostMallocDefault);
  return (GPRTContext)ctx;
}

hipStream_t Context::aux_stream() {
  if (!aux && !hip_ok(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking), "hipStreamCreate(aux)")) aux = nullptr;
  return aux;
}

hipEvent_t Context::order_event(size_t i) {
  while (orderEvents.size() <= i) {
    hipEvent_t e = nullptr;
    if (!hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate(order)")) return nullptr;
    orderEvents.push_back(e);
  }
  return orderEvents[i];
}

GPRT_API void gprtContextDestroy(GPRTContext _ctx) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) return;
  hipStreamSynchronize(ctx->stream);
  if (ctx->aux) hipStreamSynchronize(ctx->aux);
  ctx->timer.drain();
  for (auto e : ctx->timer.pool) hipEventDestroy(e);
  if (ctx->arena) hipFree(ctx->arena);
  if (ctx->traceScratch) hipFree(ctx->traceScratch);
  if (ctx->blasPool) hipFree(ctx->blasPool);
  if (ctx->pinnedWord) hipHostFree(ctx->pinnedWord);
  hipEventDestroy(ctx->profBegin);
  hipEventDestroy(ctx->profEnd);
  for (hipEvent_t e : ctx->orderEvents) hipEventDestroy(e);
  if (ctx->aux) hipStreamDestroy(ctx->aux);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

GPRT_API GPRTBuffer gprtHostBufferCreate(GPRTContext c, size_t size, size_t count, const void *init, size_t al) {
  return buffer_create(c, BufKind::Host, size, count, init, al);
}
GPRT_API GPRTBuffer gprtDeviceBufferCreate(GPRTContext c, size_t size, size_t count, const void *init, size_t al) {
  return buffer_create(c, BufKind::Device, size, count, init, al);
}
GPRT_API GPRTBuffer gprtSharedBufferCreate(GPRTContext c, size_t size, size_t count, const void *init, size_t al) {
  return buffer_create(c, BufKind::Shared, size, count, init, al);
}

GPRT_API void gprtBufferClear(GPRTBuffer _b) {
  Buffer *b = (Buffer *)_b;
  if (!b) {
    fail("gprtBufferClear: null buffer");
    return;
  }
  if (b->kind == BufKind::Host) {
    hipStreamSynchronize(b->ctx->stream);
    memset(b->h, 0, b->size);
    return;
  }
  size_t words = b->size / 4;
  fill_u32(b->ctx, b->ctx->stream, b->d, 0u, words);
  if (b->size & 3) hipMemsetAsync((char *)b->d + words * 4, 0, b->size & 3, b->ctx->stream);
}

GPRT_API void gprtBufferDestroy(GPRTBuffer _b) {
  Buffer *b = (Buffer *)_b;
  if (!b) return;
  hipStreamSynchronize(b->ctx->stream);
  buffer_free(b);
  delete b;
}

GPRT_API size_t gprtBufferGetSize(GPRTBuffer _b, int) {
  Buffer *b = (Buffer *)_b;
  return b ? b->size : 0;
}

GPRT_API void *gprtBufferGetHostPointer(GPRTBuffer _b, int) {
  Buffer *b = (Buffer *)_b;
  return b ? b->h : nullptr;
}

GPRT_API void *gprtBufferGetDevicePointer(GPRTBuffer _b, int) {
  Buffer *b = (Buffer *)_b;
  return b ? b->d : nullptr;
}

GPRT_API void gprtBufferMap(GPRTBuffer _b, int) {
  Buffer *b = (Buffer *)_b;
  if (!b) {
    fail("gprtBufferMap: null buffer");
    return;
  }
  // Wait for every enqueued kernel that may write this buffer before copying it out.
  if (!stream_wait(b->ctx, b->ctx->stream, "gprtBufferMap")) return;
  if (b->kind == BufKind::Host) {
    b->mapped = true;
    return;
  }
  if (!b->h && b->size) {
    if (!hip_ok(hipHostMalloc(&b->h, b->size, hipHostMallocDefault), "hipHostMalloc(staging)")) return;
  }
  if (b->size && !hip_ok(hipMemcpy(b->h, b->d, b->size, hipMemcpyDeviceToHost), "hipMemcpy(map)")) return;
  b->mapped = true;
}

GPRT_API void gprtBufferUnmap(GPRTBuffer _b, int) {
  Buffer *b = (Buffer *)_b;
  if (!b) {
    fail("gprtBufferUnmap: null buffer");
    return;
  }
  if (b->kind != BufKind::Host && b->h && b->size) {
    if (!hip_ok(hipMemcpy(b->d, b->h, b->size, hipMemcpyHostToDevice), "hipMemcpy(unmap)")) return;
  }
  b->mapped = false;
}

GPRT_API void gprtBufferResize(GPRTContext, GPRTBuffer _b, size_t size, size_t count, bool preserve, int) {
  Buffer *b = (Buffer *)_b;
  if (!b) {
    fail("gprtBufferResize: null buffer");
    return;
  }
  buffer_realloc(b, size * count, preserve);
}

GPRT_API void gprtBufferCopy(GPRTContext _ctx, GPRTBuffer _src, GPRTBuffer _dst, size_t srcOffset, size_t dstOffset,
                             size_t size, size_t count, int, int) {
  Context *ctx = (Context *)_ctx;
  Buffer *src = (Buffer *)_src, *dst = (Buffer *)_dst;
  if (!ctx || !src || !dst) {
    fail("gprtBufferCopy: null argument");
    return;
  }
  size_t bytes = size * count;
  if ((srcOffset + count) * size > src->size || (dstOffset + count) * size > dst->size) {
    fail("gprtBufferCopy: range out of bounds (src %zu B, dst %zu B, copy %zu B)", src->size, dst->size, bytes);
    return;
  }
  if (!bytes) return;
  hip_ok(hipMemcpyAsync((char *)dst->d + dstOffset * size, (char *)src->d + srcOffset * size, bytes,
                        hipMemcpyDefault, ctx->stream),
         "hipMemcpyAsync(copy)");
}

// ------------------------------------------------------------------------------------------------
// Scan family (gprt/gprt.cpp:8739-8818)
// ------------------------------------------------------------------------------------------------
static uint32_t buffer_scan(GPRTContext _ctx, GPRTBuffer _in, GPRTBuffer _out, GPRTBuffer _scratch, int mode,
                            bool selectPositive, const char *fn) {
  Context *ctx = (Context *)_ctx;
  Buffer *in = (Buffer *)_in, *out = (Buffer *)_out, *scratch = (Buffer *)_scratch;
  if (!ctx || !in || !out) {
    fail("%s: null argument", fn);
    return (uint32_t)-1;
  }
  if (mode != 0 && in == out) {
    fail("%s: input and output cannot be the same buffer", fn);
    return (uint32_t)-1;
  }
  size_t n = in->size / 4;
  if (out->size < n * 4) {
    fail("%s: output holds %zu bytes, needs %zu", fn, out->size, n * 4);
    return (uint32_t)-1;
  }
  if (n == 0) return 0;
  if (n >= (1ull << 32)) {
    fail("%s: %zu items exceeds the 32-bit item count", fn, n);
    return (uint32_t)-1;
  }
  size_t need = scan_scratch_bytes(n);
  void *sp;
  if (scratch) {
    if (scratch->kind == BufKind::Host) {
      fail("%s: scratch must be a device buffer", fn);
      return (uint32_t)-1;
    }
    if (scratch->size < need && !buffer_realloc(scratch, need, false)) return (uint32_t)-1;
    sp = scratch->d;
  } else {
    sp = ctx->ensure_arena(need);
    if (!sp) return (uint32_t)-1;
  }
  // The last 8 bytes of the state hold the total; copy it out through a pinned word.
  uint32_t *totalDev = (uint32_t *)sp;
  scan_run(ctx, ctx->stream, (const uint32_t *)in->d, (uint32_t *)out->d, n, mode, selectPositive, sp, totalDev);
  if (!hip_ok(hipMemcpyAsync(ctx->pinnedWord, totalDev, 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync"))
    return (uint32_t)-1;
  if (!stream_wait(ctx, ctx->stream, fn)) return (uint32_t)-1;
  return *ctx->pinnedWord;
}

// gprtBufferExclusiveSum with the library's static state of the context's device instead of a
// scratch buffer: no allocation of any kind (the total comes back through the context's pinned
// word, allocated with the context). The state is cleared by a fill kernel on every call.
GPRT_API uint32_t gprtBufferExclusiveSumNoAlloc(GPRTContext _ctx, GPRTBuffer _in, GPRTBuffer _out) {
  const StageRange range("gprtBufferExclusiveSumNoAlloc");
  const char *fn = "gprtBufferExclusiveSumNoAlloc";
  Context *ctx = (Context *)_ctx;
  Buffer *in = (Buffer *)_in, *out = (Buffer *)_out;
  if (!ctx || !in || !out) {
    fail("%s: null argument", fn);
    return (uint32_t)-1;
  }
  const size_t n = in->size / 4;
  if (n > scan_noalloc_max_items()) {
    fail("%s: %zu items exceed the %zu-item limit of the library's static scan state (use gprtBufferExclusiveSum)",
         fn, n, scan_noalloc_max_items());
    return (uint32_t)-1;
  }
  if (out->size < n * 4) {
    fail("%s: output holds %zu bytes, needs %zu", fn, out->size, n * 4);
    return (uint32_t)-1;
  }
  if (n == 0) return 0;
  void *state = scan_noalloc_state(ctx->device);
  if (!state) {
    fail("%s: no static scan state on device %d", fn, ctx->device);
    return (uint32_t)-1;
  }
  scan_run(ctx, ctx->stream, (const uint32_t *)in->d, (uint32_t *)out->d, n, 0, false, state, nullptr);
  if (!hip_ok(hipMemcpyAsync(ctx->pinnedWord, state, 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync"))
    return (uint32_t)-1;
  if (!stream_wait(ctx, ctx->stream, fn)) return (uint32_t)-1;
  return *ctx->pinnedWord;
}

GPRT_API uint32_t gprtBufferExclusiveSum(GPRTContext c, GPRTBuffer in, GPRTBuffer out, GPRTBuffer scratch) {
  const StageRange range("gprtBufferExclusiveSum");
  return buffer_scan(c, in, out, scratch, 0, false, "gprtBufferExclusiveSum");
}
GPRT_API uint32_t gprtBufferPartition(GPRTContext c, GPRTBuffer in, bool selectPositive, GPRTBuffer out,
                                      GPRTBuffer scratch) {
  const StageRange range("gprtBufferPartition");
  return buffer_scan(c, in, out, scratch, 1, selectPositive, "gprtBufferPartition");
}
GPRT_API uint32_t gprtBufferSelect(GPRTContext c, GPRTBuffer in, bool selectPositive, GPRTBuffer out,
                                   GPRTBuffer scratch) {
  const StageRange range("gprtBufferSelect");
  return buffer_scan(c, in, out, scratch, 2, selectPositive, "gprtBufferSelect");
}

// ------------------------------------------------------------------------------------------------
// Sort family (gprt/gprt.cpp:8821-9073)
// ------------------------------------------------------------------------------------------------
static void buffer_sort(Context *ctx, Buffer *keys, Buffer *vals, size_t n, int keyBits, int valBits,
                        Buffer *scratch, const char *fn) {
  if (n == 0 || n == 1) return;
  if (n >= (1ull << 30)) {
    fail("%s: %zu keys exceeds the 2^30 limit of the look-back counters", fn, n);
    return;
  }
  const SortConfig cfg = ctx->sortCfg;
  size_t kb = n * (keyBits / 8), vb = n * (valBits / 8);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t stateBytes = sort_scratch_bytes(n, keyBits, valBits, 0, keyBits, cfg);
  size_t need = al(kb) + al(vb) + stateBytes;
  void *sp;
  if (scratch) {
    if (scratch->kind == BufKind::Host) {
      fail("%s: scratch must be a device buffer", fn);
      return;
    }
    if (scratch->size < need && !buffer_realloc(scratch, need, false)) return;
    sp = scratch->d;
  } else {
    sp = ctx->ensure_arena(need);
    if (!sp) return;
  }
  char *p = (char *)sp;
  void *keysAlt = p;
  void *valsAlt = valBits ? p + al(kb) : nullptr;
  void *state = p + al(kb) + al(vb);
  sort_run(ctx, ctx->stream, keys->d, vals ? vals->d : nullptr, keysAlt, valsAlt, n, keyBits, valBits, 0, keyBits,
           state, cfg);
}

GPRT_API void gprtBufferSort(GPRTContext _ctx, GPRTBuffer _buf, GPRTBuffer _scratch) {
  const StageRange range("gprtBufferSort");
  Context *ctx = (Context *)_ctx;
  Buffer *b = (Buffer *)_buf;
  if (!ctx || !b) {
    fail("gprtBufferSort: null argument");
    return;
  }
  if (b->kind == BufKind::Host) {
    fail("gprtBufferSort: keys must be a device buffer");
    return;
  }
  buffer_sort(ctx, b, nullptr, b->size / 8, 64, 0, (Buffer *)_scratch, "gprtBufferSort");
}

GPRT_API void gprtBufferSortPayload(GPRTContext _ctx, GPRTBuffer _keys, GPRTBuffer _vals, GPRTBuffer _scratch) {
  const StageRange range("gprtBufferSortPayload");
  Context *ctx = (Context *)_ctx;
  Buffer *k = (Buffer *)_keys, *v = (Buffer *)_vals;
  if (!ctx || !k) {
    fail("gprtBufferSortPayload: null argument");
    return;
  }
  if (v && v->size != k->size) {
    fail("Keys and Values buffers must be equal in size");  // gprt/gprt.cpp:8832
    return;
  }
  if (k->kind == BufKind::Host || (v && v->kind == BufKind::Host)) {
    fail("gprtBufferSortPayload: keys and values must be device buffers");
    return;
  }
  buffer_sort(ctx, k, v, k->size / 8, 64, v ? 64 : 0, (Buffer *)_scratch, "gprtBufferSortPayload");
}

GPRT_API void gprtBufferSortKV(GPRTContext _ctx, GPRTBuffer _keys, GPRTBuffer _vals, uint32_t count, uint32_t keyBits,
                               uint32_t valueBits, GPRTBuffer _scratch) {
  const StageRange range("gprtBufferSortKV");
  Context *ctx = (Context *)_ctx;
  Buffer *k = (Buffer *)_keys, *v = (Buffer *)_vals;
  if (!ctx || !k) {
    fail("gprtBufferSortKV: null argument");
    return;
  }
  if (keyBits != 32 && keyBits != 64) {
    fail("gprtBufferSortKV: keyBits must be 32 or 64 (got %u)", keyBits);
    return;
  }
  if (valueBits != 0 && valueBits != 32 && valueBits != 64) {
    fail("gprtBufferSortKV: valueBits must be 0, 32 or 64 (got %u)", valueBits);
    return;
  }
  if (valueBits && !v) {
    fail("gprtBufferSortKV: valueBits=%u but no value buffer", valueBits);
    return;
  }
  if ((size_t)count * (keyBits / 8) > k->size) {
    fail("gprtBufferSortKV: %u keys of %u bits exceed the key buffer (%zu B)", count, keyBits, k->size);
    return;
  }
  if (valueBits && (size_t)count * (valueBits / 8) > v->size) {
    fail("gprtBufferSortKV: %u values of %u bits exceed the value buffer (%zu B)", count, valueBits, v->size);
    return;
  }
  if (k->kind == BufKind::Host || (v && v->kind == BufKind::Host)) {
    fail("gprtBufferSortKV: keys and values must be device buffers");
    return;
  }
  buffer_sort(ctx, k, valueBits ? v : nullptr, count, (int)keyBits, (int)valueBits, (Buffer *)_scratch,
              "gprtBufferSortKV");
}

// Width-explicit sorts: the count follows from the key buffer's size and the key width.
static void sort_width(GPRTContext _ctx, GPRTBuffer _keys, GPRTBuffer _vals, bool payload, uint32_t bits,
                       GPRTBuffer _scratch, const char *name) {
  Context *ctx = (Context *)_ctx;
  Buffer *k = (Buffer *)_keys, *v = (Buffer *)_vals;
  if (!ctx || !k || (payload && !v)) {
    fail("%s: null argument", name);
    return;
  }
  const size_t count = k->size / (bits / 8);
  if (payload && v->size < count * (bits / 8)) {
    fail("%s: %zu values of %u bits exceed the value buffer (%zu B)", name, count, bits, v->size);
    return;
  }
  if (k->kind == BufKind::Host || (payload && v->kind == BufKind::Host)) {
    fail("%s: keys and values must be device buffers", name);
    return;
  }
  buffer_sort(ctx, k, payload ? v : nullptr, count, (int)bits, payload ? (int)bits : 0,
              (Buffer *)_scratch, name);
}

GPRT_API void gprtBufferSort32(GPRTContext ctx, GPRTBuffer keys, GPRTBuffer scratch) {
  const StageRange range("gprtBufferSort32");
  sort_width(ctx, keys, nullptr, false, 32, scratch, "gprtBufferSort32");
}

GPRT_API void gprtBufferSort64(GPRTContext ctx, GPRTBuffer keys, GPRTBuffer scratch) {
  const StageRange range("gprtBufferSort64");
  sort_width(ctx, keys, nullptr, false, 64, scratch, "gprtBufferSort64");
}

GPRT_API void gprtBufferSortPayload32(GPRTContext ctx, GPRTBuffer keys, GPRTBuffer values, GPRTBuffer scratch) {
  const StageRange range("gprtBufferSortPayload32");
  sort_width(ctx, keys, values, true, 32, scratch, "gprtBufferSortPayload32");
}

GPRT_API void gprtBufferSortPayload64(GPRTContext ctx, GPRTBuffer keys, GPRTBuffer values, GPRTBuffer scratch) {
  const StageRange range("gprtBufferSortPayload64");
  sort_width(ctx, keys, values, true, 64, scratch, "gprtBufferSortPayload64");
}

GPRT_API void gprtSetSortConfig(GPRTContext _ctx, uint32_t digitBits, uint32_t tileKeys, uint32_t wavesPerCU) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) {
    fail("gprtSetSortConfig: null context");
    return;
  }
  if (digitBits != 4 && digitBits != 6 && digitBits != 8) {
    fail("gprtSetSortConfig: digitBits %u is not one of 4, 6, 8", digitBits);
    return;
  }
  if (tileKeys != 0 && tileKeys != 2048 && tileKeys != 4096 && tileKeys != 8192 && tileKeys != 16384) {
    fail("gprtSetSortConfig: tileKeys %u is not one of 0 (automatic), 2048, 4096, 8192, 16384", tileKeys);
    return;
  }
  if (wavesPerCU > 64) {
    fail("gprtSetSortConfig: wavesPerCU %u exceeds 64", wavesPerCU);
    return;
  }
  ctx->sortCfg.digitBits = (int)digitBits;
  ctx->sortCfg.tileKeys = (int)tileKeys;
  ctx->sortCfg.wavesPerCU = (int)wavesPerCU;
}

GPRT_API void gprtGetSortConfig(GPRTContext _ctx, uint32_t *digitBits, uint32_t *tileKeys, uint32_t *wavesPerCU) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) {
    fail("gprtGetSortConfig: null context");
    return;
  }
  if (digitBits) *digitBits = (uint32_t)ctx->sortCfg.digitBits;
  if (tileKeys) *tileKeys = (uint32_t)ctx->sortCfg.tileKeys;
  if (wavesPerCU) *wavesPerCU = (uint32_t)ctx->sortCfg.wavesPerCU;
}

// ------------------------------------------------------------------------------------------------
// Synchronisation and profiling (gprt/gprt.cpp:9482-9545)
// ------------------------------------------------------------------------------------------------
GPRT_API uint64_t gprtComputeSynchronize(GPRTContext _ctx) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) return 0;
  stream_wait(ctx, ctx->stream, "gprtComputeSynchronize");
  return 0;
}

GPRT_API uint64_t gprtDeviceSynchronize(GPRTContext _ctx) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) return 0;
  if (stream_wait(ctx, ctx->stream, "gprtDeviceSynchronize")) hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return 0;
}

GPRT_API void gprtSetSyncTimeout(GPRTContext _ctx, double seconds) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) {
    fail("gprtSetSyncTimeout: null context");
    return;
  }
  if (!(seconds >= 0.0)) {
    fail("gprtSetSyncTimeout: %g s is not a duration (0 waits without limit)", seconds);
    return;
  }
  ctx->syncTimeout = seconds;
}

GPRT_API void gprtBeginProfile(GPRTContext _ctx) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) return;
  hipEventRecord(ctx->profBegin, ctx->stream);
  ctx->profiling = true;
}

GPRT_API float gprtEndProfile(GPRTContext _ctx) {
  Context *ctx = (Context *)_ctx;
  if (!ctx || !ctx->profiling) return 0.f;
  hipEventRecord(ctx->profEnd, ctx->stream);
  hipEventSynchronize(ctx->profEnd);
  float ms = 0.f;
  hipEventElapsedTime(&ms, ctx->profBegin, ctx->profEnd);
  ctx->profiling = false;
  return ms;
}

GPRT_API void gprtKernelTimerEnable(GPRTContext _ctx, bool enable) {
  Context *ctx = (Context *)_ctx;
  if (ctx) ctx->timer.enabled = enable;
}

GPRT_API void gprtKernelTimerReset(GPRTContext _ctx) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) return;
  ctx->timer.drain();
  ctx->timer.totals.clear();
}

GPRT_API double gprtKernelTimerQuery(GPRTContext _ctx, const char *name, uint32_t *launches) {
  Context *ctx = (Context *)_ctx;
  if (!ctx || !name) return 0.0;
  ctx->timer.drain();
  double ms = 0.0;
  uint32_t n = 0;
  for (auto &kv : ctx->timer.totals) {
    if (kv.first.find(name) != std::string::npos) {
      ms += kv.second.first;
      n += kv.second.second;
    }
  }
  if (launches) *launches = n;
  return ms;
}

GPRT_API size_t gprtKernelTimerReport(GPRTContext _ctx, char *out, size_t cap) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) return 0;
  ctx->timer.drain();
  std::string s;
  char line[256];
  for (auto &kv : ctx->timer.totals) {
    snprintf(line, sizeof(line), "%s,%u,%.6f\n", kv.first.c_str(), kv.second.second, kv.second.first);
    s += line;
  }
  if (out && cap) {
    size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
    memcpy(out, s.data(), n);
    out[n] = 0;
  }
  return s.size() + 1;
}
