/*
 * gprt_hip.h — C-ABI of the MI355X (gfx950) software acceleration-structure path.
 *
 * The entry points keep the names, argument order and meaning of GPRT's public host API
 * (gprt/gprt_host.h) for the path they replace: buffers, radix sort, exclusive scan /
 * partition / select, triangle accel create / build / update, synchronisation and profiling.
 * Every exported symbol is `extern "C"`, takes opaque handles, plain pointers and sizes, and
 * carries no C++ or torch types, so cgo / ctypes / JNI can bind it directly.
 *
 * Each declaration cites the reference declaration it stands in for (file:line in GPRT).
 * Declarations under "Extensions" have no reference counterpart; each says why it exists.
 *
 * Error convention (gprt/gprt.cpp:205-251): an invalid argument or a failed HIP call prints
 * "#gprt error: <message>" and raises SIGINT, as the reference's LOG_ERROR does.
 * gprtSetErrorMode(GPRT_ERROR_RECORD) switches to a recording mode in which the message is kept
 * for gprtGetLastError() and the call returns without touching the GPU (used by tests).
 */
#ifndef GPRT_HIP_H
#define GPRT_HIP_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
#define GPRT_API extern "C" __attribute__((visibility("default")))
#define GPRT_IF_CPP(a) a
#else
#define GPRT_API __attribute__((visibility("default")))
#define GPRT_IF_CPP(a)
#endif

/* Opaque handles: gprt/gprt_host.h:104-117 (GPRT_DEFINE_HANDLE). */
typedef struct GPRTContext_T *GPRTContext;
typedef struct GPRTBuffer_T *GPRTBuffer;
typedef struct GPRTGeom_T *GPRTGeom;
typedef struct GPRTGeomType_T *GPRTGeomType;
typedef struct GPRTAccel_T *GPRTAccel;

/* gprt/gprt_host.h:286-288 */
typedef enum {
  GPRT_UNKNOWN,
  GPRT_AABBS,
  GPRT_TRIANGLES,
  GPRT_SPHERES,
  GPRT_LSS,
  GPRT_SOLIDS,
  GPRT_PARTICLES
} GPRTGeomKind;

/* gprt/gprt_host.h:349-355 */
typedef enum {
  GPRT_BUILD_MODE_UNINITIALIZED,
  GPRT_BUILD_MODE_FAST_BUILD_NO_UPDATE,
  GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE,
  GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE,
  GPRT_BUILD_MODE_FAST_TRACE_AND_UPDATE,
} GPRTBuildMode;

/* gprt/gprt_host.h:1423-1469 — same members, same order, same defaults. */
typedef struct GPRTBuildParams {
  GPRTBuildMode buildMode GPRT_IF_CPP(= GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE);
  bool allowCompaction GPRT_IF_CPP(= false);
  bool minimizeMemory GPRT_IF_CPP(= false);
  bool hasMotionBlur GPRT_IF_CPP(= false);
  bool test0 GPRT_IF_CPP(= false);
  bool test1 GPRT_IF_CPP(= false);
} GPRTBuildParams;

/* ---------------------------------------------------------------------------------------------
 * Context (gprt/gprt_host.h:992-995). One HIP device and one in-order stream per context.
 * numDevices must be 1: the reference has no multi-GPU path either (gprt/gprt.cpp:5373-5374).
 * ------------------------------------------------------------------------------------------- */
GPRT_API GPRTContext gprtContextCreate(int32_t *requestedDeviceIDs GPRT_IF_CPP(= nullptr),
                                       int numDevices GPRT_IF_CPP(= 1));
GPRT_API void gprtContextDestroy(GPRTContext context);

/* ---------------------------------------------------------------------------------------------
 * Buffers (gprt/gprt_host.h:1899-2210). `size` is the element size in bytes, `count` the number
 * of elements, exactly as in the reference. Device and shared buffers live in HBM and are read
 * and written by the host through map/unmap, which first waits for the context's stream (the
 * reference's map does not, gprt/gprt.cpp:6774; waiting here removes that hazard).
 * ------------------------------------------------------------------------------------------- */
GPRT_API GPRTBuffer gprtHostBufferCreate(GPRTContext context, size_t size, size_t count GPRT_IF_CPP(= 1),
                                         const void *init GPRT_IF_CPP(= nullptr),
                                         size_t alignment GPRT_IF_CPP(= 16)); /* gprt_host.h:1899 */
GPRT_API GPRTBuffer gprtDeviceBufferCreate(GPRTContext context, size_t size, size_t count GPRT_IF_CPP(= 1),
                                           const void *init GPRT_IF_CPP(= nullptr),
                                           size_t alignment GPRT_IF_CPP(= 16)); /* gprt_host.h:1946 */
GPRT_API GPRTBuffer gprtSharedBufferCreate(GPRTContext context, size_t size, size_t count GPRT_IF_CPP(= 1),
                                           const void *init GPRT_IF_CPP(= nullptr),
                                           size_t alignment GPRT_IF_CPP(= 16)); /* gprt_host.h:1995 */
GPRT_API void gprtBufferClear(GPRTBuffer buffer);                            /* gprt_host.h:2037 */
GPRT_API void gprtBufferDestroy(GPRTBuffer buffer);                          /* gprt_host.h:2053 */
GPRT_API size_t gprtBufferGetSize(GPRTBuffer buffer, int deviceID GPRT_IF_CPP(= 0)); /* :2067 */
GPRT_API void *gprtBufferGetHostPointer(GPRTBuffer buffer, int deviceID GPRT_IF_CPP(= 0)); /* :2082 */
GPRT_API void gprtBufferMap(GPRTBuffer buffer, int deviceID GPRT_IF_CPP(= 0));     /* :2108 */
GPRT_API void gprtBufferUnmap(GPRTBuffer buffer, int deviceID GPRT_IF_CPP(= 0));   /* :2116 */
GPRT_API void *gprtBufferGetDevicePointer(GPRTBuffer buffer, int deviceID GPRT_IF_CPP(= 0)); /* :2128 */
GPRT_API void gprtBufferResize(GPRTContext context, GPRTBuffer buffer, size_t size, size_t count,
                               bool preserveContents, int deviceID GPRT_IF_CPP(= 0)); /* :2150 */
GPRT_API void gprtBufferCopy(GPRTContext context, GPRTBuffer src, GPRTBuffer dst, size_t srcOffset,
                             size_t dstOffset, size_t size, size_t count, int srcDeviceID GPRT_IF_CPP(= 0),
                             int dstDeviceID GPRT_IF_CPP(= 0)); /* :2187 */

/* ---------------------------------------------------------------------------------------------
 * Scan family (gprt/gprt_host.h:2246, :2295, :2344; stub in gprt/gprt.cpp:8739-8818).
 * Input is 32-bit; the element count is bytes/4 (gprt/gprt.cpp:8750 in the commented-out path).
 * ExclusiveSum wraps modulo 2^32 (no 2^30 limit: look-back words are 64-bit) and returns the
 * total. Partition places selected items first in input order and unselected items at the tail
 * in reverse order (gprt/gprt_scan.slang:89-92); Select compacts the selected items. Selection
 * is "v >= 0" when selectPositive, else "v < 0" (tests/t03-bufferScans/hostCode.cpp:117).
 * scratch may be null (internal allocation) or undersized (resized in place), as documented at
 * gprt/gprt_host.h:2238-2241. The return value requires a host wait (the reference maps scratch).
 * ------------------------------------------------------------------------------------------- */
GPRT_API uint32_t gprtBufferExclusiveSum(GPRTContext context, GPRTBuffer input, GPRTBuffer output,
                                         GPRTBuffer scratch GPRT_IF_CPP(= nullptr));
GPRT_API uint32_t gprtBufferPartition(GPRTContext context, GPRTBuffer input, bool selectPositive,
                                      GPRTBuffer output, GPRTBuffer scratch GPRT_IF_CPP(= nullptr));
GPRT_API uint32_t gprtBufferSelect(GPRTContext context, GPRTBuffer input, bool selectPositive, GPRTBuffer output,
                                   GPRTBuffer scratch GPRT_IF_CPP(= nullptr));

/* ---------------------------------------------------------------------------------------------
 * Radix sort (gprt/gprt_host.h:2384, :2416; gprt/gprt.cpp:8821-9073). Stable LSD sort, ascending,
 * in place; the payload follows its key. As in the reference (gprt/gprt.cpp:8837) these two
 * entry points treat keys as 64-bit (count = bytes/8) and the payload as 64-bit; the key and
 * payload buffers must be the same size (gprt/gprt.cpp:8829-8835). Enqueued; returns at once.
 * ------------------------------------------------------------------------------------------- */
GPRT_API void gprtBufferSort(GPRTContext context, GPRTBuffer buffer, GPRTBuffer scratch GPRT_IF_CPP(= nullptr));
GPRT_API void gprtBufferSortPayload(GPRTContext context, GPRTBuffer keys, GPRTBuffer values,
                                    GPRTBuffer scratch GPRT_IF_CPP(= nullptr));

/* ---------------------------------------------------------------------------------------------
 * Triangle geometry and accel (gprt/gprt_host.h:406, :476, :498, :1308, :1421, :1506, :1529,
 * :1553, :1557, :1580). The accel is a binary LBVH: 30-bit Morton keys of triangle centroids
 * normalised to the scene box, packed as morton<<32|primID, sorted; the Karras radix tree (same
 * topology and node numbering) is built bottom-up together with its boxes, every node merged once
 * by the second of its children to arrive. Build and update are enqueued on the stream.
 * ------------------------------------------------------------------------------------------- */
GPRT_API GPRTGeomType gprtGeomTypeCreate(GPRTContext context, GPRTGeomKind kind, size_t recordSize);
GPRT_API void gprtGeomTypeDestroy(GPRTGeomType geomType);
GPRT_API GPRTGeom gprtGeomCreate(GPRTContext context, GPRTGeomType type);
GPRT_API void gprtGeomDestroy(GPRTGeom geometry);
GPRT_API void gprtTrianglesSetVertices(GPRTGeom triangles, GPRTBuffer vertices, uint32_t count,
                                       uint32_t stride GPRT_IF_CPP(= 12), uint32_t offset GPRT_IF_CPP(= 0));
GPRT_API void gprtTrianglesSetIndices(GPRTGeom triangles, GPRTBuffer indices, uint32_t count,
                                      uint32_t stride GPRT_IF_CPP(= 12), uint32_t offset GPRT_IF_CPP(= 0));
GPRT_API GPRTAccel gprtTriangleAccelCreate(GPRTContext context, GPRTGeom geom, unsigned int flags GPRT_IF_CPP(= 0));

/* Box primitives (gprt/gprt_host.h:614, :1282): `count` boxes {minX, minY, minZ, maxX, maxY, maxZ}
 * at offset + i * stride (stride >= 24, default 2 * sizeof(float3)) of a GPRT_AABBS geometry, and
 * an accel over them built by the same LBVH builder (keys from the box centres (lo + hi) * 0.5).
 * The reference runs a user intersection program on the boxes; here the closest hit is the box
 * surface: the entry distance when the ray enters the box after tmin, else the exit distance when
 * it starts inside (u = v = 0). Box accels can be built alone or in batches, updated, traced,
 * exported and given a proposal; instance accels reference triangle accels only. */
GPRT_API void gprtAABBsSetPositions(GPRTGeom aabbs, GPRTBuffer positions, uint32_t count,
                                    uint32_t stride GPRT_IF_CPP(= 24), uint32_t offset GPRT_IF_CPP(= 0));
GPRT_API GPRTAccel gprtAABBAccelCreate(GPRTContext context, GPRTGeom geom, unsigned int flags GPRT_IF_CPP(= 0));
GPRT_API void gprtAccelDestroy(GPRTAccel accel);
GPRT_API void gprtAccelBuild(GPRTContext context, GPRTAccel accel,
                             GPRTBuildParams buildParams GPRT_IF_CPP(= GPRTBuildParams()));
GPRT_API void gprtAccelUpdate(GPRTContext context, GPRTAccel accel);
GPRT_API size_t gprtAccelGetSize(GPRTAccel accel, int deviceID GPRT_IF_CPP(= 0));
GPRT_API const void *gprtAccelGetDeviceAddress(GPRTAccel accel, int deviceID GPRT_IF_CPP(= 0));

/* ---------------------------------------------------------------------------------------------
 * User compute programs (modules gprt/gprt_host.h:403-404, compute handles :1007 / :1025,
 * launches :2537-2586; limits gprt/gprt_shared.h:60-67).
 *
 * A module is device code for gfx950: a code object or an offload bundle holding one, as
 * `hipcc --genco --offload-arch=gfx950` writes it (the reference takes SPIR-V). A compute handle
 * names one of its kernels by symbol (declare it `extern "C" __global__`). A launch runs
 * numGroups workgroups of groupSize threads on the context's stream, ordered after every earlier
 * call on the context (sorts, builds, traces, copies) and before every later one, so a program
 * can fill the buffers a build or trace reads. Its arguments are the kernel's parameters, in
 * order, each at its natural alignment (the HIP kernel-argument layout; for a single record
 * argument this is the reference's push-constant block byte for byte), at most
 * GPRT_PUSH_CONSTANTS_LIMIT bytes in all. Each group count must be at most
 * GPRT_WORKGROUP_LIMIT and a group at most GPRT_THREADGROUP_LIMIT threads, as the reference's
 * launch asserts (gprt/gprt_host.h:2549-2554); a group count of 0 launches nothing. Launches
 * return 0 (the reference returns its timeline counter). Destroying a module waits for the
 * context's stream; compute handles of a destroyed module report an error when launched.
 * ------------------------------------------------------------------------------------------- */
typedef struct GPRTModule_T *GPRTModule;
typedef struct GPRTCompute_T *GPRTCompute;
#define GPRT_PUSH_CONSTANTS_LIMIT 256 /* gprt/gprt_shared.h:60 */
#define GPRT_WORKGROUP_LIMIT 65535    /* gprt/gprt_shared.h:66 */
#define GPRT_THREADGROUP_LIMIT 1024   /* gprt/gprt_shared.h:67 */
GPRT_API GPRTModule gprtModuleCreate(GPRTContext context, const void *code, size_t bytes); /* :403 */
GPRT_API void gprtModuleDestroy(GPRTModule module);                                        /* :404 */
GPRT_API GPRTCompute gprtComputeCreate(GPRTContext context, GPRTModule module, const char *entrypoint); /* :1007 */
GPRT_API void gprtComputeDestroy(GPRTCompute compute);                                                /* :1025 */
/* The untyped launch under the templates below (the reference's _gprtComputeLaunch,
 * gprt/gprt_host.h:2539): `args` holds `bytes` bytes of packed arguments. */
GPRT_API uint64_t gprtComputeLaunchArgs(GPRTCompute compute, const uint32_t numGroups[3], const uint32_t groupSize[3],
                                        const void *args, size_t bytes);

#ifdef __cplusplus
/* C++ conveniences with the reference's spelling of the same calls. */
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>
using GPRTProgram = std::vector<uint8_t>; /* gprt/gprt_host.h:136 */
template <typename... T> struct GPRTComputeOf_T;
template <typename... T> using GPRTComputeOf = GPRTComputeOf_T<T...> *; /* typed handle, as :131 */
struct GPRTUint3 {
  uint32_t x, y, z;
};
inline GPRTModule gprtModuleCreate(GPRTContext context, const GPRTProgram &code) {
  return gprtModuleCreate(context, code.data(), code.size());
}
template <typename... Uniforms>
GPRTComputeOf<Uniforms...> gprtComputeCreate(GPRTContext context, GPRTModule module, const std::string &entrypoint) {
  return (GPRTComputeOf<Uniforms...>)gprtComputeCreate(context, module, entrypoint.c_str());
}
template <typename... Uniforms> void gprtComputeDestroy(GPRTComputeOf<Uniforms...> compute) {
  gprtComputeDestroy((GPRTCompute)compute);
}
/* Bytes the arguments take when each starts at its natural alignment. */
template <typename... Args> constexpr size_t gprtPackedSize() {
  size_t off = 0;
  ((off = (off + alignof(Args) - 1) / alignof(Args) * alignof(Args) + sizeof(Args)), ...);
  return off;
}
/* Packs the arguments as a kernel's parameter list; returns the bytes written. */
template <typename... Args> size_t gprtPackArgs(unsigned char *block, const Args &...args) {
  static_assert((std::is_trivially_copyable<Args>::value && ...), "arguments must be trivially copyable");
  size_t off = 0;
  ((off = (off + alignof(Args) - 1) / alignof(Args) * alignof(Args), std::memcpy(block + off, &args, sizeof(Args)),
    off += sizeof(Args)),
   ...);
  return off;
}
template <typename... Uniforms>
uint64_t gprtComputeLaunch(GPRTCompute compute, GPRTUint3 numGroups, GPRTUint3 groupSize, Uniforms... uniforms) {
  static_assert(gprtPackedSize<Uniforms...>() <= GPRT_PUSH_CONSTANTS_LIMIT,
                "Total size of arguments exceeds GPRT_PUSH_CONSTANTS_LIMIT bytes");
  alignas(16) unsigned char block[GPRT_PUSH_CONSTANTS_LIMIT] = {};
  const size_t bytes = gprtPackArgs(block, uniforms...);
  const uint32_t g[3] = {numGroups.x, numGroups.y, numGroups.z}, t[3] = {groupSize.x, groupSize.y, groupSize.z};
  return gprtComputeLaunchArgs(compute, g, t, block, bytes);
}
template <typename... Uniforms>
uint64_t gprtComputeLaunch(GPRTComputeOf<Uniforms...> compute, GPRTUint3 numGroups, GPRTUint3 groupSize,
                           Uniforms... uniforms) {
  return gprtComputeLaunch<Uniforms...>((GPRTCompute)compute, numGroups, groupSize, uniforms...);
}
#endif

/* ---------------------------------------------------------------------------------------------
 * Synchronisation and profiling (gprt/gprt_host.h:2588-2595). Begin/EndProfile time the work
 * enqueued between the two calls on the context's stream, in milliseconds.
 * ------------------------------------------------------------------------------------------- */
GPRT_API uint64_t gprtComputeSynchronize(GPRTContext context);
GPRT_API uint64_t gprtDeviceSynchronize(GPRTContext context);
GPRT_API void gprtBeginProfile(GPRTContext context);
GPRT_API float gprtEndProfile(GPRTContext context);

/* =============================================================================================
 * Extensions (no reference counterpart)
 * =========================================================================================== */

/* Error handling for callers that cannot take SIGINT (language bindings, test runners). */
#define GPRT_ERROR_RAISE 0  /* default: print and raise(SIGINT), as gprt/gprt.cpp:247-251 */
#define GPRT_ERROR_RECORD 1 /* print, record the message, return without GPU work */
GPRT_API void gprtSetErrorMode(int mode);
/* Host-wait deadline (the reference waits with vkWaitSemaphores and syncTDR = 6 s,
 * gprt/gprt.cpp:144, :6785-6840, and treats a timeout as an error). gprtComputeSynchronize,
 * gprtDeviceSynchronize, gprtBufferMap and the scan family's returned totals poll the context's
 * stream and report work that outlives `seconds` through the error path (default 6; 0 waits
 * without limit). The work itself is not cancelled. */
GPRT_API void gprtSetSyncTimeout(GPRTContext context, double seconds);
GPRT_API const char *gprtGetLastError(void); /* NULL if no error since the last clear */
GPRT_API void gprtClearLastError(void);

/* The width-explicit sorts (gprtBufferSort32/64, gprtBufferSortPayload32/64, gprtBufferSortKV)
 * are declared extensions and live in gprt_hip_ext.h. */

/* No counterpart in the reference: gprtBufferExclusiveSum for callers that cannot allocate
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc). Same result
 * and returned total as gprtBufferExclusiveSum(context, input, output, scratch); instead of a
 * scratch buffer its look-back state is one static device array per GPU, held until the process
 * ends (1 MiB), so the call makes no allocation of any kind. Input is at most 2^31 items
 * (8 GiB); larger inputs fail through the error path. Calls on one GPU share that array: calls
 * that overlap on one GPU (two contexts' streams, or two host threads) race and must be ordered
 * by the caller. Waits for the context's stream to return the total, as the original does. */
GPRT_API uint32_t gprtBufferExclusiveSumNoAlloc(GPRTContext context, GPRTBuffer input, GPRTBuffer output);

/* Radix sort knobs used by every sort and accel build on this context (the autotuner's search
 * space): digitBits in {4, 6, 8}, tileKeys (keys per workgroup tile) in {2048, 4096, 8192, 16384}
 * or 0 (automatic: 16384 for accel builds averaging at least 2^20 primitives per mesh, else 8192;
 * for the buffer sorts 16384 from 2^22 keys, else 8192; 64-bit keys or payloads use 4096 unless
 * 2048 is asked for), wavesPerCU (0 = no cap) caps the resident waves of the accel
 * builds' sort passes through LDS padding (no effect at 16384 keys, whose 1024-thread tiles
 * already run one per CU). Results do not depend on them. Default: 8, 0, 0. */
GPRT_API void gprtSetSortConfig(GPRTContext context, uint32_t digitBits, uint32_t tileKeys, uint32_t wavesPerCU);
GPRT_API void gprtGetSortConfig(GPRTContext context, uint32_t *digitBits, uint32_t *tileKeys, uint32_t *wavesPerCU);

/* Batched build: every accel in `accels` (each over its own triangle geometry) is built by one
 * launch sequence over a shared tile table — bounds, keys, segmented radix sort, hierarchy and
 * boxes of all meshes together — captured into a hipGraph on the first call and replayed by later calls
 * with the same accels and geometry shapes (vertex positions may change between calls). The
 * accels share one device store until they are rebuilt alone or destroyed. Each accel's result
 * is identical to gprtAccelBuild on it. The reference builds every accel separately with a
 * device synchronisation per build (gprt/gprt.cpp:3216, :3271, :3285). */
GPRT_API void gprtAccelBuildBatch(GPRTContext context, GPRTAccel *accels, uint32_t count,
                                  GPRTBuildParams buildParams GPRT_IF_CPP(= GPRTBuildParams()));

/* Project-written autotuner for batched builds: times the graph-captured batch build for every
 * digitBits {4,6,8} x tileKeys {2048,4096,8192} x wavesPerCU {0,8,16}, and tileKeys 16384 (one
 * workgroup per CU, so without a cap) for each digit width: 30 configurations (median of `reps` launches),
 * keeps the fastest configuration in the context (gprtGetSortConfig reads it back), leaves the
 * accels built with it and writes a JSON report. Returns the index of the best configuration. */
GPRT_API uint32_t gprtAccelBuildBatchAutotune(GPRTContext context, GPRTAccel *accels, uint32_t count, uint32_t reps,
                                              char *report, size_t cap);

/* Ray batch closest hit against a built triangle accel. rays: `count` records of
 * {float ox, oy, oz, tmin; float dx, dy, dz, tmax} (32 B); hits: `count` records of
 * {float t; int32 primID; float u, v} (16 B), primID = -1 on a miss. A hit is accepted when
 * tmin < t < tbest, or t == tbest with a smaller primID, so the result is the lexicographic
 * minimum of (t, primID) over all triangles the watertight ray–triangle test accepts and does not
 * depend on traversal order. The reference traces through vkCmdTraceRaysKHR
 * (gprt/gprt.cpp:9404); this is the software form of that closest-hit query. */
typedef struct GPRTRay {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
} GPRTRay;
typedef struct GPRTHit {
  float t;
  int32_t primID;
  float u, v;
} GPRTHit;
GPRT_API void gprtAccelTraceClosestHit(GPRTContext context, GPRTAccel accel, GPRTBuffer rays, GPRTBuffer hits,
                                       uint32_t count);

/* Learned treelet proposal (no reference counterpart). A 102-128-128-128-4096 MLP in bf16 on
 * MFMA (FP32 accumulate) maps a ray (origin, direction with 8-band sin/cos encoding) to one of
 * the treelets obtained by cutting the accel's LBVH at `cutDepth` (<= 12: at most 4096 treelets).
 * It is trained on the device (hand-written forward / backward / Adam) with labels taken from
 * the plain traversal's hits. gprtAccelTraceClosestHitEx with a proposal first traverses the
 * proposed treelet to get a hit that tightens tmax, then runs the exact traversal from the root:
 * hits are identical to the plain trace (the exact traversal always decides the result).
 * Parameters use the flat fp32 layout W1t b1 W2t b2 W3t b3 W4t b4 with Wt[out][in] and the
 * first layer's input padded to 128 (gprtProposalParamCount() values). */
typedef struct GPRTProposal_T *GPRTProposal;
GPRT_API GPRTProposal gprtProposalCreate(GPRTContext context, uint32_t maxBatch, uint64_t seed);
GPRT_API void gprtProposalDestroy(GPRTContext context, GPRTProposal proposal);
/* Cuts the accel's tree at cutDepth; returns the number of treelets. Rebind after a rebuild.
 * Treelets are numbered from the tree alone (breadth-first, left to right within a level), so a
 * trained network bound again to the same tree, in this process or another, keeps its classes.
 * For an instance accel the classes are its top-level leaves (cutDepth is ignored; at most
 * 4096). */
GPRT_API uint32_t gprtProposalBindAccel(GPRTContext context, GPRTProposal proposal, GPRTAccel accel,
                                        uint32_t cutDepth);
GPRT_API size_t gprtProposalParamCount(void);
GPRT_API void gprtProposalSetParams(GPRTContext context, GPRTProposal proposal, const float *params, size_t count);
GPRT_API void gprtProposalGetParams(GPRTContext context, GPRTProposal proposal, float *params, size_t count);
/* Weight checkpoints (the on-disk form of a trained proposal; no counterpart in the reference,
 * which has no serialisation). Little-endian, version 1: a 64-byte header {char magic[8] =
 * "GPRTMLP\0", u32 version = 1, u32 headerBytes = 64, u64 parameter count, u32 input, hidden,
 * hidden layers, output widths (102, 128, 3, 4096), u32 Adam step, u32 flags (bit 0: Adam
 * moments follow), u32 CRC-32 (IEEE 802.3, as zlib's crc32) of the payload, 12 reserved zero
 * bytes}, then the fp32 parameters in the layout above and, with bit 0, the first and second
 * Adam moments. Serialize returns the checkpoint's size and writes it when `capacity` suffices
 * (query with out = NULL). Deserialize checks every header field and the CRC and returns false
 * (through the error path) on any mismatch; it restores the optimizer state when present, else
 * resets it as gprtProposalSetParams does. Treelet bindings are not part of a checkpoint. */
GPRT_API size_t gprtProposalSerialize(GPRTContext context, GPRTProposal proposal, void *out, size_t capacity,
                                      bool withOptimizerState);
GPRT_API bool gprtProposalDeserialize(GPRTContext context, GPRTProposal proposal, const void *in, size_t bytes);
/* Epochs of minibatch Adam over the rays the plain traversal hit (hits: GPRTHit per ray, as
 * written by gprtAccelTraceClosestHit). Returns the last epoch's mean loss; *accuracy = top-1. */
GPRT_API float gprtProposalTrain(GPRTContext context, GPRTProposal proposal, GPRTBuffer rays, GPRTBuffer hits,
                                 uint32_t count, uint32_t epochs, float learningRate, float *accuracy);
/* Same with explicit labels (int32 class per ray, < 0 = not labelled), e.g. the leafIDs of an
 * instance-accel trace, whose classes are its top-level leaves. */
GPRT_API float gprtProposalTrainLabels(GPRTContext context, GPRTProposal proposal, GPRTBuffer rays,
                                       GPRTBuffer labels, uint32_t count, uint32_t epochs, float learningRate,
                                       float *accuracy);
/* Top-1 treelet per ray (int32), fused: the 4096 logits are never written to memory. */
GPRT_API void gprtProposalInfer(GPRTContext context, GPRTProposal proposal, GPRTBuffer rays, uint32_t count,
                                GPRTBuffer classes);
/* Test access: fp32 logits (count x 4096, count <= maxBatch) and the gradient of the mean
 * cross-entropy over `count` labelled rays (int32 labels) in the flat parameter layout. */
GPRT_API void gprtProposalLogits(GPRTContext context, GPRTProposal proposal, GPRTBuffer rays, uint32_t count,
                                 GPRTBuffer logits);
GPRT_API void gprtProposalGradients(GPRTContext context, GPRTProposal proposal, GPRTBuffer rays, GPRTBuffer labels,
                                    uint32_t count, float *grads, size_t gradCount);
/* classRoot[c]: treelet root (inner node index, or ~leaf position); primClass[prim]: its treelet. */
GPRT_API void gprtProposalExportTreelets(GPRTContext context, GPRTProposal proposal, GPRTBuffer classRoot,
                                         GPRTBuffer primClass);
/* Confidence gate of proposal-first traces: a ray whose top logit is below `threshold` is traced
 * without its proposal (for such a ray the treelet walk more likely costs than saves node fetches).
 * -INFINITY (the default) keeps the proposal of every ray; +INFINITY drops all, and the trace then
 * skips the inference too (it is the plain trace). Hits never depend on it. gprtProposalInfer is
 * not gated. */
GPRT_API void gprtProposalSetGate(GPRTContext context, GPRTProposal proposal, float threshold);
GPRT_API float gprtProposalGetGate(GPRTProposal proposal);
/* Sets the gate from measured cost on a sample of rays (use rays of the workload's distribution,
 * not the ones to be timed): each ray is traced plainly and proposal-first with its inner node
 * fetches counted, and the threshold that minimises the sample's total fetches is kept (the empty
 * choice costs nothing, so the gate never makes the sample dearer than the plain trace). Returns
 * the share of the sample that keeps its proposal. */
GPRT_API float gprtProposalCalibrateGate(GPRTContext context, GPRTProposal proposal, GPRTAccel accel, GPRTBuffer rays,
                                         uint32_t count);
/* Overlap of proposal-first traces (a tuning knob; results do not depend on it): with chunks >= 2
 * the rays are split into that many chunks (of at least 65,536 rays), each chunk's inference runs
 * on a second stream of the context while the previous chunk is traversed, and the traversal runs
 * `traceBlocksPerCU` persistent workgroups per CU (0: as many as the plain trace; fewer leave room
 * for the inference on the same CUs). chunks < 2 (the default): inference, then one traversal.
 * Later calls on the context are ordered after the whole trace, as for any call. */
GPRT_API void gprtProposalSetOverlap(GPRTContext context, GPRTProposal proposal, uint32_t chunks,
                                     uint32_t traceBlocksPerCU);
/* Closest hit with an optional proposal (may be NULL) and optional per-ray count of inner nodes
 * fetched (visits, uint32, may be NULL). */
GPRT_API void gprtAccelTraceClosestHitEx(GPRTContext context, GPRTAccel accel, GPRTProposal proposal,
                                         GPRTBuffer rays, GPRTBuffer hits, uint32_t count, GPRTBuffer visits);

/* Instance accels (two-level scenes). gprtInstanceAccelCreate / gprtAccelGetInstance keep the
 * reference's names and the gprt::Instance record (gprt/gprt_host.h:1418, :1570;
 * gprt/gprt_shared.h:92-101): a 3x4 row-major transform [A | b] (world = A * object + b),
 * instanceCustomIndex:24, mask:8 (0 hides the instance), geometryIndex:24, flags:8, and
 * accelAddress = gprtAccelGetDeviceAddress of a built triangle accel. Build with gprtAccelBuild
 * after the triangle accels it names are built (and again after they are rebuilt or the
 * records change). Traces (gprtAccelTraceClosestHit / Ex / Detail) return, per ray, the closest
 * hit over all instances under the order (t, instance index, primID), primID being local to
 * the instance's geometry; t is the ray parameter of the world-space ray. At most 65536
 * instances; an address that does not name a live, built triangle accel of this context
 * disables its instance. */
typedef struct GPRTInstance {
  float transform[3][4];
  uint32_t instanceCustomIndex : 24;
  uint32_t mask : 8;
  uint32_t geometryIndex : 24;
  uint32_t flags : 8;
  uint64_t accelAddress;
} GPRTInstance;
GPRT_API GPRTAccel gprtInstanceAccelCreate(GPRTContext context, uint32_t numInstances, GPRTBuffer instances);
GPRT_API GPRTInstance gprtAccelGetInstance(GPRTAccel blas);
/* Closest hit with optional outputs (each may be NULL): instanceIDs (int32, -1 on a miss; 0 for
 * a triangle accel), leafIDs (int32: for an instance accel the top-level leaf holding the hit,
 * the class the proposal learns for such accels; -1 otherwise) and visits (uint32, inner nodes
 * fetched). `proposal` may be NULL. */
GPRT_API void gprtAccelTraceClosestHitDetail(GPRTContext context, GPRTAccel accel, GPRTProposal proposal,
                                             GPRTBuffer rays, GPRTBuffer hits, uint32_t count,
                                             GPRTBuffer instanceIDs, GPRTBuffer leafIDs, GPRTBuffer visits);

/* Export of the built LBVH in the reference's development layout (gprt/development/gprt_lbvh.h:28-55):
 * nodes int4[2N-1] {left, right, parent, leaf primID or -1} with inner nodes [0, N-2] and leaf
 * i at N-1+i; aabbs float3[2(2N-1)] {min, max} per node; keys uint64[N] sorted
 * (morton30<<32 | primID). Buffers are resized as needed. Used by parity tests. */
GPRT_API void gprtAccelExportLBVH(GPRTContext context, GPRTAccel accel, GPRTBuffer nodes, GPRTBuffer aabbs,
                                  GPRTBuffer keys);
GPRT_API uint32_t gprtAccelGetPrimCount(GPRTAccel accel);

/* ---------------------------------------------------------------------------------------------
 * Development LBVH (gprt/development/gprt_lbvh.h:28-211, the reference's software LBVH over points,
 * edges or triangles, whose results are plain buffers). After gprtLBVHBuild:
 *   mortonCodes uint32[numPrims]   30-bit Morton codes of the key points, sorted (ties by primID)
 *   ids         uint32[numPrims]   primIDs in that order
 *   nodes       int4[2 numPrims-1] {left, right, parent, leaf primID or -1}; inner nodes [0, n-2]
 *                                  with the root at 0, leaf i at n-1+i (gprt_lbvh.h:41-51)
 *   aabbs       float3[2 (2 numPrims-1)]  {min, max} per node
 * The key point and box of a primitive follow gprt/development/gprt_lbvh.hlsl: a point itself
 * (:105-110, :228-243); an edge (uint2 into the vertices) its two end points and their midpoint
 * (p1 + p2) / 2 (:112-119, :131-137, :245-260); a triangle (uint3) its vertices and centroid
 * (p1 + p2 + p3) / 3 (:121-129, :139-146, :262-277). Vertices are float3 (12 B); the topology is
 * the reference's Karras tree over (code, primID). Create allocates the four output buffers (the
 * caller may read them with gprtBufferMap); Build enqueues the build on the context's stream;
 * Destroy frees them. The reference defines GPRT_LBVH_EDGES and the edge kernels but no create
 * function for them; gprtEdgesLBVHCreate is that missing one.
 * ------------------------------------------------------------------------------------------- */
typedef enum { GPRT_LBVH_POINTS = 1, GPRT_LBVH_EDGES = 2, GPRT_LBVH_TRIANGLES = 4 } GPRTLBVHType; /* :58-62 */
typedef struct GPRTLBVH {                                                                         /* :64-85 */
  GPRTLBVHType type;
  uint32_t numPrims;
  GPRTBuffer mortonCodes;
  GPRTBuffer ids;
  GPRTBuffer nodes;
  GPRTBuffer aabbs;
  void *impl; /* library state */
} GPRTLBVH;
GPRT_API GPRTLBVH gprtPointsLBVHCreate(GPRTContext context, GPRTBuffer vertices, uint32_t count); /* :88 */
GPRT_API GPRTLBVH gprtEdgesLBVHCreate(GPRTContext context, GPRTBuffer vertices, GPRTBuffer edges, uint32_t count);
GPRT_API GPRTLBVH gprtTriangleLBVHCreate(GPRTContext context, GPRTBuffer vertices, GPRTBuffer indices,
                                         uint32_t count);                                         /* :125 */
GPRT_API void gprtLBVHBuild(GPRTContext context, GPRTLBVH *lbvh);                                 /* :163 */
GPRT_API void gprtLBVHDestroy(GPRTLBVH *lbvh);                                                    /* :197 */
/* Root box {minX, minY, minZ, maxX, maxY, maxZ} of a built accel (for an instance accel, of its
 * top level in world space): the union of every primitive's box, as the build computed it. Waits
 * for the context's stream. The reference exposes no bounds query; this lets a caller check or
 * place a built accel without exporting its tree. */
GPRT_API void gprtAccelGetBounds(GPRTContext context, GPRTAccel accel, float *bounds);

/* Per-kernel device timing. While enabled, every kernel the library launches is bracketed by
 * HIP events on the stream it runs on; gprtKernelTimerQuery returns the summed milliseconds
 * and launch count of the kernels whose name contains `name` (after a host wait). */
GPRT_API void gprtKernelTimerEnable(GPRTContext context, bool enable);
GPRT_API void gprtKernelTimerReset(GPRTContext context);
GPRT_API double gprtKernelTimerQuery(GPRTContext context, const char *name, uint32_t *launches);
/* Writes up to `cap` bytes of "name,launches,total_ms\n" lines; returns the bytes needed. */
GPRT_API size_t gprtKernelTimerReport(GPRTContext context, char *out, size_t cap);

/* Synthetic workloads on the device (benchmarks, tests): `meshes` blob meshes, mesh i with
 * sizes[i] triangles from seed seed0 + i, written concatenated (float3 vertices, uint3 indices
 * local to each mesh) into verts / idx (resized as needed); and dst = src + U(-amplitude,
 * amplitude) per float (the dynamic scene's per-frame perturbation). Same streams as
 * gprt_amd/scenes.py. */
GPRT_API void gprtSynthBlobMeshes(GPRTContext context, const int64_t *sizes, uint32_t meshes, uint64_t seed0,
                                  float offset, GPRTBuffer verts, GPRTBuffer idx);
GPRT_API void gprtSynthPerturb(GPRTContext context, GPRTBuffer src, GPRTBuffer dst, uint64_t floats, uint64_t seed,
                               float amplitude);

/* Bounce rays for a traced batch (benchmarks, tests): `outCount` cosine-weighted rays (tmin as
 * given, tmax 1e30) from the hit points of the first `rayCount` rays of `rays` / `hits` (as
 * written by a closest-hit trace of `accel`). Misses are skipped and hits are cycled in ray
 * order; the normal is the world-space geometric normal facing the incoming ray; seeded as
 * gprt_amd/scenes.bounce_rays. For an instance accel, instanceIDs holds each hit's instance
 * (from gprtAccelTraceClosestHitDetail). With no hit at all every output ray is empty
 * (tmax < tmin). */
GPRT_API void gprtSynthBounceRays(GPRTContext context, GPRTAccel accel, GPRTBuffer rays, GPRTBuffer hits,
                                  GPRTBuffer instanceIDs, uint32_t rayCount, GPRTBuffer out, uint32_t outCount,
                                  uint64_t seed, float tmin);

/* Serialised accels (the on-disk form of a built triangle or AABB accel; no counterpart in the
 * reference, which has no serialisation). Little-endian, version 1: a 64-byte header {char
 * magic[8] = "GPRTBVH\0", u32 version = 1, u32 headerBytes = 64, u32 kind (0 triangles, 1
 * AABBs), u32 primitive count n, u32 node bytes (64), u32 record bytes (48), u32 build mode, u32
 * CRC-32 (IEEE 802.3, as zlib's crc32) of the payload, 24 reserved zero bytes}, then the inner
 * nodes (n - 1 x 64 B), the leaf records (n x 48 B), the sorted Morton codes and primIDs (n x 4 B
 * each) and the mesh header (16 floats: scene box, root box). Serialize returns the size and
 * writes it when `capacity` suffices (query with out = NULL). Deserialize loads the data into an
 * accel created over geometry of the same kind and primitive count (the geometry the data was
 * built from: leaf records hold copies of the primitives, and an update refits from the
 * geometry), which is then built: it traces, exports and, in an update mode, updates as after
 * gprtAccelBuild. Every header field and the CRC are checked; a mismatch fails through the error
 * path and leaves the accel as it was. Instance accels are not serialised (rebuild them over
 * deserialised triangle accels, as their records name accels by address). */
GPRT_API size_t gprtAccelSerialize(GPRTContext context, GPRTAccel accel, void *out, size_t capacity);
GPRT_API bool gprtAccelDeserialize(GPRTContext context, GPRTAccel accel, const void *in, size_t bytes);

/* Library version string, for logs. */
GPRT_API const char *gprtVersion(void);

#endif /* GPRT_HIP_H */
