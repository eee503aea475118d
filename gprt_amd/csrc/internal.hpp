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
#include "../../include/gprt_hip_ext.h"
#include "../../include/gprt_hip_device.h"

namespace gprt {

// Descriptor of a triangle or AABB accel in the context's pool (gprtAccelGetDeviceAddress): the
// layout device programs read (include/gprt_hip_device.h).
using BlasDesc = gprt_device::AccelDescriptor;
constexpr uint32_t kBlasMagic = gprt_device::kTriangleAccelMagic;
constexpr uint32_t kBoxAccelMagic = gprt_device::kBoxAccelMagic;

// ------------------------------------------------------------------------------------------------
// Errors (gprt/gprt.cpp:205-251: print "#gprt error:" then raise(SIGINT)).
// ------------------------------------------------------------------------------------------------
void fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
bool hip_ok(hipError_t e, const char *what);
// CRC-32 (reflected IEEE 802.3 polynomial, as zlib's crc32) of checkpoints and serialised accels.
uint32_t crc32_update(uint32_t crc, const uint8_t *p, size_t n);

// Waits for the work enqueued on `s` with the context's deadline (gprtSetSyncTimeout; 6 s by
// default, the reference's syncTDR of gprt/gprt.cpp:144): the stream is polled, and a wait that
// outlives the deadline fails through the error path instead of blocking (a hung kernel).
struct Context;
bool stream_wait(Context *ctx, hipStream_t s, const char *what);

// A named range around a C-ABI call for rocprofv3 --marker-trace (roctx).
struct StageRange {
  explicit StageRange(const char *name);
  ~StageRange();
};

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
  int tileKeys = 0;      // keys per tile: 256 threads x {8, 16}, 512 x 16 or 1024 x 16; 0: automatic (8192, or
                         // 16384 for accel builds averaging at least 2^20 primitives per mesh and for
                         // buffer sorts of at least 2^22 keys)
  int threads = 256;     // 256 or 512
  int wavesPerCU = 0;    // LBVH sort passes: occupancy cap through LDS padding (0 = none)
};

struct Context {
  SortConfig sortCfg;  // used by every sort and LBVH build issued on this context
  double syncTimeout = 6.0;  // seconds a host wait may take before it is reported (0: no limit)
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
  // Pool of triangle-accel descriptors (device, fixed capacity so addresses stay valid): the
  // address an instance record names (gprt::Instance::accelAddress) is a slot of this pool, which
  // lets the instance-accel build check every address on the device before reading through it.
  BlasDesc *blasPool = nullptr;
  std::vector<uint32_t> blasFree;
  uint32_t blasNext = 0;
  uint64_t blasEpoch = 0;  // bumped whenever a descriptor is retired or rewritten
  // Bumped whenever a triangle accel's nodes or leaf records change (build, refit, load): an
  // instance accel traces copies taken at its own build, so a proposal step that walks a
  // triangle accel's live arrays is valid only while this is unchanged since that build.
  uint64_t blasContentEpoch = 0;
  std::vector<struct Accel *> slotAccel;  // host view: pool slot -> live triangle or AABB accel
  // Second in-order stream (default priority) on which proposal inference runs ahead of the
  // traversal it feeds, and the ordering events between the two (created on first use).
  hipStream_t aux = nullptr;
  std::vector<hipEvent_t> orderEvents;
  hipStream_t aux_stream();
  hipEvent_t order_event(size_t i);
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

// Per-mesh descriptor read by the segmented LBVH kernels (device copy in BVHStore::dDesc).
struct MeshDesc {
  const char *verts;
  const char *idx;
  uint32_t vstride, nverts, istride, ntri;
  uint32_t primOffset;   // first key / leaf / triangle slot of this mesh
  uint32_t innerOffset;  // first inner node of this mesh (= primOffset - mesh index)
  uint32_t tileOffset;   // first tile of this mesh
  uint32_t ntiles;       // ceil(ntri / tileKeys)
  uint32_t kind;         // PrimKind: what a primitive is and where it is read from
  uint32_t noRecords;    // 1: the build writes no leaf records (the caller fills them afterwards)
};

// Primitive kinds of the LBVH builder. Triangles: uint3 indices at idx + i * istride into float3
// vertices. Boxes: {lo.xyz, hi.xyz} at verts + i * vstride (idx unused). Edges: uint2 indices at
// idx + i * istride. Points: the float3 vertex i itself.
enum PrimKind : uint32_t { kPrimTriangles = 0, kPrimBoxes = 1, kPrimEdges = 2, kPrimPoints = 3 };

// Device storage for the LBVHs of one or more meshes built together (one accel = one mesh).
// Every kernel of the build is launched over the tile table, so a batch of thousands of uneven
// meshes costs the same number of launches as a single mesh. Layout in bvh.hpp.
struct BVHStore {
  int refs = 0;
  uint32_t meshes = 0, prims = 0, inner = 0, tiles = 0;
  uint32_t primCap = 0, meshCap = 0, tileCap = 0;
  SortConfig cfg;    // tile size / digit width the arrays were laid out for
  int passes = 0;
  void *block = nullptr;
  size_t blockBytes = 0;
  MeshDesc *dDesc = nullptr;
  uint32_t *dTileMesh = nullptr;
  uint32_t *codes = nullptr, *ids = nullptr, *codesTmp = nullptr, *idsTmp = nullptr;
  void *nodes = nullptr;  // BVHNode[inner]
  void *tris = nullptr;   // TriRecord[prims]
  void *centroids = nullptr;  // float[prims][3]: centroids in primitive order (written by the bounds pass)
  uint32_t *parentLeaf = nullptr, *parentInner = nullptr;  // filled on demand by lbvh_parents
  float *headers = nullptr;  // [meshes][16]
  // Finished subtrees passed between the hierarchy levels (32 B each, in leaf order) and the
  // key delta before each: list0 from the 1024-key sub-tiles, list1 from the 65536-key regions.
  void *list0 = nullptr, *list1 = nullptr;      // [prims]
  void *listD0 = nullptr, *listD1 = nullptr;    // [prims] int8
  uint32_t *subCount = nullptr;     // [tiles * tileKeys / 1024]: outputs of each 1024-key sub-tile (zeroed range)
  uint32_t *regionCount = nullptr;  // [tiles]: outputs of each region, at its first tile (zeroed range)
  // Zeroed at the start of every build (one contiguous range):
  uint32_t *zeroBase = nullptr;
  size_t zeroWords = 0;
  uint32_t *hist = nullptr;      // [meshes][passes][R]
  uint32_t *status = nullptr;    // [passes][tiles][R]
  uint32_t *tickets = nullptr;   // [passes]
  // Host copies of the plan (to detect when a cached graph can be replayed): the mesh table, and
  // tile -> mesh followed by the first tile of each 65536-key region and the meshes with more
  // than one region (the device copy is dTileMesh).
  std::vector<MeshDesc> desc;
  std::vector<uint32_t> tileMesh;
  uint32_t regions = 0, bigMeshes = 0;
  std::vector<uint32_t> regionFirst;  // [meshes + 1]: each mesh's first region
  hipGraphExec_t graph = nullptr;
  bool graphValid = false;
};

// Device descriptor of a built triangle accel (64 B, one pool slot per accel, rewritten after
// every build): what an instance accel needs to traverse it and to read its triangles.
constexpr uint32_t kBlasPoolSlots = 65536;

// Triangles and AABBs are primitive accels (built by the LBVH builder); Instances is the two-level
// accel over triangle accels.
enum class AccelKind { Triangles, Instances, AABBs, Edges, Points };
struct Tlas;

struct Accel {
  Context *ctx = nullptr;
  AccelKind kind = AccelKind::Triangles;
  Geom *geom = nullptr;
  unsigned flags = 0;
  GPRTBuildMode mode = GPRT_BUILD_MODE_UNINITIALIZED;
  bool built = false;
  // The build leaves the leaf records unwritten: the owner fills them from the sorted primitive
  // ids (an instance accel's top-level tree, whose records are copies of its instances' records).
  bool recordsExternal = false;
  BVHStore *store = nullptr;  // shared by every accel of one batch build
  uint32_t mesh = 0;          // this accel's index in the store
  uint32_t n = 0;             // triangles
  // Views into the store for this mesh (used by trace / export / update).
  void *nodes = nullptr;
  void *tris = nullptr;
  float *header = nullptr;
  // Triangle accels: descriptor pool slot and the host copy last written there.
  uint32_t slot = 0xFFFFFFFFu;
  BlasDesc hdesc{};
  // Instance accels.
  Tlas *tlas = nullptr;
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
// The library's static look-back state for gprtBufferExclusiveSumNoAlloc: looked up once per
// device (with that device current) by gprtContextCreate; its device address per device.
bool scan_noalloc_init(int device);
void *scan_noalloc_state(int device);
size_t scan_noalloc_max_items();
void scan_run(Context *ctx, hipStream_t s, const uint32_t *in, uint32_t *out, size_t n, int mode, bool selectPositive,
              void *scratch, uint32_t *totalDev);

size_t sort_scratch_bytes(size_t n, int keyBits, int valBits, int loBit, int hiBit, const SortConfig &cfg);
// Sorts n keys (keyBits 32|64) with optional payload (valBits 0|32|64) by key bits [loBit, hiBit).
// keysAlt/valsAlt are ping-pong buffers of the same size; the result lands in keys/vals.
bool sort_run(Context *ctx, hipStream_t s, void *keys, void *vals, void *keysAlt, void *valsAlt, size_t n, int keyBits,
              int valBits, int loBit, int hiBit, void *scratch, const SortConfig &cfg);

// Fill a device range with a 32-bit pattern (graph-safe; no hipMemset nodes).
void fill_u32(Context *ctx, hipStream_t s, void *dst, uint32_t value, size_t words);

// Trace work counter (first 256 B) + per-lane stack spill of `spillWordsPerLane` words.
uint32_t *ensure_trace_scratch(Context *ctx, hipStream_t s, int spillWordsPerLane);

// Triangle-accel descriptors (instance.hip).
bool blas_slot_acquire(Context *ctx, Accel *a);
void blas_slot_release(Accel *a);
bool blas_desc_refresh(Context *ctx, hipStream_t s, Accel *a);

// Instance accels (instance.hip).
bool tlas_build(Context *ctx, hipStream_t s, Accel *a);
void tlas_release(Accel *a);
size_t tlas_size(const Accel *a);
uint32_t tlas_leaves(const Accel *a);
// Proposal classes of an instance accel and the class of each top-level leaf (-1: disabled).
uint32_t tlas_classes(const Accel *a);
const int32_t *tlas_leaf_class(const Accel *a);
bool tlas_trace(Context *ctx, hipStream_t s, Accel *a, const void *rays, void *hits, uint32_t count,
                const int32_t *proposal, uint32_t numClasses, int32_t *instOut, int32_t *leafOut, uint32_t *visits,
                unsigned blocksPerCU = 0);

// LBVH (segmented: one store may hold many meshes)
void store_release(BVHStore *st);
void accel_detach(Accel *a);
// Builds every accel in `accels` into one store with one launch sequence (captured into a
// hipGraph and replayed when `useGraph` and the batch shape is unchanged).
// layoutOnly: lay out the store and the accels' views without building (deserialisation).
bool lbvh_build_batch(Context *ctx, hipStream_t s, Accel **accels, uint32_t count, bool useGraph,
                      bool layoutOnly = false);
bool lbvh_refit(Context *ctx, hipStream_t s, Accel *a);
bool lbvh_export(Context *ctx, hipStream_t s, Accel *a, int32_t *nodes, float *aabbs, uint64_t *keys);
// Fills the store's parent links of one built triangle accel from its nodes.
void lbvh_parents(Context *ctx, hipStream_t s, Accel *a);

// Trace
bool trace_closest(Context *ctx, hipStream_t s, Accel *a, const void *rays, void *hits, uint32_t count);
// proposal / classRoot / visits may be null. proposal[i] = treelet class of ray i (or -1).
// blocksPerCU: persistent workgroups per CU (0: the default); fewer leave room for kernels of
// another stream on the same CUs.
bool trace_closest_ex(Context *ctx, hipStream_t s, Accel *a, const void *rays, void *hits, uint32_t count,
                      const int32_t *proposal, const int32_t *classRoot, uint32_t numClasses, uint32_t *visits,
                      unsigned blocksPerCU = 0);

}  // namespace gprt
