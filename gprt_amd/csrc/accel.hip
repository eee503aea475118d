// C-ABI of the triangle geometry / accel family: gprtGeomTypeCreate, gprtGeomCreate,
// gprtTrianglesSetVertices/Indices, gprtTriangleAccelCreate, gprtAccelBuild/Update/Destroy
// (gprt/gprt_host.h:406-498, :1308, :1421-1557; gprt/gprt.cpp:3149-3575, :9152-9262) and the
// ray-batch trace and LBVH export extensions.
#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "bvh.hpp"
#include "internal.hpp"

using namespace gprt;

GPRT_API GPRTGeomType gprtGeomTypeCreate(GPRTContext _ctx, GPRTGeomKind kind, size_t recordSize) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) {
    fail("gprtGeomTypeCreate: null context");
    return nullptr;
  }
  if (kind != GPRT_TRIANGLES && kind != GPRT_AABBS) {
    fail("gprtGeomTypeCreate: only GPRT_TRIANGLES and GPRT_AABBS geometry have a software accel in this build "
         "(kind %d)", (int)kind);
    return nullptr;
  }
  GeomType *t = new GeomType{ctx, kind, recordSize};
  return (GPRTGeomType)t;
}

GPRT_API void gprtGeomTypeDestroy(GPRTGeomType t) { delete (GeomType *)t; }

GPRT_API GPRTGeom gprtGeomCreate(GPRTContext _ctx, GPRTGeomType _type) {
  Context *ctx = (Context *)_ctx;
  GeomType *type = (GeomType *)_type;
  if (!ctx || !type) {
    fail("gprtGeomCreate: null argument");
    return nullptr;
  }
  Geom *g = new Geom();
  g->ctx = ctx;
  g->type = type;
  return (GPRTGeom)g;
}

GPRT_API void gprtGeomDestroy(GPRTGeom g) { delete (Geom *)g; }

GPRT_API void gprtTrianglesSetVertices(GPRTGeom _g, GPRTBuffer _v, uint32_t count, uint32_t stride, uint32_t offset) {
  Geom *g = (Geom *)_g;
  Buffer *v = (Buffer *)_v;
  if (!g || !v) {
    fail("gprtTrianglesSetVertices: null argument");
    return;
  }
  if (stride < 12 || (stride & 3) || (offset & 3)) {
    fail("gprtTrianglesSetVertices: stride %u / offset %u must be 4-byte aligned with stride >= 12", stride, offset);
    return;
  }
  if (count && (size_t)offset + (size_t)(count - 1) * stride + 12 > v->size) {
    fail("gprtTrianglesSetVertices: %u vertices (stride %u, offset %u) exceed the buffer (%zu B)", count, stride,
         offset, v->size);
    return;
  }
  g->vbuf = v;
  g->vcount = count;
  g->vstride = stride;
  g->voffset = offset;
}

GPRT_API void gprtTrianglesSetIndices(GPRTGeom _g, GPRTBuffer _i, uint32_t count, uint32_t stride, uint32_t offset) {
  Geom *g = (Geom *)_g;
  Buffer *ib = (Buffer *)_i;
  if (!g || !ib) {
    fail("gprtTrianglesSetIndices: null argument");
    return;
  }
  if (stride < 12 || (stride & 3) || (offset & 3)) {
    fail("gprtTrianglesSetIndices: stride %u / offset %u must be 4-byte aligned with stride >= 12", stride, offset);
    return;
  }
  if (count && (size_t)offset + (size_t)(count - 1) * stride + 12 > ib->size) {
    fail("gprtTrianglesSetIndices: %u triangles (stride %u, offset %u) exceed the buffer (%zu B)", count, stride,
         offset, ib->size);
    return;
  }
  g->ibuf = ib;
  g->icount = count;
  g->istride = stride;
  g->ioffset = offset;
}

// Box primitives (gprt/gprt_host.h:614): count boxes {minX, minY, minZ, maxX, maxY, maxZ} at
// offset + i * stride. They take the vertex slots of the geometry; the primitive count is count.
GPRT_API void gprtAABBsSetPositions(GPRTGeom _g, GPRTBuffer _p, uint32_t count, uint32_t stride, uint32_t offset) {
  Geom *g = (Geom *)_g;
  Buffer *pb = (Buffer *)_p;
  if (!g || !pb) {
    fail("gprtAABBsSetPositions: null argument");
    return;
  }
  if (g->type->kind != GPRT_AABBS) {
    fail("gprtAABBsSetPositions: geometry is not GPRT_AABBS");
    return;
  }
  if (stride < 24 || (stride & 3) || (offset & 3)) {
    fail("gprtAABBsSetPositions: stride %u / offset %u must be 4-byte aligned with stride >= 24", stride, offset);
    return;
  }
  if (count && (size_t)offset + (size_t)(count - 1) * stride + 24 > pb->size) {
    fail("gprtAABBsSetPositions: %u boxes (stride %u, offset %u) exceed the buffer (%zu B)", count, stride, offset,
         pb->size);
    return;
  }
  g->vbuf = pb;
  g->vcount = count;
  g->vstride = stride;
  g->voffset = offset;
  g->icount = count;
}

// An accel over box primitives (gprt/gprt_host.h:1282). It is built by the same LBVH builder; its
// closest hit is the box surface (boxprim_hit), standing in for the user intersection program
// that the reference's AABB geometry runs. Its descriptor slot (what device programs trace
// through) carries the box magic word, which instance records do not accept.
GPRT_API GPRTAccel gprtAABBAccelCreate(GPRTContext _ctx, GPRTGeom _g, unsigned int flags) {
  Context *ctx = (Context *)_ctx;
  Geom *g = (Geom *)_g;
  if (!ctx || !g) {
    fail("gprtAABBAccelCreate: null argument");
    return nullptr;
  }
  if (g->type->kind != GPRT_AABBS) {
    fail("gprtAABBAccelCreate: geometry is not GPRT_AABBS");
    return nullptr;
  }
  Accel *a = new Accel();
  a->ctx = ctx;
  a->kind = AccelKind::AABBs;
  a->geom = g;
  a->flags = flags;
  if (!blas_slot_acquire(ctx, a)) {
    delete a;
    return nullptr;
  }
  return (GPRTAccel)a;
}

GPRT_API GPRTAccel gprtTriangleAccelCreate(GPRTContext _ctx, GPRTGeom _g, unsigned int flags) {
  Context *ctx = (Context *)_ctx;
  Geom *g = (Geom *)_g;
  if (!ctx || !g) {
    fail("gprtTriangleAccelCreate: null argument");
    return nullptr;
  }
  if (g->type->kind != GPRT_TRIANGLES) {
    fail("gprtTriangleAccelCreate: geometry is not GPRT_TRIANGLES");
    return nullptr;
  }
  Accel *a = new Accel();
  a->ctx = ctx;
  a->geom = g;
  a->flags = flags;
  if (!blas_slot_acquire(ctx, a)) {
    delete a;
    return nullptr;
  }
  return (GPRTAccel)a;
}

GPRT_API void gprtAccelDestroy(GPRTAccel _a) {
  Accel *a = (Accel *)_a;
  if (!a) return;
  hipStreamSynchronize(a->ctx->stream);
  if (a->kind == AccelKind::Instances) {
    tlas_release(a);
  } else {
    accel_detach(a);
    blas_slot_release(a);
  }
  delete a;
}

static bool validate_geom(Geom *g, const char *fn) {
  if (g->type->kind == GPRT_AABBS) {
    if (!g->vbuf || g->icount == 0) {
      fail("%s: box geometry has no boxes (gprtAABBsSetPositions)", fn);
      return false;
    }
    if (g->icount >= (1u << 30)) {
      fail("%s: %u boxes exceeds the 2^30 limit", fn, g->icount);
      return false;
    }
    // The buffer may have been resized since gprtAABBsSetPositions checked it.
    if ((size_t)g->voffset + (size_t)(g->icount - 1) * g->vstride + 24 > g->vbuf->size) {
      fail("%s: %u boxes (stride %u, offset %u) exceed the box buffer (%zu B)", fn, g->icount, g->vstride,
           g->voffset, g->vbuf->size);
      return false;
    }
    return true;
  }
  if (!g->vbuf || !g->ibuf) {
    fail("%s: triangle geometry needs vertices and indices", fn);
    return false;
  }
  if (g->icount == 0 || g->vcount == 0) {
    fail("%s: geometry has no triangles or no vertices", fn);
    return false;
  }
  if (g->icount >= (1u << 30)) {
    fail("%s: %u triangles exceeds the 2^30 limit", fn, g->icount);
    return false;
  }
  // The buffers may have been resized since gprtTrianglesSetVertices / SetIndices checked them;
  // the kernels read every index triple and clamp vertex indices to vcount.
  if ((size_t)g->voffset + (size_t)(g->vcount - 1) * g->vstride + 12 > g->vbuf->size) {
    fail("%s: %u vertices (stride %u, offset %u) exceed the vertex buffer (%zu B)", fn, g->vcount, g->vstride,
         g->voffset, g->vbuf->size);
    return false;
  }
  if ((size_t)g->ioffset + (size_t)(g->icount - 1) * g->istride + 12 > g->ibuf->size) {
    fail("%s: %u triangles (stride %u, offset %u) exceed the index buffer (%zu B)", fn, g->icount, g->istride,
         g->ioffset, g->ibuf->size);
    return false;
  }
  return true;
}

GPRT_API void gprtAccelBuild(GPRTContext _ctx, GPRTAccel _a, GPRTBuildParams params) {
  const StageRange range("gprtAccelBuild");
  Context *ctx = (Context *)_ctx;
  Accel *a = (Accel *)_a;
  if (!ctx || !a) {
    fail("gprtAccelBuild: null argument");
    return;
  }
  if (params.buildMode == GPRT_BUILD_MODE_UNINITIALIZED) params.buildMode = GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE;
  if (a->kind == AccelKind::Instances) {
    a->mode = params.buildMode;
    a->built = tlas_build(ctx, ctx->stream, a);
    return;
  }
  if (!validate_geom(a->geom, "gprtAccelBuild")) return;
  a->mode = params.buildMode;
  // A single build detaches the accel from any batch store it shared.
  if (a->store && a->store->refs > 1) accel_detach(a);
  lbvh_build_batch(ctx, ctx->stream, &a, 1, false);
}

// ------------------------------------------------------------------------------------------------
// Serialised accels: a 64-byte header, then the built arrays of one triangle or AABB accel (inner
// nodes, leaf records, sorted keys, mesh header) as the library lays them out.
// ------------------------------------------------------------------------------------------------
namespace {

struct AccelFileHeader {
  char magic[8];
  uint32_t version, headerBytes;
  uint32_t kind;  // 0 triangles, 1 AABBs
  uint32_t prims;
  uint32_t nodeBytes, recordBytes;
  uint32_t buildMode;
  uint32_t crc;
  uint32_t reserved[6];
};
static_assert(sizeof(AccelFileHeader) == 64, "serialised accel header is 64 bytes");
constexpr char kAccelMagic[8] = {'G', 'P', 'R', 'T', 'B', 'V', 'H', '\0'};
constexpr uint32_t kAccelVersion = 1;

// Payload pieces of an accel of n primitives, in file order: nodes, records, codes, ids, header.
struct AccelPieces {
  size_t bytes[5];
  size_t total() const { return bytes[0] + bytes[1] + bytes[2] + bytes[3] + bytes[4]; }
};
AccelPieces accel_pieces(uint32_t n) {
  return {{(size_t)(n - 1) * sizeof(BVHNode), (size_t)n * sizeof(TriRecord), (size_t)n * 4, (size_t)n * 4, 16 * 4}};
}

void *accel_piece(Accel *a, int k) {
  const MeshDesc &d = a->store->desc[a->mesh];
  switch (k) {
    case 0: return a->nodes;
    case 1: return a->tris;
    case 2: return a->store->codes + d.primOffset;
    case 3: return a->store->ids + d.primOffset;
    default: return a->header;
  }
}

}  // namespace

GPRT_API size_t gprtAccelSerialize(GPRTContext _ctx, GPRTAccel _a, void *out, size_t capacity) {
  const StageRange range("gprtAccelSerialize");
  Context *ctx = (Context *)_ctx;
  Accel *a = (Accel *)_a;
  if (!ctx || !a) {
    fail("gprtAccelSerialize: null argument");
    return 0;
  }
  if (a->kind == AccelKind::Instances) {
    fail("gprtAccelSerialize: instance accels are not serialised (serialise their triangle accels)");
    return 0;
  }
  if (!a->built || !a->store) {
    fail("gprtAccelSerialize: the accel is not built");
    return 0;
  }
  const AccelPieces pc = accel_pieces(a->n);
  const size_t bytes = sizeof(AccelFileHeader) + pc.total();
  if (!out || capacity < bytes) return bytes;
  uint8_t *dst = (uint8_t *)out + sizeof(AccelFileHeader);
  for (int k = 0; k < 5; ++k) {
    if (pc.bytes[k] && !hip_ok(hipMemcpyAsync(dst, accel_piece(a, k), pc.bytes[k], hipMemcpyDeviceToHost, ctx->stream),
                               "hipMemcpyAsync(serialise)"))
      return 0;
    dst += pc.bytes[k];
  }
  if (!hip_ok(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize(serialise)")) return 0;
  AccelFileHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kAccelMagic, 8);
  h.version = kAccelVersion;
  h.headerBytes = sizeof(AccelFileHeader);
  h.kind = a->kind == AccelKind::AABBs ? 1u : 0u;
  h.prims = a->n;
  h.nodeBytes = sizeof(BVHNode);
  h.recordBytes = sizeof(TriRecord);
  h.buildMode = (uint32_t)a->mode;
  h.crc = crc32_update(0u, (const uint8_t *)out + sizeof(AccelFileHeader), pc.total());
  std::memcpy(out, &h, sizeof(h));
  return bytes;
}

GPRT_API bool gprtAccelDeserialize(GPRTContext _ctx, GPRTAccel _a, const void *in, size_t bytes) {
  const StageRange range("gprtAccelDeserialize");
  Context *ctx = (Context *)_ctx;
  Accel *a = (Accel *)_a;
  if (!ctx || !a || !in) {
    fail("gprtAccelDeserialize: null argument");
    return false;
  }
  if (a->kind == AccelKind::Instances) {
    fail("gprtAccelDeserialize: instance accels are not serialised (build them over deserialised triangle accels)");
    return false;
  }
  if (bytes < sizeof(AccelFileHeader)) {
    fail("gprtAccelDeserialize: %zu bytes is shorter than the 64-byte header", bytes);
    return false;
  }
  AccelFileHeader h;
  std::memcpy(&h, in, sizeof(h));
  if (std::memcmp(h.magic, kAccelMagic, 8) != 0) {
    fail("gprtAccelDeserialize: not a serialised accel (bad magic)");
    return false;
  }
  if (h.version != kAccelVersion || h.headerBytes != sizeof(AccelFileHeader) || h.nodeBytes != sizeof(BVHNode) ||
      h.recordBytes != sizeof(TriRecord)) {
    fail("gprtAccelDeserialize: unsupported version %u (header %u, node %u, record %u bytes)", h.version, h.headerBytes,
         h.nodeBytes, h.recordBytes);
    return false;
  }
  if (h.buildMode > (uint32_t)GPRT_BUILD_MODE_FAST_TRACE_AND_UPDATE) {
    fail("gprtAccelDeserialize: unknown build mode %u", h.buildMode);
    return false;
  }
  const uint32_t kind = a->kind == AccelKind::AABBs ? 1u : 0u;
  if (h.kind != kind) {
    fail("gprtAccelDeserialize: the data is a %s accel, the target a %s accel", h.kind ? "AABB" : "triangle",
         kind ? "AABB" : "triangle");
    return false;
  }
  if (!validate_geom(a->geom, "gprtAccelDeserialize")) return false;
  if (h.prims == 0 || h.prims != a->geom->icount) {
    fail("gprtAccelDeserialize: the data holds %u primitives, the accel's geometry %u", h.prims, a->geom->icount);
    return false;
  }
  const AccelPieces pc = accel_pieces(h.prims);
  if (bytes != sizeof(AccelFileHeader) + pc.total()) {
    fail("gprtAccelDeserialize: %zu bytes, expected %zu", bytes, sizeof(AccelFileHeader) + pc.total());
    return false;
  }
  const uint8_t *src = (const uint8_t *)in + sizeof(AccelFileHeader);
  if (crc32_update(0u, src, pc.total()) != h.crc) {
    fail("gprtAccelDeserialize: checksum mismatch (the data is damaged)");
    return false;
  }
  if (a->store && a->store->refs > 1) accel_detach(a);
  if (!lbvh_build_batch(ctx, ctx->stream, &a, 1, false, true)) return false;
  for (int k = 0; k < 5; ++k) {
    if (pc.bytes[k] && !hip_ok(hipMemcpyAsync(accel_piece(a, k), src, pc.bytes[k], hipMemcpyHostToDevice, ctx->stream),
                               "hipMemcpyAsync(deserialise)"))
      return false;
    src += pc.bytes[k];
  }
  if (!hip_ok(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize(deserialise)")) return false;
  a->mode = (GPRTBuildMode)h.buildMode;
  a->built = true;
  return true;
}

// The accels of a batch: each non-null, a triangle or box accel, named once (a batch store counts
// one reference per accel), over valid geometry.
static bool collect_batch(GPRTAccel *_accels, uint32_t count, std::vector<Accel *> &accels, const char *fn) {
  accels.resize(count);
  std::unordered_set<Accel *> seen;
  for (uint32_t i = 0; i < count; ++i) {
    accels[i] = (Accel *)_accels[i];
    if (!accels[i]) {
      fail("%s: accel %u is null", fn, i);
      return false;
    }
    if (!seen.insert(accels[i]).second) {
      fail("%s: accel %u appears twice", fn, i);
      return false;
    }
    if (accels[i]->kind == AccelKind::Instances) {
      fail("%s: accel %u is an instance accel", fn, i);
      return false;
    }
    if (!validate_geom(accels[i]->geom, fn)) return false;
  }
  return true;
}

GPRT_API void gprtAccelBuildBatch(GPRTContext _ctx, GPRTAccel *_accels, uint32_t count, GPRTBuildParams params) {
  const StageRange range("gprtAccelBuildBatch");
  Context *ctx = (Context *)_ctx;
  if (!ctx || !_accels) {
    fail("gprtAccelBuildBatch: null argument");
    return;
  }
  if (count == 0) return;
  std::vector<Accel *> accels;
  if (!collect_batch(_accels, count, accels, "gprtAccelBuildBatch")) return;
  if (params.buildMode == GPRT_BUILD_MODE_UNINITIALIZED) params.buildMode = GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE;
  for (auto *a : accels) a->mode = params.buildMode;
  lbvh_build_batch(ctx, ctx->stream, accels.data(), count, true);
}

GPRT_API void gprtAccelUpdate(GPRTContext _ctx, GPRTAccel _a) {
  const StageRange range("gprtAccelUpdate");
  Context *ctx = (Context *)_ctx;
  Accel *a = (Accel *)_a;
  if (!ctx || !a) {
    fail("gprtAccelUpdate: null argument");
    return;
  }
  if (!a->built) {
    fail("gprtAccelUpdate: accel has not been built");
    return;
  }
  // gprt/gprt.cpp:3404-3408: update requires an *_AND_UPDATE build mode.
  if (a->mode != GPRT_BUILD_MODE_FAST_BUILD_AND_UPDATE && a->mode != GPRT_BUILD_MODE_FAST_TRACE_AND_UPDATE) {
    fail("Attempting to update an accel that was not built with an *_AND_UPDATE build mode");
    return;
  }
  if (a->kind == AccelKind::Instances) {
    // The top level is rebuilt from the records and the current sub-root boxes.
    a->built = tlas_build(ctx, ctx->stream, a);
    return;
  }
  if (!validate_geom(a->geom, "gprtAccelUpdate")) return;
  if (a->geom->icount != a->n) {
    fail("gprtAccelUpdate: primitive count changed (%u -> %u); rebuild instead", a->n, a->geom->icount);
    return;
  }
  lbvh_refit(ctx, ctx->stream, a);
}

GPRT_API size_t gprtAccelGetSize(GPRTAccel _a, int) {
  Accel *a = (Accel *)_a;
  if (a && a->kind == AccelKind::Instances) return tlas_size(a);
  // Nodes + leaf triangle records of this accel (the traversal-resident part).
  return (a && a->built) ? (size_t)(a->n > 1 ? a->n - 1 : 1) * sizeof(BVHNode) + (size_t)a->n * sizeof(TriRecord)
                         : 0;
}

// A triangle or AABB accel's address is its descriptor slot (what gprt::Instance::accelAddress
// names, and what device programs trace through); it stays the same for the accel's lifetime. An
// instance accel's is its top-level node array.
GPRT_API const void *gprtAccelGetDeviceAddress(GPRTAccel _a, int) {
  Accel *a = (Accel *)_a;
  if (!a) return nullptr;
  if (a->kind == AccelKind::Instances || a->slot == 0xFFFFFFFFu) return a->nodes;
  return a->ctx->blasPool ? (const void *)(a->ctx->blasPool + a->slot) : nullptr;
}

GPRT_API uint32_t gprtAccelGetPrimCount(GPRTAccel _a) {
  Accel *a = (Accel *)_a;
  return a ? a->n : 0;
}

// The root box lives in the accel's 16-float header (floats 8..13), written by the build's
// hierarchy pass (or by the top level of an instance accel's build).
GPRT_API void gprtAccelGetBounds(GPRTContext _ctx, GPRTAccel _a, float *bounds) {
  Context *ctx = (Context *)_ctx;
  Accel *a = (Accel *)_a;
  if (!ctx || !a || !bounds) {
    fail("gprtAccelGetBounds: null argument");
    return;
  }
  if (!a->built || !a->header) {
    fail("gprtAccelGetBounds: accel has not been built");
    return;
  }
  float b[6];
  if (!hip_ok(hipMemcpyAsync(b, a->header + 8, sizeof(b), hipMemcpyDeviceToHost, ctx->stream),
              "hipMemcpyAsync(root box)"))
    return;
  if (!stream_wait(ctx, ctx->stream, "gprtAccelGetBounds")) return;
  std::memcpy(bounds, b, sizeof(b));
}

GPRT_API void gprtAccelTraceClosestHit(GPRTContext _ctx, GPRTAccel _a, GPRTBuffer _rays, GPRTBuffer _hits,
                                       uint32_t count) {
  const StageRange range("gprtAccelTraceClosestHit");
  Context *ctx = (Context *)_ctx;
  Accel *a = (Accel *)_a;
  Buffer *rays = (Buffer *)_rays, *hits = (Buffer *)_hits;
  if (!ctx || !a || !rays || !hits) {
    fail("gprtAccelTraceClosestHit: null argument");
    return;
  }
  if (!a->built) {
    fail("gprtAccelTraceClosestHit: accel has not been built");
    return;
  }
  if ((size_t)count * sizeof(GPRTRay) > rays->size) {
    fail("gprtAccelTraceClosestHit: %u rays exceed the ray buffer (%zu B)", count, rays->size);
    return;
  }
  if ((size_t)count * sizeof(GPRTHit) > hits->size) {
    fail("gprtAccelTraceClosestHit: %u hits exceed the hit buffer (%zu B)", count, hits->size);
    return;
  }
  if (a->kind == AccelKind::Instances)
    tlas_trace(ctx, ctx->stream, a, rays->d, hits->d, count, nullptr, 0, nullptr, nullptr, nullptr);
  else
    trace_closest(ctx, ctx->stream, a, rays->d, hits->d, count);
}

GPRT_API void gprtAccelExportLBVH(GPRTContext _ctx, GPRTAccel _a, GPRTBuffer _nodes, GPRTBuffer _aabbs,
                                  GPRTBuffer _keys) {
  Context *ctx = (Context *)_ctx;
  Accel *a = (Accel *)_a;
  Buffer *nodes = (Buffer *)_nodes, *aabbs = (Buffer *)_aabbs, *keys = (Buffer *)_keys;
  if (!ctx || !a || !nodes || !aabbs || !keys) {
    fail("gprtAccelExportLBVH: null argument");
    return;
  }
  if (!a->built) {
    fail("gprtAccelExportLBVH: accel has not been built");
    return;
  }
  if (a->kind == AccelKind::Instances) {
    fail("gprtAccelExportLBVH: expects a triangle or box accel");
    return;
  }
  const size_t n = a->n;
  const size_t total = 2 * n - 1;
  if (nodes->size < total * 16 && !buffer_realloc(nodes, total * 16, false)) return;
  if (aabbs->size < total * 24 && !buffer_realloc(aabbs, total * 24, false)) return;
  if (keys->size < n * 8 && !buffer_realloc(keys, n * 8, false)) return;
  lbvh_export(ctx, ctx->stream, a, (int32_t *)nodes->d, (float *)aabbs->d, (uint64_t *)keys->d);
}

// Autotuner for batched builds: times the graph-captured batch build of `accels` for every
// digit width {4, 6, 8} x tile {2048, 4096, 8192} keys x sort-pass occupancy {none, 8, 16} waves
// per CU, and the 16384-key tile (one workgroup per CU, no cap) per digit width: 30 configurations
// (median of `reps` graph launches, HIP events on the context's stream), keeps the
// fastest configuration in the context and leaves the accels built with it. Results do not depend
// on the configuration. Writes a JSON report into `report` (up to `cap` bytes); returns the index
// of the best configuration.
GPRT_API uint32_t gprtAccelBuildBatchAutotune(GPRTContext _ctx, GPRTAccel *_accels, uint32_t count, uint32_t reps,
                                              char *report, size_t cap) {
  const StageRange range("gprtAccelBuildBatchAutotune");
  Context *ctx = (Context *)_ctx;
  if (!ctx || !_accels || count == 0) {
    fail("gprtAccelBuildBatchAutotune: null argument or empty batch");
    return 0;
  }
  if (reps == 0) reps = 3;
  std::vector<Accel *> accels;
  if (!collect_batch(_accels, count, accels, "gprtAccelBuildBatchAutotune")) return 0;
  // An accel keeps the build mode of its last build (default parameters if it has none).
  for (auto *a : accels)
    if (a->mode == GPRT_BUILD_MODE_UNINITIALIZED) a->mode = GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE;
  const int bits[3] = {4, 6, 8}, tiles[4] = {2048, 4096, 8192, 16384}, waves[3] = {0, 8, 16};
  const bool timer = ctx->timer.enabled;
  ctx->timer.enabled = false;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  std::string js = "{\"configs\": [";
  float bestMs = 1e30f;
  uint32_t best = 0, idx = 0;
  SortConfig bestCfg = ctx->sortCfg;
  for (int b : bits)
    for (int t : tiles)
      for (int w : waves) {
        if (t == 16384 && w != 0) continue;  // one workgroup per CU already: a cap changes nothing
        SortConfig cfg = ctx->sortCfg;
        cfg.digitBits = b;
        cfg.tileKeys = t;
        cfg.wavesPerCU = w;
        ctx->sortCfg = cfg;
        if (!lbvh_build_batch(ctx, ctx->stream, accels.data(), count, true)) break;  // capture + warm-up
        std::vector<float> ms(reps);
        for (uint32_t r = 0; r < reps; ++r) {
          hipEventRecord(e0, ctx->stream);
          lbvh_build_batch(ctx, ctx->stream, accels.data(), count, true);
          hipEventRecord(e1, ctx->stream);
          hipEventSynchronize(e1);
          hipEventElapsedTime(&ms[r], e0, e1);
        }
        std::sort(ms.begin(), ms.end());
        const float med = ms[reps / 2];
        char line[160];
        // %.9g gives the float back exactly, so a reader's argmin agrees with the strict < below
        snprintf(line, sizeof(line), "%s{\"digitBits\": %d, \"tileKeys\": %d, \"wavesPerCU\": %d, \"ms\": %.9g}",
                 idx ? ", " : "", b, t, w, med);
        js += line;
        if (med < bestMs) {
          bestMs = med;
          best = idx;
          bestCfg = cfg;
        }
        ++idx;
      }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  ctx->sortCfg = bestCfg;
  lbvh_build_batch(ctx, ctx->stream, accels.data(), count, true);
  ctx->timer.enabled = timer;
  char tail[200];
  snprintf(tail, sizeof(tail), "], \"best\": {\"digitBits\": %d, \"tileKeys\": %d, \"wavesPerCU\": %d, \"ms\": %.9g}}",
           bestCfg.digitBits, bestCfg.tileKeys, bestCfg.wavesPerCU, bestMs);
  js += tail;
  if (report && cap) {
    const size_t n = std::min(js.size(), cap - 1);
    std::memcpy(report, js.data(), n);
    report[n] = 0;
  }
  return best;
}
