// The development LBVH interface of the reference (gprt/development/gprt_lbvh.h:28-211): an LBVH
// over points, edges or triangles whose results are plain buffers in the reference's layout —
// sorted 30-bit Morton codes and primIDs, int4 nodes {left, right, parent, leaf primID or -1}
// (inner nodes [0, n-2], leaf i at n-1+i) and float3 box pairs per node. The reference builds it
// with six compute dispatches in the order bounds, codes, sort, make nodes, split, hierarchy
// (gprt_lbvh.h:163-195); here the accel builder produces the same tree and boxes (bounds, keys,
// segmented one-sweep sort, bottom-up hierarchy with boxes) and an export pass writes the
// buffers. The kind of primitive selects the bounds and key point exactly as the reference's
// kernels do: a point itself (gprt_lbvh.hlsl:105-110, :228-243), an edge's two end points and
// their midpoint (:112-119, :131-137, :245-260), a triangle's three vertices and their centroid
// (:121-129, :139-146, :262-277).
#include <cstring>

#include "internal.hpp"

namespace gprt {
namespace {

struct LBVHImpl {
  GeomType type{};
  Geom geom{};
  Accel acc;
};

bool lbvh_geom_ok(const LBVHImpl *im, const char *fn) {
  const Geom &g = im->geom;
  const uint32_t n = g.icount;
  if (!g.vbuf || g.vbuf->size < 12) {
    fail("%s: the vertex buffer holds no float3", fn);
    return false;
  }
  const AccelKind k = im->acc.kind;
  if (k == AccelKind::Points && (size_t)n * 12 > g.vbuf->size) {
    fail("%s: %u points exceed the vertex buffer (%zu B)", fn, n, g.vbuf->size);
    return false;
  }
  if (k != AccelKind::Points) {
    const size_t rec = k == AccelKind::Edges ? 8 : 12;
    if (!g.ibuf || (size_t)n * rec > g.ibuf->size) {
      fail("%s: %u %s exceed the index buffer (%zu B)", fn, n, k == AccelKind::Edges ? "edges" : "triangles",
           g.ibuf ? g.ibuf->size : (size_t)0);
      return false;
    }
  }
  return true;
}

GPRTLBVH lbvh_create(GPRTContext _ctx, GPRTLBVHType type, GPRTBuffer _verts, GPRTBuffer _prims, uint32_t count,
                     const char *fn) {
  GPRTLBVH out;
  std::memset(&out, 0, sizeof(out));
  Context *ctx = (Context *)_ctx;
  Buffer *verts = (Buffer *)_verts, *prims = (Buffer *)_prims;
  if (!ctx || !verts || (type != GPRT_LBVH_POINTS && !prims)) {
    fail("%s: null argument", fn);
    return out;
  }
  if (count == 0 || count >= (1u << 30)) {
    fail("%s: %u primitives (1 .. 2^30 - 1)", fn, count);
    return out;
  }
  LBVHImpl *im = new LBVHImpl();
  im->type.ctx = ctx;
  im->type.kind = type == GPRT_LBVH_TRIANGLES ? GPRT_TRIANGLES : GPRT_UNKNOWN;
  Geom &g = im->geom;
  g.ctx = ctx;
  g.type = &im->type;
  g.vbuf = verts;
  g.vstride = 12;
  g.vcount = (uint32_t)std::min<size_t>(verts->size / 12, 0xFFFFFFFFu);
  g.icount = count;
  if (type != GPRT_LBVH_POINTS) {
    g.ibuf = prims;
    g.istride = type == GPRT_LBVH_EDGES ? 8 : 12;
  }
  im->acc.ctx = ctx;
  im->acc.kind = type == GPRT_LBVH_POINTS ? AccelKind::Points
                 : type == GPRT_LBVH_EDGES ? AccelKind::Edges
                                           : AccelKind::Triangles;
  im->acc.geom = &im->geom;
  im->acc.mode = GPRT_BUILD_MODE_FAST_BUILD_NO_UPDATE;
  if (!lbvh_geom_ok(im, fn)) {
    delete im;
    return out;
  }
  const size_t nodes = 2 * (size_t)count - 1;
  out.mortonCodes = gprtDeviceBufferCreate(_ctx, 4, count, nullptr, 16);
  out.ids = gprtDeviceBufferCreate(_ctx, 4, count, nullptr, 16);
  out.nodes = gprtDeviceBufferCreate(_ctx, 16, nodes, nullptr, 16);
  out.aabbs = gprtDeviceBufferCreate(_ctx, 24, nodes, nullptr, 16);
  if (!out.mortonCodes || !out.ids || !out.nodes || !out.aabbs) {
    for (GPRTBuffer b : {out.mortonCodes, out.ids, out.nodes, out.aabbs})
      if (b) gprtBufferDestroy(b);
    delete im;
    std::memset(&out, 0, sizeof(out));
    return out;
  }
  out.type = type;
  out.numPrims = count;
  out.impl = im;
  return out;
}

}  // namespace
}  // namespace gprt

using namespace gprt;

GPRT_API GPRTLBVH gprtPointsLBVHCreate(GPRTContext context, GPRTBuffer vertices, uint32_t count) {
  return lbvh_create(context, GPRT_LBVH_POINTS, vertices, nullptr, count, "gprtPointsLBVHCreate");
}

GPRT_API GPRTLBVH gprtEdgesLBVHCreate(GPRTContext context, GPRTBuffer vertices, GPRTBuffer edges, uint32_t count) {
  return lbvh_create(context, GPRT_LBVH_EDGES, vertices, edges, count, "gprtEdgesLBVHCreate");
}

GPRT_API GPRTLBVH gprtTriangleLBVHCreate(GPRTContext context, GPRTBuffer vertices, GPRTBuffer indices,
                                         uint32_t count) {
  return lbvh_create(context, GPRT_LBVH_TRIANGLES, vertices, indices, count, "gprtTriangleLBVHCreate");
}

GPRT_API void gprtLBVHBuild(GPRTContext _ctx, GPRTLBVH *lbvh) {
  const StageRange range("gprtLBVHBuild");
  Context *ctx = (Context *)_ctx;
  if (!ctx || !lbvh || !lbvh->impl) {
    fail("gprtLBVHBuild: null argument or LBVH not created");
    return;
  }
  LBVHImpl *im = (LBVHImpl *)lbvh->impl;
  // The inputs may have been resized since the LBVH was created.
  im->geom.vcount = (uint32_t)std::min<size_t>(im->geom.vbuf->size / 12, 0xFFFFFFFFu);
  if (!lbvh_geom_ok(im, "gprtLBVHBuild")) return;
  const size_t n = lbvh->numPrims, total = 2 * n - 1;
  Buffer *codes = (Buffer *)lbvh->mortonCodes, *ids = (Buffer *)lbvh->ids, *nodes = (Buffer *)lbvh->nodes,
         *aabbs = (Buffer *)lbvh->aabbs;
  if (!codes || !ids || !nodes || !aabbs) {
    fail("gprtLBVHBuild: an output buffer is missing");
    return;
  }
  if (codes->size < n * 4 && !buffer_realloc(codes, n * 4, false)) return;
  if (ids->size < n * 4 && !buffer_realloc(ids, n * 4, false)) return;
  if (nodes->size < total * 16 && !buffer_realloc(nodes, total * 16, false)) return;
  if (aabbs->size < total * 24 && !buffer_realloc(aabbs, total * 24, false)) return;
  Accel *a = &im->acc;
  if (!lbvh_build_batch(ctx, ctx->stream, &a, 1, false)) return;
  if (!lbvh_export(ctx, ctx->stream, a, (int32_t *)nodes->d, (float *)aabbs->d, nullptr)) return;
  const BVHStore *st = a->store;
  const uint32_t off = st->desc[a->mesh].primOffset;
  hip_ok(hipMemcpyAsync(codes->d, st->codes + off, n * 4, hipMemcpyDeviceToDevice, ctx->stream),
         "hipMemcpyAsync(morton codes)");
  hip_ok(hipMemcpyAsync(ids->d, st->ids + off, n * 4, hipMemcpyDeviceToDevice, ctx->stream), "hipMemcpyAsync(ids)");
}

GPRT_API void gprtLBVHDestroy(GPRTLBVH *lbvh) {
  if (!lbvh) return;
  LBVHImpl *im = (LBVHImpl *)lbvh->impl;
  if (im) {
    hipStreamSynchronize(im->acc.ctx->stream);
    accel_detach(&im->acc);
    delete im;
  }
  for (GPRTBuffer b : {lbvh->mortonCodes, lbvh->ids, lbvh->nodes, lbvh->aabbs})
    if (b) gprtBufferDestroy(b);
  std::memset(lbvh, 0, sizeof(*lbvh));
}
