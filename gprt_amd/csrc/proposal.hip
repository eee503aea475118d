// C-ABI of the learned treelet proposal: create / bind to an accel (treelet cut) / train on
// plain-traversal hits / infer / proposal-first trace, plus parameter and gradient access for
// tests. None of these has a reference counterpart (see the extension block of gprt_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh.hpp"
#include "device_common.hpp"
#include "internal.hpp"
#include "mlp.hpp"

using namespace gprt;

namespace {

// A label counts when it names one of the `limit` classes.
__global__ void __launch_bounds__(256) k_label_flags(const int32_t *__restrict__ labels, uint32_t n, uint32_t limit,
                                                     uint32_t *__restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = labels[i] >= 0 && (uint32_t)labels[i] < limit ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_compact(const GPRTRay *__restrict__ rays, const int32_t *__restrict__ labels,
                                                 const uint32_t *__restrict__ flags,
                                                 const uint32_t *__restrict__ offs, uint32_t n,
                                                 GPRTRay *__restrict__ outRays, int32_t *__restrict__ outLabels) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flags[i]) return;
  const uint32_t o = offs[i];
  outRays[o] = rays[i];
  outLabels[o] = labels[i];
}

// Gate calibration: each ray's top logit becomes a sort key that orders logits from the largest
// down (its order-preserving bits inverted; -0 taken as +0, which compares equal to it, and NaN as
// -infinity, which no threshold keeps), and its node-fetch difference (proposal-first minus plain,
// two's complement) the payload.
__global__ void __launch_bounds__(256) k_gate_keys(uint32_t *__restrict__ logitKey, uint32_t *__restrict__ v0Delta,
                                                   const uint32_t *__restrict__ v1, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f = __uint_as_float(logitKey[i]);
  if (f != f) f = -INFINITY;
  if (f == 0.0f) f = 0.0f;
  logitKey[i] = ~f2ord(f);
  v0Delta[i] = v1[i] - v0Delta[i];
}

__global__ void __launch_bounds__(256) k_inst_of_hits(const GPRTHit *__restrict__ hits, uint32_t n,
                                                      int32_t *__restrict__ inst) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inst[i] = hits[i].primID >= 0 ? 0 : -1;
}

void instance_ids_of_hits(Context *ctx, hipStream_t s, const GPRTHit *hits, uint32_t n, int32_t *inst) {
  launch(ctx, s, "trace_instance_ids", k_inst_of_hits, dim3((n + 255) / 256), dim3(256), 0, hits, n, inst);
}

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// He-uniform initialisation: W ~ U(-sqrt(6/fan_in), sqrt(6/fan_in)), biases 0; padded input
// columns of the first layer stay 0. Deterministic in `seed`.
std::vector<float> init_params(uint64_t seed) {
  std::vector<float> w(kMlpParams, 0.f);
  uint64_t ctr = mix64(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull);
  for (int layer = 0; layer < 4; ++layer) {
    const int fanIn = layer == 0 ? kMlpIn : kMlpHid;
    const int rows = layer < 3 ? kMlpHid : kMlpOut;
    const float lim = std::sqrt(6.0f / (float)fanIn);
    float *W = w.data() + mlp_weight_offset(layer);
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < fanIn; ++c) {
        const uint64_t u = mix64(ctr += 0x9E3779B97F4A7C15ull);
        const float x = (float)((u >> 40) * (1.0 / 16777216.0));
        W[(size_t)r * (layer == 0 ? kMlpInPad : kMlpHid) + c] = (2.0f * x - 1.0f) * lim;
      }
  }
  return w;
}

bool ensure_ray_scratch(Proposal *p, size_t bytes, hipStream_t s) {
  if (bytes <= p->rayScratchBytes) return true;
  if (p->rayScratch) {
    hipStreamSynchronize(s);
    hipFree(p->rayScratch);
    p->rayScratch = nullptr;
    p->rayScratchBytes = 0;
  }
  if (!hip_ok(hipMalloc(&p->rayScratch, bytes), "hipMalloc(proposal rays)")) return false;
  p->rayScratchBytes = bytes;
  return true;
}

}  // namespace

GPRT_API GPRTProposal gprtProposalCreate(GPRTContext _ctx, uint32_t maxBatch, uint64_t seed) {
  Context *ctx = (Context *)_ctx;
  if (!ctx) {
    fail("gprtProposalCreate: null context");
    return nullptr;
  }
  if (maxBatch < 64 || maxBatch > (1u << 16)) {
    fail("gprtProposalCreate: maxBatch %u outside [64, 65536]", maxBatch);
    return nullptr;
  }
  Proposal *p = new Proposal();
  if (!mlp_alloc(ctx, p, maxBatch)) {
    delete p;
    return nullptr;
  }
  std::vector<float> w = init_params(seed);
  if (!mlp_set_params(ctx, ctx->stream, p, w.data())) {
    mlp_release(p);
    delete p;
    return nullptr;
  }
  hipStreamSynchronize(ctx->stream);
  return (GPRTProposal)p;
}

GPRT_API void gprtProposalDestroy(GPRTContext _ctx, GPRTProposal _p) {
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  if (!p) return;
  if (ctx) hipStreamSynchronize(ctx->stream);
  if (p->rayScratch) hipFree(p->rayScratch);
  mlp_release(p);
  delete p;
}

GPRT_API uint32_t gprtProposalBindAccel(GPRTContext _ctx, GPRTProposal _p, GPRTAccel _a, uint32_t cutDepth) {
  const StageRange range("gprtProposalBindAccel");
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  Accel *a = (Accel *)_a;
  if (!ctx || !p || !a) {
    fail("gprtProposalBindAccel: null argument");
    return 0;
  }
  if (!a->built) {
    fail("gprtProposalBindAccel: accel has not been built");
    return 0;
  }
  if (cutDepth > (uint32_t)kMlpCutDepth) {
    fail("gprtProposalBindAccel: cutDepth %u exceeds %d (at most 4096 treelets)", cutDepth, kMlpCutDepth);
    return 0;
  }
  if (a->kind == AccelKind::Instances) {
    // Classes are fixed groups of each instance's triangles (stable across rebuilds); cutDepth is
    // unused.
    p->accel = a;
    p->instanceBound = true;
    p->numClasses = tlas_classes(a);
    p->treePrims = a->n;
    return p->numClasses;
  }
  p->instanceBound = false;
  if (!mlp_treelets(ctx, ctx->stream, p, a, (int)cutDepth)) return 0;
  return p->numClasses;
}

GPRT_API void gprtProposalSetParams(GPRTContext _ctx, GPRTProposal _p, const float *host, size_t count) {
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  if (!ctx || !p || !host) {
    fail("gprtProposalSetParams: null argument");
    return;
  }
  if (count != kMlpParams) {
    fail("gprtProposalSetParams: expected %zu parameters, got %zu", (size_t)kMlpParams, count);
    return;
  }
  mlp_set_params(ctx, ctx->stream, p, host);
  hipStreamSynchronize(ctx->stream);
}

GPRT_API void gprtProposalGetParams(GPRTContext _ctx, GPRTProposal _p, float *host, size_t count) {
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  if (!ctx || !p || !host) {
    fail("gprtProposalGetParams: null argument");
    return;
  }
  if (count != kMlpParams) {
    fail("gprtProposalGetParams: expected %zu parameters, got %zu", (size_t)kMlpParams, count);
    return;
  }
  hip_ok(hipMemcpyAsync(host, p->w, kMlpParams * 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync(params)");
  hipStreamSynchronize(ctx->stream);
}

GPRT_API size_t gprtProposalParamCount(void) { return kMlpParams; }

// ------------------------------------------------------------------------------------------------
// Weight checkpoints: a 64-byte header, the fp32 parameters and optionally the Adam moments.
// ------------------------------------------------------------------------------------------------
namespace {

struct CheckpointHeader {
  char magic[8];
  uint32_t version, headerBytes;
  uint64_t params;
  uint32_t in, hidden, layers, out;
  uint32_t step, flags, crc;
  uint32_t reserved[3];
};
static_assert(sizeof(CheckpointHeader) == 64, "checkpoint header is 64 bytes");
constexpr char kCheckpointMagic[8] = {'G', 'P', 'R', 'T', 'M', 'L', 'P', '\0'};
constexpr uint32_t kCheckpointVersion = 1;

}  // namespace

GPRT_API size_t gprtProposalSerialize(GPRTContext _ctx, GPRTProposal _p, void *out, size_t capacity,
                                      bool withOptimizerState) {
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  if (!ctx || !p) {
    fail("gprtProposalSerialize: null argument");
    return 0;
  }
  const size_t arrays = withOptimizerState ? 3 : 1;
  const size_t bytes = sizeof(CheckpointHeader) + arrays * kMlpParams * 4;
  if (!out || capacity < bytes) return bytes;
  CheckpointHeader h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.magic, kCheckpointMagic, 8);
  h.version = kCheckpointVersion;
  h.headerBytes = sizeof(CheckpointHeader);
  h.params = kMlpParams;
  h.in = kMlpIn;
  h.hidden = kMlpHid;
  h.layers = 3;
  h.out = kMlpOut;
  h.step = p->step;
  h.flags = withOptimizerState ? 1u : 0u;
  uint8_t *payload = (uint8_t *)out + sizeof(CheckpointHeader);
  const float *src[3] = {p->w, p->m, p->v};
  for (size_t k = 0; k < arrays; ++k)
    if (!hip_ok(hipMemcpyAsync(payload + k * kMlpParams * 4, src[k], kMlpParams * 4, hipMemcpyDeviceToHost,
                               ctx->stream),
                "hipMemcpyAsync(checkpoint)"))
      return 0;
  if (!hip_ok(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize(checkpoint)")) return 0;
  h.crc = crc32_update(0u, payload, arrays * kMlpParams * 4);
  std::memcpy(out, &h, sizeof(h));
  return bytes;
}

GPRT_API bool gprtProposalDeserialize(GPRTContext _ctx, GPRTProposal _p, const void *in, size_t bytes) {
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  if (!ctx || !p || !in) {
    fail("gprtProposalDeserialize: null argument");
    return false;
  }
  if (bytes < sizeof(CheckpointHeader)) {
    fail("gprtProposalDeserialize: %zu bytes is shorter than the 64-byte header", bytes);
    return false;
  }
  CheckpointHeader h;
  std::memcpy(&h, in, sizeof(h));
  if (std::memcmp(h.magic, kCheckpointMagic, 8) != 0) {
    fail("gprtProposalDeserialize: not a proposal checkpoint (bad magic)");
    return false;
  }
  if (h.version != kCheckpointVersion || h.headerBytes != sizeof(CheckpointHeader)) {
    fail("gprtProposalDeserialize: unsupported checkpoint version %u (header %u bytes)", h.version, h.headerBytes);
    return false;
  }
  if (h.params != kMlpParams || h.in != (uint32_t)kMlpIn || h.hidden != (uint32_t)kMlpHid || h.layers != 3 ||
      h.out != (uint32_t)kMlpOut || (h.flags & ~1u) != 0u) {
    fail("gprtProposalDeserialize: network shape %llu params (%u-%u x%u-%u), flags %u does not match this build's "
         "%zu params (%d-%d x3-%d)",
         (unsigned long long)h.params, h.in, h.hidden, h.layers, h.out, h.flags, kMlpParams, kMlpIn, kMlpHid, kMlpOut);
    return false;
  }
  const size_t arrays = (h.flags & 1u) ? 3 : 1;
  const size_t payloadBytes = arrays * kMlpParams * 4;
  if (bytes != sizeof(CheckpointHeader) + payloadBytes) {
    fail("gprtProposalDeserialize: %zu bytes, expected %zu", bytes, sizeof(CheckpointHeader) + payloadBytes);
    return false;
  }
  const uint8_t *payload = (const uint8_t *)in + sizeof(CheckpointHeader);
  if (crc32_update(0u, payload, payloadBytes) != h.crc) {
    fail("gprtProposalDeserialize: checksum mismatch (the checkpoint is damaged)");
    return false;
  }
  if (!mlp_set_params(ctx, ctx->stream, p, (const float *)payload)) return false;
  if (arrays == 3) {
    if (!hip_ok(hipMemcpyAsync(p->m, payload + kMlpParams * 4, kMlpParams * 4, hipMemcpyHostToDevice, ctx->stream),
                "hipMemcpyAsync(checkpoint m)") ||
        !hip_ok(hipMemcpyAsync(p->v, payload + 2 * kMlpParams * 4, kMlpParams * 4, hipMemcpyHostToDevice,
                               ctx->stream),
                "hipMemcpyAsync(checkpoint v)"))
      return false;
    p->step = h.step;
  }
  return hip_ok(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize(checkpoint)");
}

GPRT_API void gprtProposalLogits(GPRTContext _ctx, GPRTProposal _p, GPRTBuffer _rays, uint32_t count,
                                 GPRTBuffer _logits) {
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  Buffer *rays = (Buffer *)_rays, *logits = (Buffer *)_logits;
  if (!ctx || !p || !rays || !logits) {
    fail("gprtProposalLogits: null argument");
    return;
  }
  if (count > p->maxBatch || (size_t)count * sizeof(GPRTRay) > rays->size ||
      (size_t)count * kMlpOut * 4 > logits->size) {
    fail("gprtProposalLogits: count %u exceeds the batch (%u) or the buffers", count, p->maxBatch);
    return;
  }
  mlp_forward_batch(ctx, ctx->stream, p, rays->d, count);
  hip_ok(hipMemcpyAsync(logits->d, p->Z, (size_t)count * kMlpOut * 4, hipMemcpyDeviceToDevice, ctx->stream),
         "hipMemcpyAsync(logits)");
}

GPRT_API void gprtProposalGradients(GPRTContext _ctx, GPRTProposal _p, GPRTBuffer _rays, GPRTBuffer _labels,
                                    uint32_t count, float *hostGrads, size_t gradCount) {
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  Buffer *rays = (Buffer *)_rays, *labels = (Buffer *)_labels;
  if (!ctx || !p || !rays || !labels || !hostGrads) {
    fail("gprtProposalGradients: null argument");
    return;
  }
  if (gradCount != kMlpParams || count == 0 || count > p->maxBatch ||
      (size_t)count * sizeof(GPRTRay) > rays->size || (size_t)count * 4 > labels->size) {
    fail("gprtProposalGradients: bad sizes (count %u, batch %u, grads %zu)", count, p->maxBatch, gradCount);
    return;
  }
  hipStream_t s = ctx->stream;
  hip_ok(hipMemcpyAsync(p->labels, labels->d, (size_t)count * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
  fill_u32(ctx, s, p->stats, 0u, 2);
  mlp_forward_batch(ctx, s, p, rays->d, count);
  mlp_backward_batch(ctx, s, p, count, count);
  hip_ok(hipMemcpyAsync(hostGrads, p->g, kMlpParams * 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(grads)");
  hipStreamSynchronize(s);
}

namespace {

// Labels come from plain-trace hits (triangle accels: the treelet of the hit primitive) or are
// given per ray; unlabelled rays are compacted away, then `epochs` passes of minibatch Adam run
// over the rest. Returns the last epoch's mean loss.
float train_core(Context *ctx, Proposal *p, const void *rays, const void *hits, const int32_t *userLabels,
                 uint32_t count, uint32_t epochs, float lr, float *accuracy) {
  hipStream_t s = ctx->stream;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  // Scratch: labels | flags | offsets | compacted rays | compacted labels | scan state
  const size_t need = al((size_t)count * 4) * 4 + al((size_t)count * sizeof(GPRTRay)) + al(scan_scratch_bytes(count));
  if (!ensure_ray_scratch(p, need, s)) return NAN;
  char *b = (char *)p->rayScratch;
  int32_t *labels = (int32_t *)b;
  uint32_t *flags = (uint32_t *)(b + al((size_t)count * 4));
  uint32_t *offs = (uint32_t *)(b + 2 * al((size_t)count * 4));
  int32_t *cLabels = (int32_t *)(b + 3 * al((size_t)count * 4));
  GPRTRay *cRays = (GPRTRay *)(b + 4 * al((size_t)count * 4));
  void *scanState = b + 4 * al((size_t)count * 4) + al((size_t)count * sizeof(GPRTRay));
  const unsigned g = (count + 255) / 256;
  if (hits)
    mlp_labels(ctx, s, p, hits, count, labels);
  else if (p->instanceBound)
    mlp_labels_of_leaves(ctx, s, p, userLabels, count, labels);  // top-level leaf -> its treelet
  else
    hip_ok(hipMemcpyAsync(labels, userLabels, (size_t)count * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
  launch(ctx, s, "mlp_label_flags", k_label_flags, dim3(g), dim3(256), 0, (const int32_t *)labels, count,
         p->numClasses ? p->numClasses : (uint32_t)kMlpOut, flags);
  scan_run(ctx, s, flags, offs, count, 0, false, scanState, nullptr);
  launch(ctx, s, "mlp_compact", k_compact, dim3(g), dim3(256), 0, (const GPRTRay *)rays, (const int32_t *)labels,
         (const uint32_t *)flags, (const uint32_t *)offs, count, cRays, cLabels);
  uint32_t labelled = 0;
  hip_ok(hipMemcpyAsync(&labelled, scanState, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(labelled)");
  hipStreamSynchronize(s);
  float lastLoss = 0.f, acc = 0.f;
  for (uint32_t e = 0; e < epochs; ++e) {
    fill_u32(ctx, s, p->stats, 0u, 2);
    uint32_t batches = 0;
    for (uint32_t b0 = 0; b0 < labelled; b0 += p->maxBatch) {
      const uint32_t n = std::min<uint32_t>(p->maxBatch, labelled - b0);
      hip_ok(hipMemcpyAsync(p->labels, cLabels + b0, (size_t)n * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
      mlp_forward_batch(ctx, s, p, cRays + b0, n);
      mlp_backward_batch(ctx, s, p, n, n);
      mlp_adam(ctx, s, p, lr);
      ++batches;
    }
    float st[2];
    hip_ok(hipMemcpyAsync(st, p->stats, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync(stats)");
    hipStreamSynchronize(s);
    uint32_t correct;
    std::memcpy(&correct, &st[1], 4);
    lastLoss = batches ? st[0] / (float)batches : 0.f;
    acc = labelled ? (float)correct / (float)labelled : 0.f;
  }
  if (accuracy) *accuracy = acc;
  return lastLoss;
}


}  // namespace

GPRT_API float gprtProposalTrain(GPRTContext _ctx, GPRTProposal _p, GPRTBuffer _rays, GPRTBuffer _hits,
                                 uint32_t count, uint32_t epochs, float lr, float *accuracy) {
  const StageRange range("gprtProposalTrain");
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  Buffer *rays = (Buffer *)_rays, *hits = (Buffer *)_hits;
  if (!ctx || !p || !rays || !hits) {
    fail("gprtProposalTrain: null argument");
    return NAN;
  }
  if (!p->accel) {
    fail("gprtProposalTrain: bind the proposal to an accel first (gprtProposalBindAccel)");
    return NAN;
  }
  if (p->instanceBound) {
    fail("gprtProposalTrain: hits of an instance accel do not name their top-level leaf; use "
         "gprtProposalTrainLabels with the leafIDs of gprtAccelTraceClosestHitDetail");
    return NAN;
  }
  if ((size_t)count * sizeof(GPRTRay) > rays->size || (size_t)count * sizeof(GPRTHit) > hits->size) {
    fail("gprtProposalTrain: %u rays exceed the ray or hit buffer", count);
    return NAN;
  }
  if (count == 0) return 0.f;
  return train_core(ctx, p, rays->d, hits->d, nullptr, count, epochs, lr, accuracy);
}

GPRT_API float gprtProposalTrainLabels(GPRTContext _ctx, GPRTProposal _p, GPRTBuffer _rays, GPRTBuffer _labels,
                                       uint32_t count, uint32_t epochs, float lr, float *accuracy) {
  const StageRange range("gprtProposalTrainLabels");
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  Buffer *rays = (Buffer *)_rays, *labels = (Buffer *)_labels;
  if (!ctx || !p || !rays || !labels) {
    fail("gprtProposalTrainLabels: null argument");
    return NAN;
  }
  if ((size_t)count * sizeof(GPRTRay) > rays->size || (size_t)count * 4 > labels->size) {
    fail("gprtProposalTrainLabels: %u rays exceed the ray or label buffer", count);
    return NAN;
  }
  if (count == 0) return 0.f;
  return train_core(ctx, p, rays->d, nullptr, (const int32_t *)labels->d, count, epochs, lr, accuracy);
}

GPRT_API void gprtProposalInfer(GPRTContext _ctx, GPRTProposal _p, GPRTBuffer _rays, uint32_t count,
                                GPRTBuffer _classes) {
  const StageRange range("gprtProposalInfer");
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  Buffer *rays = (Buffer *)_rays, *cls = (Buffer *)_classes;
  if (!ctx || !p || !rays || !cls) {
    fail("gprtProposalInfer: null argument");
    return;
  }
  if ((size_t)count * sizeof(GPRTRay) > rays->size || (size_t)count * 4 > cls->size) {
    fail("gprtProposalInfer: %u rays exceed the ray or class buffer", count);
    return;
  }
  if (count == 0) return;
  mlp_infer(ctx, ctx->stream, p, rays->d, count, (int32_t *)cls->d, nullptr, -INFINITY);
}

GPRT_API void gprtProposalExportTreelets(GPRTContext _ctx, GPRTProposal _p, GPRTBuffer _classRoot,
                                         GPRTBuffer _primClass) {
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  Buffer *cr = (Buffer *)_classRoot, *pc = (Buffer *)_primClass;
  if (!ctx || !p || !cr || !pc || !p->accel) {
    fail("gprtProposalExportTreelets: null argument or unbound proposal");
    return;
  }
  if (p->instanceBound) {
    fail("gprtProposalExportTreelets: bound to an instance accel, whose classes are groups of its instances' "
         "triangles");
    return;
  }

  if (cr->size < (size_t)p->numClasses * 4 && !buffer_realloc(cr, (size_t)p->numClasses * 4, false)) return;
  if (pc->size < (size_t)p->treePrims * 4 && !buffer_realloc(pc, (size_t)p->treePrims * 4, false)) return;
  hip_ok(hipMemcpyAsync(cr->d, p->classRoot, (size_t)p->numClasses * 4, hipMemcpyDeviceToDevice, ctx->stream),
         "hipMemcpyAsync");
  hip_ok(hipMemcpyAsync(pc->d, p->primClass, (size_t)p->treePrims * 4, hipMemcpyDeviceToDevice, ctx->stream),
         "hipMemcpyAsync");
}

namespace {

void trace_detail(const char *fn, Context *ctx, Accel *a, Proposal *p, Buffer *rays, Buffer *hits, uint32_t count,
                  Buffer *inst, Buffer *leaf, Buffer *visits) {
  if (!ctx || !a || !rays || !hits) {
    fail("%s: null argument", fn);
    return;
  }
  if (!a->built) {
    fail("%s: accel has not been built", fn);
    return;
  }
  if ((size_t)count * sizeof(GPRTRay) > rays->size || (size_t)count * sizeof(GPRTHit) > hits->size ||
      (visits && (size_t)count * 4 > visits->size) || (inst && (size_t)count * 4 > inst->size) ||
      (leaf && (size_t)count * 4 > leaf->size)) {
    fail("%s: %u rays exceed a buffer", fn, count);
    return;
  }
  if (p && p->accel != a) {
    fail("%s: the proposal is bound to another accel", fn);
    return;
  }
  if (p && p->treePrims != a->n) {
    fail("%s: the accel was rebuilt with a different size; bind the proposal again", fn);
    return;
  }
  if (count == 0) return;
  // A gate of +infinity admits no ray (every top logit is below it), so the classes would all be
  // discarded: the trace runs without the proposal and without its inference. Hits are the same.
  if (p && p->gate == INFINITY) p = nullptr;
  hipStream_t s = ctx->stream;
  int32_t *cls = nullptr;
  uint32_t *vis = visits ? (uint32_t *)visits->d : nullptr;
  // One traversal of rays [r0, r0 + m) with their classes.
  auto traverse_range = [&](uint32_t r0, uint32_t m, unsigned blocksPerCU) {
    const GPRTRay *r = (const GPRTRay *)rays->d + r0;
    GPRTHit *h = (GPRTHit *)hits->d + r0;
    const int32_t *c = cls ? cls + r0 : nullptr;
    uint32_t *v = vis ? vis + r0 : nullptr;
    if (a->kind == AccelKind::Instances)
      tlas_trace(ctx, s, a, r, h, m, c, p ? p->numClasses : 0, inst ? (int32_t *)inst->d + r0 : nullptr,
                 leaf ? (int32_t *)leaf->d + r0 : nullptr, v, blocksPerCU);
    else
      trace_closest_ex(ctx, s, a, r, h, m, c, p ? p->classRoot : nullptr, p ? p->numClasses : 0, v, blocksPerCU);
  };
  const uint32_t chunks = p ? std::min<uint32_t>(p->overlapChunks, (count + 65535) / 65536) : 0;
  if (p) {
    if (!ensure_ray_scratch(p, (size_t)count * 4, s)) return;
    cls = (int32_t *)p->rayScratch;
  }
  hipStream_t s2 = chunks >= 2 ? ctx->aux_stream() : nullptr;
  if (chunks >= 2 && s2 && ctx->order_event(chunks)) {
    // Chunk k is inferred on the second stream while chunk k - 1 is traversed on the main stream;
    // the second stream starts after the work already enqueued on the main one, and the main
    // stream waits for each chunk's classes before traversing it (so everything later on the main
    // stream is ordered after all of it).
    const uint32_t step = ((count + chunks - 1) / chunks + 255) / 256 * 256;
    hipEvent_t start = ctx->order_event(chunks);
    hip_ok(hipEventRecord(start, s), "hipEventRecord");
    hip_ok(hipStreamWaitEvent(s2, start, 0), "hipStreamWaitEvent");
    for (uint32_t k = 0, r0 = 0; r0 < count; ++k, r0 += step) {
      const uint32_t m = std::min(step, count - r0);
      mlp_infer(ctx, s2, p, (const GPRTRay *)rays->d + r0, m, cls + r0, nullptr, p->gate);
      hip_ok(hipEventRecord(ctx->order_event(k), s2), "hipEventRecord");
    }
    for (uint32_t k = 0, r0 = 0; r0 < count; ++k, r0 += step) {
      hip_ok(hipStreamWaitEvent(s, ctx->order_event(k), 0), "hipStreamWaitEvent");
      traverse_range(r0, std::min(step, count - r0), p->overlapBlocks);
    }
  } else {
    if (p) mlp_infer(ctx, s, p, rays->d, count, cls, nullptr, p->gate);
    traverse_range(0, count, 0);
  }
  if (a->kind == AccelKind::Instances) return;
  if (inst || leaf) {
    // Triangle accel: instance 0 on a hit, no top-level leaf.
    Buffer *hb = hits;
    if (inst) instance_ids_of_hits(ctx, s, (const GPRTHit *)hb->d, count, (int32_t *)inst->d);
    if (leaf) fill_u32(ctx, s, leaf->d, 0xFFFFFFFFu, count);
  }
}

}  // namespace

GPRT_API void gprtProposalSetOverlap(GPRTContext _ctx, GPRTProposal _p, uint32_t chunks, uint32_t traceBlocksPerCU) {
  Proposal *p = (Proposal *)_p;
  if (!_ctx || !p) {
    fail("gprtProposalSetOverlap: null argument");
    return;
  }
  if (chunks > 64) {
    fail("gprtProposalSetOverlap: %u chunks (at most 64)", chunks);
    return;
  }
  p->overlapChunks = chunks;
  p->overlapBlocks = traceBlocksPerCU;
}

GPRT_API void gprtProposalSetGate(GPRTContext _ctx, GPRTProposal _p, float threshold) {
  Proposal *p = (Proposal *)_p;
  if (!_ctx || !p) {
    fail("gprtProposalSetGate: null argument");
    return;
  }
  if (std::isnan(threshold)) {
    fail("gprtProposalSetGate: the threshold is NaN");
    return;
  }
  p->gate = threshold;
}

GPRT_API float gprtProposalGetGate(GPRTProposal _p) {
  Proposal *p = (Proposal *)_p;
  return p ? p->gate : NAN;
}

// The gate from measured traversal cost: for every ray of the sample the inner nodes fetched
// plainly and proposal-first (ungated) are counted; proposing to the rays whose top logit is at
// least t costs sum over those rays of (proposal-first - plain) fetches, and the threshold with
// the smallest total is kept (none proposed costs 0, so the gate never makes the sample dearer).
GPRT_API float gprtProposalCalibrateGate(GPRTContext _ctx, GPRTProposal _p, GPRTAccel _a, GPRTBuffer _rays,
                                         uint32_t count) {
  const StageRange range("gprtProposalCalibrateGate");
  const char *fn = "gprtProposalCalibrateGate";
  Context *ctx = (Context *)_ctx;
  Proposal *p = (Proposal *)_p;
  Accel *a = (Accel *)_a;
  Buffer *rays = (Buffer *)_rays;
  if (!ctx || !p || !a || !rays) {
    fail("%s: null argument", fn);
    return NAN;
  }
  if (!a->built || p->accel != a || p->treePrims != a->n) {
    fail("%s: the proposal is not bound to this built accel (bind it after the build)", fn);
    return NAN;
  }
  if ((size_t)count * sizeof(GPRTRay) > rays->size) {
    fail("%s: %u rays exceed the ray buffer", fn, count);
    return NAN;
  }
  if (count == 0) return 0.f;
  hipStream_t s = ctx->stream;
  const size_t n = count;
  const SortConfig cfg = ctx->sortCfg;
  const size_t sortBytes = sort_scratch_bytes(n, 32, 32, 0, 32, cfg);
  const size_t main = (n * (4 + 4 + 4 + 4 + sizeof(GPRTHit)) + 255) & ~(size_t)255;
  char *blk = nullptr;
  if (!hip_ok(hipMalloc((void **)&blk, main + sortBytes), "hipMalloc(calibration)")) return NAN;
  int32_t *cls = (int32_t *)blk;
  uint32_t *logit = (uint32_t *)(blk + 4 * n);  // the top logits, then their sort keys
  uint32_t *v0 = (uint32_t *)(blk + 8 * n), *v1 = (uint32_t *)(blk + 12 * n);  // v0: then the differences
  char *hits = blk + 16 * n;  // the traces' hits, then the sort's ping-pong buffers
  mlp_infer(ctx, s, p, rays->d, count, cls, (float *)logit, -INFINITY);
  if (a->kind == AccelKind::Instances) {
    tlas_trace(ctx, s, a, rays->d, hits, count, nullptr, 0, nullptr, nullptr, v0);
    tlas_trace(ctx, s, a, rays->d, hits, count, cls, p->numClasses, nullptr, nullptr, v1);
  } else {
    trace_closest_ex(ctx, s, a, rays->d, hits, count, nullptr, nullptr, 0, v0);
    trace_closest_ex(ctx, s, a, rays->d, hits, count, cls, p->classRoot, p->numClasses, v1);
  }
  // Rays from the largest top logit down, with their fetch differences, sorted on the device.
  launch(ctx, s, "proposal_gate_keys", k_gate_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, logit, v0,
         (const uint32_t *)v1, (uint32_t)n);
  std::vector<uint32_t> key(n), delta(n);
  bool ok = sort_run(ctx, s, logit, v0, hits, hits + 4 * n, n, 32, 32, 0, 32, blk + main, cfg) &&
            hip_ok(hipMemcpyAsync(key.data(), logit, 4 * n, hipMemcpyDeviceToHost, s), "hipMemcpyAsync") &&
            hip_ok(hipMemcpyAsync(delta.data(), v0, 4 * n, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  // Whatever failed, the work already enqueued ends before its memory is freed.
  ok = stream_wait(ctx, s, fn) && ok;
  hipFree(blk);
  if (!ok) return NAN;
  // Proposing to the rays with the k largest logits costs the sum of their first k differences;
  // the cheapest k wins. A cut between equal logits cannot be expressed by a threshold.
  int64_t cum = 0, best = 0;
  size_t keep = 0;
  for (size_t k = 0; k < n; ++k) {
    cum += (int32_t)delta[k];
    if (cum < best && (k + 1 == n || key[k + 1] != key[k])) {
      best = cum;
      keep = k + 1;
    }
  }
  float gate = -INFINITY;
  if (keep == 0) {
    gate = INFINITY;
  } else if (keep < n) {
    const uint32_t u = ~key[keep - 1];  // back from the order-preserving bits to the logit
    const uint32_t bits = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    std::memcpy(&gate, &bits, 4);
  }
  p->gate = gate;
  return (float)keep / (float)n;
}

GPRT_API void gprtAccelTraceClosestHitEx(GPRTContext _ctx, GPRTAccel _a, GPRTProposal _p, GPRTBuffer _rays,
                                         GPRTBuffer _hits, uint32_t count, GPRTBuffer _visits) {
  const StageRange range("gprtAccelTraceClosestHitEx");
  trace_detail("gprtAccelTraceClosestHitEx", (Context *)_ctx, (Accel *)_a, (Proposal *)_p, (Buffer *)_rays,
               (Buffer *)_hits, count, nullptr, nullptr, (Buffer *)_visits);
}

GPRT_API void gprtAccelTraceClosestHitDetail(GPRTContext _ctx, GPRTAccel _a, GPRTProposal _p, GPRTBuffer _rays,
                                             GPRTBuffer _hits, uint32_t count, GPRTBuffer _inst, GPRTBuffer _leaf,
                                             GPRTBuffer _visits) {
  const StageRange range("gprtAccelTraceClosestHitDetail");
  trace_detail("gprtAccelTraceClosestHitDetail", (Context *)_ctx, (Accel *)_a, (Proposal *)_p, (Buffer *)_rays,
               (Buffer *)_hits, count, (Buffer *)_inst, (Buffer *)_leaf, (Buffer *)_visits);
}
