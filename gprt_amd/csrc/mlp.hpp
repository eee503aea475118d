// Treelet proposal network: sizes, parameter layout and host-side state.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace gprt {

struct Context;
struct Accel;

constexpr int kMlpBands = 8;
constexpr int kMlpIn = 6 + 6 * 2 * kMlpBands;  // 102
constexpr int kMlpInPad = 128;
constexpr int kMlpHid = 128;
constexpr int kMlpOut = 4096;
constexpr int kMlpCutDepth = 12;

// Flat fp32 parameter layout: W1t b1 W2t b2 W3t b3 W4t b4 (Wt[out][in], inputs padded to 128).
constexpr size_t kMlpW = (size_t)kMlpHid * kMlpInPad;  // 16384 per hidden layer
constexpr size_t kMlpW4 = (size_t)kMlpOut * kMlpHid;   // 524288
constexpr size_t kMlpParams = 3 * (kMlpW + kMlpHid) + kMlpW4 + kMlpOut;
constexpr size_t kMlpWeights = 3 * kMlpW + kMlpW4;    // bf16 copies
constexpr size_t kMlpBiases = 3 * kMlpHid + kMlpOut;  // contiguous bias vector used by inference

inline size_t mlp_weight_offset(int layer) { return (size_t)layer * (kMlpW + kMlpHid); }
inline size_t mlp_bias_offset(int layer) { return layer < 3 ? mlp_weight_offset(layer) + kMlpW : 3 * (kMlpW + kMlpHid) + kMlpW4; }
inline size_t mlp_weight_size(int layer) { return layer < 3 ? kMlpW : kMlpW4; }
inline size_t mlp_bf16_offset(int layer) { return (size_t)layer * kMlpW; }

struct Proposal {
  void *block = nullptr;
  size_t blockBytes = 0;
  uint32_t maxBatch = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // fp32 params, grads, Adam moments
  uint32_t *bfMap = nullptr;   // flat param index -> bf16 weight index (or ~0)
  uint16_t *wbf = nullptr;     // bf16 weights W1t W2t W3t W4t
  float *bias = nullptr;       // b1 b2 b3 b4 contiguous
  float *stats = nullptr;      // [0] loss sum, [1] correct count (as uint)
  uint16_t *X = nullptr, *H = nullptr, *dA = nullptr, *dZ = nullptr;
  float *Z = nullptr;
  int32_t *labels = nullptr;
  float *gemmWs = nullptr;  // split-K partial sums of the training GEMMs
  size_t gemmWsBytes = 0;
  uint32_t step = 0;
  // Treelets of the accel the proposal serves.
  void *treeBlock = nullptr;
  uint32_t treeCap = 0, treePrims = 0, numClasses = 0;
  int32_t *classRoot = nullptr, *innerClass = nullptr, *leafClass = nullptr, *primClass = nullptr;
  uint32_t *numClassesDev = nullptr;
  Accel *accel = nullptr;
  bool instanceBound = false;  // bound to an instance accel: classes of its top-level leaves
  // Per-ray scratch for proposal-first traces.
  void *rayScratch = nullptr;
  size_t rayScratchBytes = 0;
  // Confidence gate of proposal-first traces: rays whose top logit is below it are traced without
  // their proposal (-INFINITY: every ray uses its proposal).
  float gate = -INFINITY;
  // Overlap of proposal-first traces: the rays go in `overlapChunks` chunks, each inferred on the
  // context's second stream while the previous chunk is traversed on its main stream with
  // `overlapBlocks` traversal workgroups per CU (room for the inference on the same CUs).
  // overlapChunks < 2: inference, then one traversal.
  uint32_t overlapChunks = 0;
  uint32_t overlapBlocks = 0;
};

size_t mlp_param_count();
bool mlp_alloc(Context *ctx, Proposal *p, uint32_t maxBatch);
void mlp_release(Proposal *p);
bool mlp_set_params(Context *ctx, hipStream_t s, Proposal *p, const float *host);
bool mlp_refresh_bias(Context *ctx, hipStream_t s, Proposal *p);
void mlp_forward_batch(Context *ctx, hipStream_t s, Proposal *p, const void *rays, uint32_t n);
void mlp_backward_batch(Context *ctx, hipStream_t s, Proposal *p, uint32_t n, uint32_t labelled);
void mlp_adam(Context *ctx, hipStream_t s, Proposal *p, float lr);
void mlp_labels(Context *ctx, hipStream_t s, Proposal *p, const void *hits, uint32_t n, int32_t *labels);
// Classes of top-level leaves (instance accels): labels[i] = leafClass[leaves[i]], or -1.
void mlp_labels_of_leaves(Context *ctx, hipStream_t s, Proposal *p, const int32_t *leaves, uint32_t n,
                          int32_t *labels);
// Top-1 class per ray (and its top logit when `logit` is set); a ray whose top logit is below
// `gate` gets class -1 (-INFINITY: every ray keeps its class).
void mlp_infer(Context *ctx, hipStream_t s, Proposal *p, const void *rays, uint32_t n, int32_t *cls, float *logit,
               float gate);
bool mlp_treelets(Context *ctx, hipStream_t s, Proposal *p, Accel *a, int depth);

}  // namespace gprt
