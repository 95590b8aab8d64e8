// Neural-field sampler: multiresolution hash-grid encoding + 32-64-64-1 ReLU
// MLP, trained on samples of the exact sampler, with per-leaf error bounds and
// exact fallback. Device-side model description and kernel launchers.
#pragma once

#include <hip/hip_runtime_api.h>

#include "vkl_views.h"

namespace vklhip {

constexpr int kNfLevels   = 16;
constexpr int kNfFeatures = 2;
constexpr int kNfLog2T    = 19;
constexpr uint32_t kNfT   = 1u << kNfLog2T;
constexpr int kNfIn       = kNfLevels * kNfFeatures;  // 32
constexpr int kNfHidden   = 64;
constexpr size_t kNfTableEntries = (size_t)kNfLevels * kNfT * kNfFeatures;
constexpr int kNfW1 = kNfIn * kNfHidden;      // 2048
constexpr int kNfW2 = kNfHidden * kNfHidden;  // 4096
constexpr int kNfW3 = kNfHidden;              // 64

// binary16 copies read by the kernels (written by the optimizer).
struct NfWeights
{
  const uint16_t *table;  // [levels][T][2]
  const uint16_t *W1;     // [32 in][64 hidden]
  const uint16_t *W1T;    // [64 hidden][32 in]
  const uint16_t *W2;     // [64 in][64 out]
  const uint16_t *W2T;    // [64 out][64 in]
  const float *W3;        // [64]
};

struct NfModel
{
  NfWeights w;
  uint32_t res[kNfLevels];
  float o2i[12];
  float lo[3];       // index-space box lower corner
  float invExtent;   // 1 / edge of the (cubic) normalisation box
};

// fp32 gradient accumulators (unscaled sums of per-sample gradients).
struct NfGrads
{
  float *table, *W1, *W2, *W3;
  float *loss;  // sum of squared errors
};

// Training samples: index-space points drawn in leaves (fraction leafFrac)
// or uniformly in the box; written as object-space points.
hipError_t launchNfGenerate(const float *i2o, const float *boxLo, const float *boxHi, const int32_t *leafOrigins,
                            uint64_t numLeaves, float leafFrac, uint64_t n, uint64_t seed, float *xyz,
                            hipStream_t stream);
hipError_t launchNfTrainStep(const NfModel &m, uint64_t n, const float *xyz, const float *target, const NfGrads &g,
                             hipStream_t stream);
// Adam on fp32 masters; refreshes the binary16 copies; zeroes the gradients.
struct NfAdam
{
  float *table32, *W1_32, *W2_32, *W3_32;
  float *mTable, *vTable, *mW, *vW;  // mW / vW: W1 | W2 | W3 packed
  uint16_t *table16, *W1h, *W1Th, *W2h, *W2Th;
  float *W3f;
};
hipError_t launchNfAdam(const NfAdam &a, const NfGrads &g, float lrTable, float lrMlp, float beta1, float beta2,
                        float eps, uint32_t step, float gradScale, hipStream_t stream);
hipError_t launchNfInit(const NfAdam &a, uint64_t seed, hipStream_t stream);

// Inference. Pass 1 evaluates the network for every point; points whose cell
// leaf is not approved (or not a dense leaf) are appended to `fallback`
// (count in *numFallback). Pass 2 (exact) is the caller's stream sampler.
// clampBand: clamp network values in approved leaves to the leaf value range.
hipError_t launchNfInfer(const NfModel &m, const GridView &g, uint32_t attr, const uint8_t *leafOk, uint64_t n,
                         const float *xyz, float *out, uint32_t *fallback, uint32_t *numFallback, int useFallback,
                         int clampBand, hipStream_t stream);
// Exact samples at the fallback points, scattered to their slots.
hipError_t launchGatherSample(const SamplerView &s, const AttrView &a, uint32_t attr, const uint32_t *idx,
                              const uint32_t *count, uint64_t maxCount, const float *xyz, float *out,
                              hipStream_t stream);

// Per-leaf maximum |network - exact| over points drawn in each leaf.
hipError_t launchNfLeafPoints(const float *i2o, const int32_t *leafOrigins, uint64_t numLeaves, uint32_t perLeaf,
                              uint64_t seed, float *xyz, hipStream_t stream);
hipError_t launchNfLeafError(const GridView &g, const float *nf, const float *exact, uint64_t numLeaves,
                             uint32_t perLeaf, float *leafErr, hipStream_t stream);
hipError_t launchNfApprove(const GridView &g, const int32_t *leafOrigins, const float *leafErr, uint64_t numLeaves,
                           uint32_t attr, float tolerance, float safety, int useBand, uint8_t *leafOk,
                           uint32_t *numOk, hipStream_t stream);

}  // namespace vklhip
