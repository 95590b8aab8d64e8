// Neural-field sampler objects: parameters and optimizer state in HBM,
// training loop (samples drawn from the exact sampler), per-leaf error bounds
// and the two-pass inference with exact fallback.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

#include "kernels.h"
#include "neural.h"
#include "objects.h"
#include "openvkl/openvkl_hip.h"

namespace vklhip {

struct NeuralField
{
  Device *device   = nullptr;
  Sampler *sampler = nullptr;  // exact sampler (retained)
  uint32_t attr    = 0;
  NfModel model{};
  NfAdam adam{};
  NfGrads grads{};
  DeviceBuffer table16, W1h, W1Th, W2h, W2Th, W3f;
  DeviceBuffer table32, W1_32, W2_32, W3_32, mTable, vTable, mW, vW;
  DeviceBuffer gTable, gW1, gW2, gW3, gLoss;
  DeviceBuffer leafOrigins, leafOk, leafErr, i2o, boxLo, boxHi;
  uint64_t numLeaves = 0;
  uint32_t step      = 0;
  uint32_t numOk     = 0;
  float tolerance    = 0.f;
  bool finalized     = false;
  // Nearest / linear exact samples lie in their leaf's value range, so the
  // network output may be clamped to it (not so for the cubic filter).
  bool band = false;

  ~NeuralField()
  {
    if (sampler) sampler->release();
  }
};

namespace {

template <typename T>
T *alloc(DeviceBuffer &b, size_t count)
{
  b.allocate(count * sizeof(T));
  return b.as<T>();
}

void zero(DeviceBuffer &b, hipStream_t st) { checkHip(hipMemsetAsync(b.ptr, 0, b.bytes, st), "hipMemsetAsync"); }

NeuralField *nfOf(VKLNeuralField h)
{
  if (!h) throw VklError("null neural field");
  return reinterpret_cast<NeuralField *>(h);
}

}  // namespace
}  // namespace vklhip

using namespace vklhip;

#define NF_TRY(expr)     \
  Device *dev_ = nullptr; \
  try {                  \
    dev_ = (expr);
#define NF_CATCH(ret)                                                   \
  }                                                                     \
  catch (const std::exception &e)                                       \
  {                                                                     \
    if (dev_)                                                           \
      dev_->handleError(VKL_UNKNOWN_ERROR, e.what());                   \
    else                                                                \
      std::fprintf(stderr, "OPENVKL ERROR: %s\n", e.what());            \
    return ret;                                                         \
  }

extern "C" {

OPENVKL_INTERFACE VKLNeuralField vklNewNeuralField(const VKLSampler *exactSampler, unsigned int attributeIndex,
                                                   unsigned int maxResolution, unsigned int seed)
{
  NF_TRY(exactSampler && exactSampler->host ? static_cast<Sampler *>(exactSampler->host)->device : nullptr)
  if (!dev_) throw VklError("null sampler");
  Sampler *s = static_cast<Sampler *>(exactSampler->host);
  Volume *v  = s->volume;
  if (v->grid.dense) throw VklError("neural fields are built for vdb volumes");
  if (v->grid.temporalFormat) throw VklError("neural fields are built for temporally constant vdb volumes");
  (void)s->attr(attributeIndex);
  dev_->makeCurrent();
  hipStream_t st = dev_->stream;
  std::unique_ptr<NeuralField> nf(new NeuralField());
  nf->device  = dev_;
  nf->sampler = s;
  s->retain();
  nf->attr = attributeIndex;
  nf->band = s->view.filter != kFilterCubic;

  // Parameters and optimizer state.
  NfAdam &a  = nf->adam;
  a.table16  = alloc<uint16_t>(nf->table16, kNfTableEntries);
  a.W1h      = alloc<uint16_t>(nf->W1h, kNfW1);
  a.W1Th     = alloc<uint16_t>(nf->W1Th, kNfW1);
  a.W2h      = alloc<uint16_t>(nf->W2h, kNfW2);
  a.W2Th     = alloc<uint16_t>(nf->W2Th, kNfW2);
  a.W3f      = alloc<float>(nf->W3f, kNfW3);
  a.table32  = alloc<float>(nf->table32, kNfTableEntries);
  a.W1_32    = alloc<float>(nf->W1_32, kNfW1);
  a.W2_32    = alloc<float>(nf->W2_32, kNfW2);
  a.W3_32    = alloc<float>(nf->W3_32, kNfW3);
  a.mTable   = alloc<float>(nf->mTable, kNfTableEntries);
  a.vTable   = alloc<float>(nf->vTable, kNfTableEntries);
  a.mW       = alloc<float>(nf->mW, kNfW1 + kNfW2 + kNfW3);
  a.vW       = alloc<float>(nf->vW, kNfW1 + kNfW2 + kNfW3);
  NfGrads &g = nf->grads;
  g.table    = alloc<float>(nf->gTable, kNfTableEntries);
  g.W1       = alloc<float>(nf->gW1, kNfW1);
  g.W2       = alloc<float>(nf->gW2, kNfW2);
  g.W3       = alloc<float>(nf->gW3, kNfW3);
  g.loss     = alloc<float>(nf->gLoss, 1);
  for (DeviceBuffer *b : {&nf->gTable, &nf->gW1, &nf->gW2, &nf->gW3, &nf->gLoss}) zero(*b, st);
  checkHip(launchNfInit(a, 0x5eedULL + seed, st), "neural field init");

  // Model description.
  NfModel &m = nf->model;
  m.w        = {a.table16, a.W1h, a.W1Th, a.W2h, a.W2Th, a.W3f};
  std::memcpy(m.o2i, v->grid.o2i, sizeof(m.o2i));
  float lo[3], hi[3], ext = 0.f;
  for (int c = 0; c < 3; c++) {
    lo[c] = v->grid.domainBox[c] + (float)v->grid.rootOrigin[c];
    hi[c] = v->grid.domainBox[3 + c] + (float)v->grid.rootOrigin[c];
    ext   = std::max(ext, hi[c] - lo[c]);
    m.lo[c] = lo[c];
  }
  m.invExtent = 1.f / ext;
  const double nmin = 16.0, nmax = std::max<double>(maxResolution ? maxResolution : 2048, 17.0);
  const double b    = std::exp((std::log(nmax) - std::log(nmin)) / (kNfLevels - 1));
  for (int l = 0; l < kNfLevels; l++) m.res[l] = (uint32_t)std::floor(nmin * std::pow(b, l));

  // Dense leaves in packed-slot order (the index the inference uses).
  nf->numLeaves = v->denseLeafOrigins.size() / 3;
  if (nf->numLeaves) {
    alloc<int32_t>(nf->leafOrigins, nf->numLeaves * 3);
    nf->leafOrigins.upload(v->denseLeafOrigins.data(), nf->numLeaves * 12, st);
    alloc<uint8_t>(nf->leafOk, nf->numLeaves);
    zero(nf->leafOk, st);
    alloc<float>(nf->leafErr, nf->numLeaves);
  }
  alloc<float>(nf->i2o, 12);
  nf->i2o.upload(v->i2o, 48, st);
  alloc<float>(nf->boxLo, 3);
  nf->boxLo.upload(lo, 12, st);
  alloc<float>(nf->boxHi, 3);
  nf->boxHi.upload(hi, 12, st);
  checkHip(hipStreamSynchronize(st), "neural field init");
  return reinterpret_cast<VKLNeuralField>(nf.release());
  NF_CATCH(nullptr)
}

OPENVKL_INTERFACE void vklReleaseNeuralField(VKLNeuralField h) { delete reinterpret_cast<NeuralField *>(h); }

OPENVKL_INTERFACE float vklNeuralFieldTrain(VKLNeuralField h, unsigned int steps, unsigned int batchSize,
                                            float learningRate, float leafFraction, unsigned int seed,
                                            float *lossHistory)
{
  NF_TRY(nfOf(h)->device)
  NeuralField *nf = nfOf(h);
  dev_->makeCurrent();
  hipStream_t st = dev_->stream;
  DeviceBuffer pts(batchSize * 12ull), tgt(batchSize * 4ull);
  float *hostLoss = nullptr;
  checkHip(hipHostMalloc(reinterpret_cast<void **>(&hostLoss), sizeof(float) * std::max(1u, steps)), "hipHostMalloc");
  float last = 0.f;
  try {
    for (unsigned k = 0; k < steps; k++) {
      const uint64_t s = ((uint64_t)seed << 32) ^ (uint64_t)(nf->step + 1) * 0x9E3779B97F4A7C15ull;
      checkHip(launchNfGenerate(nf->i2o.as<float>(), nf->boxLo.as<float>(), nf->boxHi.as<float>(),
                                nf->leafOrigins.as<int32_t>(), nf->numLeaves, leafFraction, batchSize, s,
                                pts.as<float>(), st),
               "generate");
      checkHip(launchSample(nf->sampler->viewCopy(), nf->sampler->attr(nf->attr), nf->attr, batchSize, pts.as<float>(),
                            tgt.as<float>(), st),
               "exact targets");
      checkHip(launchNfTrainStep(nf->model, batchSize, pts.as<float>(), tgt.as<float>(), nf->grads, st), "train");
      checkHip(hipMemcpyAsync(hostLoss + k, nf->grads.loss, 4, hipMemcpyDeviceToHost, st), "loss");
      nf->step++;
      checkHip(launchNfAdam(nf->adam, nf->grads, learningRate, learningRate, 0.9f, 0.99f, 1e-15f, nf->step,
                            1.f / (float)batchSize, st),
               "adam");
    }
    checkHip(hipStreamSynchronize(st), "train");
    for (unsigned k = 0; k < steps; k++) {
      const float mse = hostLoss[k] / (float)batchSize;
      if (lossHistory) lossHistory[k] = mse;
      last = mse;
    }
  } catch (...) {
    (void)hipHostFree(hostLoss);
    throw;
  }
  (void)hipHostFree(hostLoss);
  nf->finalized = false;
  return last;
  NF_CATCH(NAN)
}

// Approve leaves whose measured error (max over pointsPerLeaf random points,
// times `safety`) is within `tolerance`; the rest fall back to the exact
// sampler. Returns the number of approved leaves.
OPENVKL_INTERFACE unsigned int vklNeuralFieldFinalize(VKLNeuralField h, float tolerance, unsigned int pointsPerLeaf,
                                                      float safety, unsigned int seed)
{
  NF_TRY(nfOf(h)->device)
  NeuralField *nf = nfOf(h);
  dev_->makeCurrent();
  hipStream_t st = dev_->stream;
  const uint64_t n = nf->numLeaves * pointsPerLeaf;
  nf->tolerance    = tolerance;
  if (n == 0) return 0;
  DeviceBuffer pts(n * 12), a(n * 4), b(n * 4), cnt(4);
  checkHip(launchNfLeafPoints(nf->i2o.as<float>(), nf->leafOrigins.as<int32_t>(), nf->numLeaves, pointsPerLeaf,
                              0xA11CEULL + seed, pts.as<float>(), st),
           "leaf points");
  checkHip(launchNfInfer(nf->model, nf->sampler->viewCopy().grid, nf->attr, nullptr, n, pts.as<float>(), a.as<float>(),
                         nullptr, nullptr, 0, 0, st),
           "network");
  checkHip(launchSample(nf->sampler->viewCopy(), nf->sampler->attr(nf->attr), nf->attr, n, pts.as<float>(), b.as<float>(),
                        st),
           "exact");
  checkHip(launchNfLeafError(nf->sampler->viewCopy().grid, a.as<float>(), b.as<float>(), nf->numLeaves, pointsPerLeaf,
                             nf->leafErr.as<float>(), st),
           "leaf error");
  zero(cnt, st);
  checkHip(launchNfApprove(nf->sampler->viewCopy().grid, nf->leafOrigins.as<int32_t>(), nf->leafErr.as<float>(),
                           nf->numLeaves, nf->attr, tolerance, safety, nf->band ? 1 : 0, nf->leafOk.as<uint8_t>(),
                           cnt.as<uint32_t>(), st),
           "approve");
  uint32_t ok = 0;
  checkHip(hipMemcpyAsync(&ok, cnt.ptr, 4, hipMemcpyDeviceToHost, st), "copy");
  checkHip(hipStreamSynchronize(st), "finalize");
  nf->numOk     = ok;
  nf->finalized = true;
  return ok;
  NF_CATCH(0)
}

// Sample the neural field at N object-space points (device pointers). With
// useFallback, points whose cell leaf is not approved (or lies outside dense
// leaves) are sampled exactly; *numFallback (host pointer, may be NULL)
// receives their count. `workspace` must hold at least 4 * N + 4 bytes of
// device memory, or be NULL (then it is allocated per call).
OPENVKL_INTERFACE void vklNeuralFieldSampleN(VKLNeuralField h, unsigned int N, const vkl_vec3f *objectCoordinates,
                                             float *samples, int useFallback, unsigned int *numFallback,
                                             void *workspace)
{
  NF_TRY(nfOf(h)->device)
  NeuralField *nf = nfOf(h);
  dev_->makeCurrent();
  hipStream_t st = dev_->stream;
  if (N == 0) return;
  if (useFallback && !nf->finalized) throw VklError("vklNeuralFieldFinalize must run before sampling with fallback");
  DeviceBuffer own;
  uint8_t *ws = static_cast<uint8_t *>(workspace);
  if (useFallback && !ws) {
    own.allocate(4ull * N + 4);
    ws = own.as<uint8_t>();
  }
  uint32_t *count = useFallback ? reinterpret_cast<uint32_t *>(ws) : nullptr;
  uint32_t *list  = useFallback ? reinterpret_cast<uint32_t *>(ws + 4) : nullptr;
  if (useFallback) checkHip(hipMemsetAsync(count, 0, 4, st), "hipMemsetAsync");
  const float *xyz = reinterpret_cast<const float *>(objectCoordinates);
  checkHip(launchNfInfer(nf->model, nf->sampler->viewCopy().grid, nf->attr, nf->leafOk.as<uint8_t>(), N, xyz, samples, list,
                         count, useFallback, useFallback && nf->band ? 1 : 0, st),
           "network");
  uint32_t fb = 0;
  if (useFallback) {
    checkHip(hipMemcpyAsync(&fb, count, 4, hipMemcpyDeviceToHost, st), "copy");
    checkHip(hipStreamSynchronize(st), "network");
    if (fb)
      checkHip(launchGatherSample(nf->sampler->viewCopy(), nf->sampler->attr(nf->attr), nf->attr, list, count, fb, xyz,
                                  samples, st),
               "fallback");
  }
  checkHip(hipStreamSynchronize(st), "neural field sample");
  if (numFallback) *numFallback = fb;
  NF_CATCH()
}

// One forward + backward pass without an optimizer step: writes the summed
// per-sample gradients of sum((y - t)^2) (host pointers, fp32; table
// gradients [levels][2^19][2]) and returns the loss sum.
OPENVKL_INTERFACE float vklNeuralFieldGradients(VKLNeuralField h, unsigned int N, const vkl_vec3f *objectCoordinates,
                                                const float *targets, float *gTable, float *gW1, float *gW2,
                                                float *gW3)
{
  NF_TRY(nfOf(h)->device)
  NeuralField *nf = nfOf(h);
  dev_->makeCurrent();
  hipStream_t st = dev_->stream;
  DeviceBuffer pts(N * 12ull), tgt(N * 4ull);
  pts.upload(objectCoordinates, N * 12ull, st);
  tgt.upload(targets, N * 4ull, st);
  checkHip(launchNfTrainStep(nf->model, N, pts.as<float>(), tgt.as<float>(), nf->grads, st), "train");
  float loss = 0.f;
  checkHip(hipMemcpyAsync(&loss, nf->grads.loss, 4, hipMemcpyDeviceToHost, st), "copy");
  checkHip(hipMemcpyAsync(gTable, nf->grads.table, kNfTableEntries * 4, hipMemcpyDeviceToHost, st), "copy");
  checkHip(hipMemcpyAsync(gW1, nf->grads.W1, kNfW1 * 4, hipMemcpyDeviceToHost, st), "copy");
  checkHip(hipMemcpyAsync(gW2, nf->grads.W2, kNfW2 * 4, hipMemcpyDeviceToHost, st), "copy");
  checkHip(hipMemcpyAsync(gW3, nf->grads.W3, kNfW3 * 4, hipMemcpyDeviceToHost, st), "copy");
  for (DeviceBuffer *b : {&nf->gTable, &nf->gW1, &nf->gW2, &nf->gW3, &nf->gLoss}) zero(*b, st);
  checkHip(hipStreamSynchronize(st), "gradients");
  return loss;
  NF_CATCH(NAN)
}

// Copies of the binary16 parameters (host pointers): table [levels][2^19][2],
// W1 [32][64], W2 [64][64], W3 [64] (fp32); and the level resolutions [16].
OPENVKL_INTERFACE void vklNeuralFieldGetParams(VKLNeuralField h, uint16_t *table, uint16_t *W1, uint16_t *W2,
                                               float *W3, unsigned int *resolutions, float *boxLo, float *invExtent)
{
  NF_TRY(nfOf(h)->device)
  NeuralField *nf = nfOf(h);
  dev_->makeCurrent();
  checkHip(hipDeviceSynchronize(), "sync");
  if (table) checkHip(hipMemcpy(table, nf->table16.ptr, kNfTableEntries * 2, hipMemcpyDeviceToHost), "copy");
  if (W1) checkHip(hipMemcpy(W1, nf->W1h.ptr, kNfW1 * 2, hipMemcpyDeviceToHost), "copy");
  if (W2) checkHip(hipMemcpy(W2, nf->W2h.ptr, kNfW2 * 2, hipMemcpyDeviceToHost), "copy");
  if (W3) checkHip(hipMemcpy(W3, nf->W3f.ptr, kNfW3 * 4, hipMemcpyDeviceToHost), "copy");
  if (resolutions) std::memcpy(resolutions, nf->model.res, sizeof(nf->model.res));
  if (boxLo) std::memcpy(boxLo, nf->model.lo, 12);
  if (invExtent) *invExtent = nf->model.invExtent;
  NF_CATCH()
}

OPENVKL_INTERFACE unsigned int vklNeuralFieldNumApprovedLeaves(VKLNeuralField h, unsigned int *numLeaves)
{
  NeuralField *nf = reinterpret_cast<NeuralField *>(h);
  if (!nf) return 0;
  if (numLeaves) *numLeaves = (unsigned int)nf->numLeaves;
  return nf->numOk;
}

}  // extern "C"
