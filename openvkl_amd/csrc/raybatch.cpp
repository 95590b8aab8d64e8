// Ray batch plans: interval and hit iteration of many (iterator context, ray
// batch) pairs of uneven size, possibly over different volumes, replayed as
// one hipGraph. Each batch is the launch vklIterateIntervalsBatch /
// vklIterateHitsBatch would make (interval kernel specialised on the
// context's depth; persistent hit kernel with its own ray counter), and the
// batches are captured on parallel branches of the graph, so small batches of
// different volumes run side by side on the GPU instead of one after the
// other. Batches are spread round-robin over at most kMaxBranches branches.
//
// A plan keeps copies of the context views it was built with (in its own
// device memory), so later changes of a context or its sampler (observers)
// do not reach it; the contexts are retained.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "kernels.h"
#include "objects.h"
#include "openvkl/openvkl_hip.h"

namespace vklhip {

constexpr unsigned kMaxBranches = 16;

struct RayBatchPlan
{
  Device *device = nullptr;
  std::vector<IteratorContext *> contexts;  // retained
  DeviceBuffer views;                       // one IteratorContextView per batch
  DeviceBuffer work;                        // one hit-batch ray counter per batch
  uint32_t numLaunches = 0;
  hipGraph_t graph         = nullptr;
  hipGraphExec_t graphExec = nullptr;

  ~RayBatchPlan()
  {
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    if (graph) (void)hipGraphDestroy(graph);
    for (IteratorContext *c : contexts) c->release();
  }
};

}  // namespace vklhip

using namespace vklhip;

extern "C" {

OPENVKL_INTERFACE VKLRayBatchPlan vklNewRayBatchPlan(unsigned int numBatches, const VKLObject *contexts,
                                                     const unsigned int *numRays, const vkl_vec3f *const *origins,
                                                     const vkl_vec3f *const *directions,
                                                     const vkl_range1f *const *tRanges, const float *const *times,
                                                     void *const *results, const unsigned int *maxPerRay,
                                                     unsigned int *const *counts)
{
  Device *dev         = nullptr;
  RayBatchPlan *p     = nullptr;
  hipStream_t main    = nullptr;
  std::vector<hipStream_t> branches;
  std::vector<hipEvent_t> events;
  bool capturing = false;
  try {
    if (numBatches == 0 || !contexts || !numRays || !origins || !directions || !results || !maxPerRay || !counts)
      throw VklError("ray batch plan: null argument", VKL_INVALID_ARGUMENT);
    p = new RayBatchPlan();
    std::vector<IteratorContextView> hostViews(numBatches);
    for (unsigned i = 0; i < numBatches; i++) {
      IteratorContext *ctx = static_cast<IteratorContext *>(contexts[i].host);
      if (!ctx || !ctx->committed)
        throw VklError("ray batch plan: context " + std::to_string(i) + " is null or not committed",
                       VKL_INVALID_ARGUMENT);
      if (!dev) dev = ctx->device;
      if (ctx->device != dev) throw VklError("ray batch plan: contexts of different devices", VKL_INVALID_ARGUMENT);
      ctx->retain();
      p->contexts.push_back(ctx);
      hostViews[i] = ctx->refreshSampler();
      const void *arrays[] = {origins[i], directions[i], results[i], counts[i],
                              tRanges ? tRanges[i] : nullptr, times ? times[i] : nullptr};
      for (const void *a : arrays)
        if (numRays[i] && a && !isDevicePointer(a))
          throw VklError("ray batch plan: ray and result arrays must be device memory", VKL_INVALID_ARGUMENT);
      if (numRays[i] && (!origins[i] || !directions[i] || !results[i] || !counts[i]))
        throw VklError("ray batch plan: null ray or result array", VKL_INVALID_ARGUMENT);
    }
    p->device = dev;
    dev->makeCurrent();
    hipStream_t st = dev->stream;
    p->views.allocate(numBatches * sizeof(IteratorContextView));
    p->views.upload(hostViews.data(), numBatches * sizeof(IteratorContextView), st);
    p->work.allocate(numBatches * sizeof(uint32_t));
    checkHip(hipStreamSynchronize(st), "ray batch plan upload");
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev->hipDevice);

    // Capture: the main stream forks the branches, batch i goes to branch
    // i mod (number of branches), and the main stream joins them all.
    const unsigned nb = numBatches < kMaxBranches ? numBatches : kMaxBranches;
    checkHip(hipStreamCreateWithFlags(&main, hipStreamNonBlocking), "hipStreamCreate");
    events.resize(nb + 1, nullptr);
    for (auto &e : events) checkHip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    branches.resize(nb, nullptr);
    for (auto &b : branches) checkHip(hipStreamCreateWithFlags(&b, hipStreamNonBlocking), "hipStreamCreate");
    checkHip(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    capturing = true;
    checkHip(hipEventRecord(events[nb], main), "fork");
    for (hipStream_t b : branches) checkHip(hipStreamWaitEvent(b, events[nb], 0), "fork wait");
    const IteratorContextView *dv = p->views.as<IteratorContextView>();
    uint32_t *work                = p->work.as<uint32_t>();
    for (unsigned i = 0; i < numBatches; i++) {
      hipStream_t b = branches[i % nb];
      if (numRays[i]) {
        const IteratorContextView &hv = hostViews[i];
        const float *o  = reinterpret_cast<const float *>(origins[i]);
        const float *d  = reinterpret_cast<const float *>(directions[i]);
        const float *tr = tRanges ? reinterpret_cast<const float *>(tRanges[i]) : nullptr;
        if (p->contexts[i]->isHit) {
          const uint32_t lanes = batchLaneCount(numRays[i], cus);
          checkHip(launchSetWork(work + i, lanes, b), "ray batch work counter");
          checkHip(launchHitBatch(hv, dv + i, p->contexts[i]->sampler->attr(hv.attributeIndex), numRays[i], o, d, tr,
                                  times ? times[i] : nullptr, static_cast<HitOut *>(results[i]), maxPerRay[i],
                                  counts[i], work + i, lanes, b),
                   "ray batch hits");
        } else {
          checkHip(launchIntervalBatch(hv.maxIteratorDepth, denseQueryOf(hv.sampler.grid), dv + i, numRays[i], o, d, tr,
                                       static_cast<IntervalOut *>(results[i]), maxPerRay[i], counts[i], b),
                   "ray batch intervals");
        }
        p->numLaunches++;
      }
    }
    for (unsigned k = 0; k < nb; k++) {
      checkHip(hipEventRecord(events[k], branches[k]), "join");
      checkHip(hipStreamWaitEvent(main, events[k], 0), "join wait");
    }
    capturing = false;
    checkHip(hipStreamEndCapture(main, &p->graph), "hipStreamEndCapture");
    checkHip(hipGraphInstantiate(&p->graphExec, p->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    for (auto &b : branches) (void)hipStreamDestroy(b);
    for (auto &e : events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(main);
    return reinterpret_cast<VKLRayBatchPlan>(p);
  } catch (const std::exception &e) {
    if (capturing) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(main, &g);
      if (g) (void)hipGraphDestroy(g);
    }
    for (auto &b : branches)
      if (b) (void)hipStreamDestroy(b);
    for (auto &ev : events)
      if (ev) (void)hipEventDestroy(ev);
    if (main) (void)hipStreamDestroy(main);
    delete p;
    const VklError *ve = dynamic_cast<const VklError *>(&e);
    if (dev)
      dev->handleError(ve ? ve->code : VKL_UNKNOWN_ERROR, e.what());
    else
      std::fprintf(stderr, "OPENVKL ERROR: %s\n", e.what());
    return nullptr;
  }
}

OPENVKL_INTERFACE void vklExecuteRayBatchPlan(VKLRayBatchPlan plan, void *stream)
{
  RayBatchPlan *p = reinterpret_cast<RayBatchPlan *>(plan);
  if (!p) return;
  try {
    p->device->makeCurrent();
    checkHip(hipGraphLaunch(p->graphExec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
  } catch (const std::exception &e) {
    p->device->handleError(VKL_UNKNOWN_ERROR, e.what());
  }
}

OPENVKL_INTERFACE unsigned int vklGetRayBatchPlanLaunchCount(VKLRayBatchPlan plan)
{
  RayBatchPlan *p = reinterpret_cast<RayBatchPlan *>(plan);
  return p ? p->numLaunches : 0;
}

OPENVKL_INTERFACE void vklReleaseRayBatchPlan(VKLRayBatchPlan plan)
{
  delete reinterpret_cast<RayBatchPlan *>(plan);
}

}  // extern "C"
