// Batched sampling plans: a device descriptor table for many (sampler,
// ragged point batch) pairs, grouped by kernel specialisation (voxel source
// type x filter); one persistent kernel launch per group, all captured in one
// hipGraph; a small autotuner for the launch shape.
//
// Autotuning: on plan creation the candidate shapes (threads per workgroup x
// points per tile x workgroups per CU) are timed with HIP events over a few
// executions of all groups and the fastest is kept. Results are cached in
// memory per workload signature (batch count and, per group, its
// specialisation key, batch count and total points rounded to a power of two)
// and, if OPENVKL_HIP_TUNE_CACHE names a file, persisted there as text lines
// "signature blocks threads tile".

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "batch.h"
#include "objects.h"
#include "openvkl/openvkl_hip.h"

namespace vklhip {

struct SampleBatchPlan
{
  Device *device = nullptr;
  std::vector<Sampler *> samplers;  // retained
  DeviceBuffer items, begins;       // all groups' items / offsets, group after group
  std::vector<BatchGroup> groups;
  uint32_t numItems = 0;
  uint64_t total    = 0;
  BatchLaunch launch{};
  hipGraph_t graph        = nullptr;
  hipGraphExec_t graphExec = nullptr;
  std::string signature;

  ~SampleBatchPlan()
  {
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    if (graph) (void)hipGraphDestroy(graph);
    for (Sampler *s : samplers) s->release();
  }
};

namespace {

std::mutex g_tuneMutex;
std::map<std::string, BatchLaunch> g_tuneCache;
bool g_tuneFileLoaded = false;

void loadTuneFile()
{
  if (g_tuneFileLoaded) return;
  g_tuneFileLoaded = true;
  const char *path = std::getenv("OPENVKL_HIP_TUNE_CACHE");
  if (!path) return;
  std::ifstream f(path);
  std::string sig;
  BatchLaunch l;
  while (f >> sig >> l.blocks >> l.threads >> l.tile) g_tuneCache[sig] = l;
}

void storeTune(const std::string &sig, const BatchLaunch &l)
{
  g_tuneCache[sig] = l;
  if (const char *path = std::getenv("OPENVKL_HIP_TUNE_CACHE")) {
    std::ofstream f(path, std::ios::app);
    f << sig << " " << l.blocks << " " << l.threads << " " << l.tile << "\n";
  }
}

std::vector<BatchLaunch> candidates(int cus)
{
  std::vector<BatchLaunch> c;
  for (uint32_t threads : {256u, 512u, 1024u})
    for (uint32_t ppt : {1u, 4u, 16u})
      for (uint32_t perCu : {2u, 8u}) c.push_back({(uint32_t)cus * perCu, threads, threads * ppt});
  return c;
}

BatchLaunch tune(SampleBatchPlan &p, hipStream_t st)
{
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  std::lock_guard<std::mutex> lock(g_tuneMutex);
  loadTuneFile();
  auto it = g_tuneCache.find(p.signature);
  if (it != g_tuneCache.end()) return it->second;
  hipEvent_t a, b;
  checkHip(hipEventCreate(&a), "hipEventCreate");
  checkHip(hipEventCreate(&b), "hipEventCreate");
  BatchLaunch best{};
  float bestMs = 1e30f;
  for (const BatchLaunch &c : candidates(cus)) {
    checkHip(launchSampleBatch(p.groups.data(), (uint32_t)p.groups.size(), c, st), "batch tune");
    checkHip(hipEventRecord(a, st), "hipEventRecord");
    for (int r = 0; r < 3; r++)
      checkHip(launchSampleBatch(p.groups.data(), (uint32_t)p.groups.size(), c, st), "batch tune");
    checkHip(hipEventRecord(b, st), "hipEventRecord");
    checkHip(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0.f;
    checkHip(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    if (ms < bestMs) {
      bestMs = ms;
      best   = c;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  storeTune(p.signature, best);
  return best;
}

}  // namespace

}  // namespace vklhip

using namespace vklhip;

extern "C" {

OPENVKL_INTERFACE VKLSampleBatchPlan vklNewSampleBatchPlan(unsigned int numBatches, const VKLSampler *samplers,
                                                           const unsigned int *attributeIndices,
                                                           const unsigned int *counts,
                                                           const vkl_vec3f *const *objectCoordinates,
                                                           float *const *samples, int autotune)
{
  Device *dev = nullptr;
  try {
    if (numBatches == 0 || !samplers) throw VklError("empty sample batch");
    Sampler *s0 = static_cast<Sampler *>(samplers[0].host);
    if (!s0) throw VklError("null sampler");
    dev = s0->device;
    dev->makeCurrent();
    auto *p   = new SampleBatchPlan();
    p->device = dev;
    // Group the batches by kernel specialisation (order kept within a group).
    std::map<uint32_t, std::vector<unsigned>> byKey;
    std::vector<BatchItem> all(numBatches);
    for (unsigned i = 0; i < numBatches; i++) {
      Sampler *s = static_cast<Sampler *>(samplers[i].host);
      if (!s || !s->committed) {
        delete p;
        throw VklError("sample batch: sampler " + std::to_string(i) + " is null or not committed");
      }
      s->retain();
      p->samplers.push_back(s);
      const uint32_t a = attributeIndices ? attributeIndices[i] : 0;
      BatchItem &it    = all[i];
      it.sampler       = s->view;
      it.attr          = s->attr(a);
      it.attrIndex     = a;
      it.xyz           = reinterpret_cast<const float *>(objectCoordinates[i]);
      it.out           = samples[i];
      it.count         = counts[i];
      if (counts[i] && (!isDevicePointer(it.xyz) || !isDevicePointer(it.out))) {
        delete p;
        throw VklError("sample batch: point and sample arrays must be device memory");
      }
      byKey[sourceKey(s->view.grid, it.attr, s->view.filter)].push_back(i);
    }
    std::vector<BatchItem> items;
    std::vector<uint64_t> begins;
    std::vector<size_t> firstOf;
    std::ostringstream sig;
    sig << "b" << numBatches;
    uint64_t total = 0;
    for (const auto &kv : byKey) {
      BatchGroup g{};
      const BatchItem &first = all[kv.second.front()];
      g.grid                 = first.sampler.grid;
      g.attr                 = first.attr;
      g.filter               = first.sampler.filter;
      g.numItems             = (uint32_t)kv.second.size();
      firstOf.push_back(items.size());
      uint64_t gt = 0;
      for (unsigned i : kv.second) {
        BatchItem it = all[i];
        it.begin     = gt;
        begins.push_back(gt);
        items.push_back(it);
        gt += it.count;
      }
      g.total = gt;
      total += gt;
      p->groups.push_back(g);
      uint64_t pow2 = 1;
      while (pow2 < gt) pow2 <<= 1;
      sig << "_k" << kv.first << "x" << kv.second.size() << "n" << pow2;
    }
    p->numItems = numBatches;
    p->total    = total;
    p->items.allocate(items.size() * sizeof(BatchItem));
    p->items.upload(items.data(), items.size() * sizeof(BatchItem), dev->stream);
    p->begins.allocate(begins.size() * sizeof(uint64_t));
    p->begins.upload(begins.data(), begins.size() * sizeof(uint64_t), dev->stream);
    for (size_t g = 0; g < p->groups.size(); g++) {
      p->groups[g].items  = p->items.as<BatchItem>() + firstOf[g];
      p->groups[g].begins = p->begins.as<uint64_t>() + firstOf[g];
    }
    p->signature = sig.str();
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev->hipDevice);
    p->launch = autotune ? tune(*p, dev->stream) : BatchLaunch{(uint32_t)cus * 8u, 256u, 1024u};
    // Capture the persistent launch in a graph (kernel node only).
    hipStream_t cap;
    checkHip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    checkHip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    checkHip(launchSampleBatch(p->groups.data(), (uint32_t)p->groups.size(), p->launch, cap), "batch capture");
    checkHip(hipStreamEndCapture(cap, &p->graph), "hipStreamEndCapture");
    checkHip(hipGraphInstantiate(&p->graphExec, p->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    (void)hipStreamDestroy(cap);
    return reinterpret_cast<VKLSampleBatchPlan>(p);
  } catch (const std::exception &e) {
    if (dev)
      dev->handleError(VKL_UNKNOWN_ERROR, e.what());
    else
      std::fprintf(stderr, "OPENVKL ERROR: %s\n", e.what());
    return nullptr;
  }
}

OPENVKL_INTERFACE void vklExecuteSampleBatchPlan(VKLSampleBatchPlan plan, void *stream)
{
  SampleBatchPlan *p = reinterpret_cast<SampleBatchPlan *>(plan);
  if (!p) return;
  try {
    p->device->makeCurrent();
    checkHip(hipGraphLaunch(p->graphExec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
  } catch (const std::exception &e) {
    p->device->handleError(VKL_UNKNOWN_ERROR, e.what());
  }
}

OPENVKL_INTERFACE void vklGetSampleBatchPlanLaunch(VKLSampleBatchPlan plan, unsigned int *blocks,
                                                   unsigned int *threads, unsigned int *tile)
{
  SampleBatchPlan *p = reinterpret_cast<SampleBatchPlan *>(plan);
  if (!p) return;
  if (blocks) *blocks = p->launch.blocks;
  if (threads) *threads = p->launch.threads;
  if (tile) *tile = p->launch.tile;
}

OPENVKL_INTERFACE void vklReleaseSampleBatchPlan(VKLSampleBatchPlan plan)
{
  delete reinterpret_cast<SampleBatchPlan *>(plan);
}

}  // extern "C"
