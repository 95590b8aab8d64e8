// Batched sampling plans: a device descriptor table for many (sampler,
// ragged point batch) pairs, one persistent kernel launch captured in a
// hipGraph, and a small autotuner for the launch shape.
//
// Autotuning: on plan creation the candidate shapes (threads per workgroup x
// points per tile x workgroups per CU) are timed with HIP events over a few
// launches and the fastest is kept. Results are cached in memory per workload
// signature (batch count, total points rounded to a power of two, fast-path
// mix) and, if OPENVKL_HIP_TUNE_CACHE names a file, persisted there as text
// lines "signature blocks threads tile".

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
  DeviceBuffer items, begins;
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
    checkHip(launchSampleBatch(p.items.as<BatchItem>(), p.begins.as<uint64_t>(), p.numItems, p.total, c, st),
             "batch tune");
    checkHip(hipEventRecord(a, st), "hipEventRecord");
    for (int r = 0; r < 3; r++)
      checkHip(launchSampleBatch(p.items.as<BatchItem>(), p.begins.as<uint64_t>(), p.numItems, p.total, c, st),
               "batch tune");
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

int32_t fastPathOf(const Sampler &s, const AttrView &a)
{
  const GridView &g = s.view.grid;
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  if (!g.dense || s.view.filter != kFilterLinear || nbz * g.layout.strideZ >= (1ull << 32)) return kBatchGeneric;
  if (a.type == kTypeFloat) return kBatchDenseLinearF32;
  if (a.type == kTypeHalf) return kBatchDenseLinearF16;
  return kBatchGeneric;
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
    std::vector<BatchItem> items(numBatches);
    std::vector<uint64_t> begins(numBatches);
    uint64_t total = 0;
    int fast[3]    = {0, 0, 0};
    for (unsigned i = 0; i < numBatches; i++) {
      Sampler *s = static_cast<Sampler *>(samplers[i].host);
      if (!s || !s->committed) {
        delete p;
        throw VklError("sample batch: sampler " + std::to_string(i) + " is null or not committed");
      }
      s->retain();
      p->samplers.push_back(s);
      const uint32_t a = attributeIndices ? attributeIndices[i] : 0;
      BatchItem &it    = items[i];
      it.sampler       = s->view;
      it.attr          = s->attr(a);
      it.attrIndex     = a;
      it.xyz           = reinterpret_cast<const float *>(objectCoordinates[i]);
      it.out           = samples[i];
      if (counts[i] && (!isDevicePointer(it.xyz) || !isDevicePointer(it.out))) {
        delete p;
        throw VklError("sample batch: point and sample arrays must be device memory");
      }
      it.begin    = total;
      it.count    = counts[i];
      it.fastPath = fastPathOf(*s, it.attr);
      fast[it.fastPath]++;
      begins[i] = total;
      total += counts[i];
    }
    p->numItems = numBatches;
    p->total    = total;
    p->items.allocate(items.size() * sizeof(BatchItem));
    p->items.upload(items.data(), items.size() * sizeof(BatchItem), dev->stream);
    p->begins.allocate(begins.size() * sizeof(uint64_t));
    p->begins.upload(begins.data(), begins.size() * sizeof(uint64_t), dev->stream);
    uint64_t pow2 = 1;
    while (pow2 < total) pow2 <<= 1;
    std::ostringstream sig;
    sig << "b" << numBatches << "_n" << pow2 << "_f" << fast[0] << "." << fast[1] << "." << fast[2];
    p->signature = sig.str();
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev->hipDevice);
    p->launch = autotune ? tune(*p, dev->stream) : BatchLaunch{(uint32_t)cus * 8u, 256u, 1024u};
    // Capture the persistent launch in a graph (kernel node only).
    hipStream_t cap;
    checkHip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    checkHip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    checkHip(launchSampleBatch(p->items.as<BatchItem>(), p->begins.as<uint64_t>(), p->numItems, p->total,
                               p->launch, cap),
             "batch capture");
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
