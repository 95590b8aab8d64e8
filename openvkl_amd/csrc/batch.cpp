// Batched sampling plans: a device descriptor table for many (sampler,
// ragged point batch) pairs, grouped for launch; one persistent kernel launch
// per group, all captured in one hipGraph; a small autotuner for the launch
// shape and the grouping.
//
// Two groupings are kept: per kernel specialisation (voxel source type x
// filter), and "merged", where the structuredRegular batches with 32-bit
// voxel offsets of one filter share one launch whatever their element types
// (the type is chosen per batch inside the kernel; its registers are the
// maximum over the six element types, which can cost occupancy).
//
// Autotuning: on plan creation the candidate shapes (threads per workgroup x
// points per tile x workgroups per CU, each with either grouping when they
// differ) are timed with HIP events over a few executions of all groups and
// the fastest is kept. Results are cached in memory per workload signature
// (batch count and, per group, its specialisation key, batch count and total
// points rounded to a power of two) and, if OPENVKL_HIP_TUNE_CACHE names a
// file, persisted there as text lines "signature blocks threads tile merged".
// Without autotuning the per-specialisation grouping is used.
//
// Tricubic batches on structuredRegular volumes with cubic records do not
// join the groups: each runs the stencil-cooperative kernel of
// vklComputeSampleN (its lanes load whole stencils together), captured in the
// same graph after the group launches.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "batch.h"
#include "kernels.h"
#include "objects.h"
#include "openvkl/openvkl_hip.h"

namespace vklhip {

struct SampleBatchPlan
{
  Device *device = nullptr;
  std::vector<Sampler *> samplers;  // retained
  // [0]: per specialisation, [1]: merged structuredRegular groups; each the
  // groups' items / offsets, group after group
  DeviceBuffer items[2], begins[2];
  std::vector<BatchGroup> groups[2];
  uint32_t numItems = 0;
  uint64_t total    = 0;
  BatchLaunch launch{};
  std::vector<BatchItem> direct;  // tricubic batches on cubic records
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
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream in(line);
    std::string sig;
    BatchLaunch l{};
    if (in >> sig >> l.blocks >> l.threads >> l.tile >> l.merged && l.merged <= 1 && l.threads && l.blocks)
      g_tuneCache[sig] = l;
  }
}

void storeTune(const std::string &sig, const BatchLaunch &l)
{
  g_tuneCache[sig] = l;
  if (const char *path = std::getenv("OPENVKL_HIP_TUNE_CACHE")) {
    std::ofstream f(path, std::ios::app);
    f << sig << " " << l.blocks << " " << l.threads << " " << l.tile << " " << l.merged << "\n";
  }
}

std::vector<BatchLaunch> candidates(int cus, bool canMerge)
{
  std::vector<BatchLaunch> c;
  for (uint32_t merged = 0; merged <= (canMerge ? 1u : 0u); merged++)
    for (uint32_t threads : {256u, 512u, 1024u})
      for (uint32_t ppt : {1u, 4u, 16u})
        for (uint32_t perCu : {2u, 8u}) c.push_back({(uint32_t)cus * perCu, threads, threads * ppt, merged});
  return c;
}

// Batches the cubic-records kernel serves (as the stream call would).
bool onCubicRecords(const BatchItem &it)
{
  const GridView &g = it.sampler.grid;
  return it.sampler.filter == kFilterCubic && it.attr.cubic && g.dense && !g.temporalFormat &&
         (it.attr.cubicStencils || g.activeSize[0] >= 4);
}

hipError_t launchPlan(const SampleBatchPlan &p, const BatchLaunch &c, hipStream_t st)
{
  const std::vector<BatchGroup> &g = p.groups[c.merged];
  hipError_t e = g.empty() ? hipSuccess : launchSampleBatch(g.data(), (uint32_t)g.size(), c, st);
  for (const BatchItem &it : p.direct) {
    if (e != hipSuccess) break;
    e = launchCubicRecords(it.sampler, it.attr, it.attrIndex, it.count, it.xyz, it.out, 1, false, st);
  }
  return e;
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
  for (const BatchLaunch &c : candidates(cus, p.groups[1].size() < p.groups[0].size())) {
    checkHip(launchPlan(p, c, st), "batch tune");
    checkHip(hipEventRecord(a, st), "hipEventRecord");
    for (int r = 0; r < 3; r++) checkHip(launchPlan(p, c, st), "batch tune");
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
      it.sampler       = s->viewCopy();
      it.attr          = s->attr(a);
      it.attrIndex     = a;
      it.xyz           = reinterpret_cast<const float *>(objectCoordinates[i]);
      it.out           = samples[i];
      it.count         = counts[i];
      if (counts[i] && (!isDevicePointer(it.xyz) || !isDevicePointer(it.out))) {
        delete p;
        throw VklError("sample batch: point and sample arrays must be device memory");
      }
    }
    // Tricubic batches on cubic records run on their own (see the top).
    std::vector<unsigned> grouped;
    for (unsigned i = 0; i < numBatches; i++) {
      if (onCubicRecords(all[i])) {
        if (all[i].count) p->direct.push_back(all[i]);
      } else {
        grouped.push_back(i);
      }
    }
    std::ostringstream sig;
    sig << "b" << numBatches << "d" << p->direct.size();
    for (int merged = 0; merged < 2; merged++) {
      // Group the batches (order kept within a group): per kernel
      // specialisation, or with the structuredRegular 32-bit-offset batches
      // of a filter in one group.
      std::map<uint32_t, std::vector<unsigned>> byKey;
      for (unsigned i : grouped) {
        const BatchItem &it = all[i];
        const int f         = it.sampler.filter;
        byKey[merged ? groupKey(it.sampler.grid, it.attr, f) : sourceKey(it.sampler.grid, it.attr, f)].push_back(i);
      }
      std::vector<BatchItem> items;
      std::vector<uint64_t> begins;
      std::vector<size_t> firstOf;
      uint64_t total = 0;
      for (const auto &kv : byKey) {
        BatchGroup g{};
        const BatchItem &first = all[kv.second.front()];
        g.grid                 = first.sampler.grid;
        g.attr                 = first.attr;
        g.filter               = first.sampler.filter;
        g.numItems             = (uint32_t)kv.second.size();
        // A merged group whose batches share one element type runs the
        // kernel specialised for it.
        if (merged && mergesDense(g.grid, g.attr))
          for (unsigned i : kv.second)
            if (all[i].attr.type != first.attr.type) g.mergedDense = 1;
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
        p->groups[merged].push_back(g);
        if (!merged) {
          uint64_t pow2 = 1;
          while (pow2 < gt) pow2 <<= 1;
          sig << "_k" << kv.first << "x" << kv.second.size() << "n" << pow2;
        }
      }
      p->total = total;
      p->items[merged].allocate(items.size() * sizeof(BatchItem));
      p->items[merged].upload(items.data(), items.size() * sizeof(BatchItem), dev->stream);
      p->begins[merged].allocate(begins.size() * sizeof(uint64_t));
      p->begins[merged].upload(begins.data(), begins.size() * sizeof(uint64_t), dev->stream);
      for (size_t g = 0; g < p->groups[merged].size(); g++) {
        p->groups[merged][g].items  = p->items[merged].as<BatchItem>() + firstOf[g];
        p->groups[merged][g].begins = p->begins[merged].as<uint64_t>() + firstOf[g];
      }
    }
    p->numItems = numBatches;
    p->signature = sig.str();
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev->hipDevice);
    if (autotune < 0 || autotune > 2) throw VklError("sample batch: autotune must be 0, 1 or 2");
    p->launch = autotune == 1 ? tune(*p, dev->stream)
                              : BatchLaunch{(uint32_t)cus * 8u, 256u, 1024u, autotune == 2 ? 1u : 0u};
    // Capture the persistent launch in a graph (kernel node only).
    hipStream_t cap;
    checkHip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    checkHip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    checkHip(launchPlan(*p, p->launch, cap), "batch capture");
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

OPENVKL_INTERFACE unsigned int vklGetSampleBatchPlanLaunchCount(VKLSampleBatchPlan plan)
{
  SampleBatchPlan *p = reinterpret_cast<SampleBatchPlan *>(plan);
  if (!p) return 0;
  unsigned int n = (unsigned int)p->direct.size();
  for (const BatchGroup &g : p->groups[p->launch.merged]) n += (g.total && g.numItems) ? 1u : 0u;
  return n;
}

OPENVKL_INTERFACE void vklReleaseSampleBatchPlan(VKLSampleBatchPlan plan)
{
  delete reinterpret_cast<SampleBatchPlan *>(plan);
}

}  // extern "C"
