// C ABI of the HIP device: the object API (openvkl/api/API.cpp:104-525) and
// the hot-path calls (openvkl/devices/cpu/api/DeviceAPI.cpp:65-533), plus the
// batch / async extensions declared in openvkl/openvkl_hip.h.
//
// Error contract (API.cpp:37-89): every entry catches exceptions and reports
// them to the owning device (last error code / message + error callback);
// std::bad_alloc maps to VKL_OUT_OF_MEMORY, everything else to the code the
// exception carries (VKL_UNKNOWN_ERROR by default). Default returns: NaN
// samples, null handles / iterators, 0 sizes, 0 from iterate calls.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

#include "binned.h"
#include "kernels.h"
#include "noalloc.h"
#include "objects.h"
#include "openvkl/openvkl_hip.h"

using namespace vklhip;

namespace vklhip {

// ---- Sampler ------------------------------------------------------------------

Sampler::Sampler(Device *d, Volume *v) : ManagedObject(d), volume(v)
{
  volume->retain();
  commit();
}

Sampler::~Sampler() { volume->release(); }

// VdbSampler::commit, VdbSampler.cpp:37-59.
void Sampler::commit()
{
  if (!volume->committed) throw VklError("the volume of this sampler is not committed");
  const int f  = getInt("filter", volume->filter);
  const int gf = getInt("gradientFilter", hasParam("filter") ? f : volume->gradientFilter);
  const int md = getInt("maxSamplingDepth", volume->maxSamplingDepth);
  for (int x : {f, gf})
    if (x != kFilterNearest && x != kFilterLinear && x != kFilterCubic) throw VklError("invalid filter");
  view.grid                  = volume->grid;
  view.grid.maxSamplingDepth = (uint32_t)std::max(md, 0);
  view.filter                = f;
  view.gradientFilter        = gf;
  binnedMode                 = getInt("binnedSampling", -1);
  applyObservation();
  committed = true;
}

void Sampler::applyObservation()
{
  std::lock_guard<std::mutex> lock(viewMutex);
  const bool observed    = liveObservers > 0 && accessBuf.ptr;
  view.grid.leafAccess   = observed ? accessBuf.as<uint32_t>() : nullptr;
  view.grid.accessStamp  = accessEpoch;
  view.grid.packedDepth  = observed ? 0xffffffffu : volume->grid.packedDepth;
  viewVersion++;
  device->makeCurrent();
  if (!deviceView.ptr) deviceView.allocate(sizeof(SamplerView));
  deviceView.upload(&view, sizeof(view), device->stream);
}

const AttrView &Sampler::attr(uint32_t a) const
{
  if (a >= volume->attrViews.size()) throw VklError("illegal attributeIndex requested on volume");
  return volume->attrViews[a];
}

// ---- Iterator contexts --------------------------------------------------------------

IteratorContext::IteratorContext(Device *d, Sampler *s, bool hit) : ManagedObject(d), sampler(s), isHit(hit)
{
  sampler->retain();
  commit();
}

IteratorContext::~IteratorContext() { sampler->release(); }

IteratorContextView IteratorContext::refreshSampler()
{
  std::lock_guard<std::mutex> lock(viewMutex);
  std::lock_guard<std::mutex> slock(sampler->viewMutex);
  if (samplerVersion != sampler->viewVersion) {
    view.sampler   = sampler->view;
    samplerVersion = sampler->viewVersion;
    device->makeCurrent();
    deviceView.upload(&view, sizeof(IteratorContextView), device->stream);
  }
  return view;
}

namespace {
// mapToMaxIteratorDepth for VDB volumes, IteratorContext.cpp:49-174.
int mapToDepth(float hint)
{
  hint = std::max(std::min(1.f, hint), 0.f);
  static const std::pair<float, int> table[4] = {{0.f, 0}, {0.2f, 1}, {0.4f, 2}, {0.8f, 3}};
  int depth = 0;
  for (const auto &m : table) {
    if (hint >= m.first)
      depth = m.second;
    else
      break;
  }
  return depth;
}
}  // namespace

// IntervalIteratorContext::commit / HitIteratorContext::commit,
// IteratorContext.cpp:189-336.
void IteratorContext::commit()
{
  const int attrIndex = getInt("attributeIndex", 0);
  if (attrIndex < 0 || (uint32_t)attrIndex >= sampler->volume->numAttributes())
    throw VklError("illegal attributeIndex requested on volume");
  {
    std::lock_guard<std::mutex> slock(sampler->viewMutex);
    view.sampler   = sampler->view;
    samplerVersion = sampler->viewVersion;
  }
  view.attributeIndex = (uint32_t)attrIndex;
  std::vector<Range1f> ranges;
  std::vector<float> values;
  if (isHit) {
    if (Data *d = getData("values")) values = d->hostArray<float>();
    for (float v : values) ranges.push_back({v, v});
    // Sparse vdb: elementary cells (depth 3); structuredRegular: depth 2.
    view.maxIteratorDepth = (uint32_t)(sampler->volume->grid.dense ? mapToDepth(0.5f) : mapToDepth(1.f));
  } else {
    if (Data *d = getData("valueRanges")) {
      std::vector<uint8_t> b = d->hostCopy();
      ranges.resize(d->numItems);
      std::memcpy(ranges.data(), b.data(), d->numItems * sizeof(Range1f));
    }
    view.maxIteratorDepth = (uint32_t)mapToDepth(getFloat("intervalResolutionHint", 0.5f));
  }
  device->makeCurrent();
  view.numRanges    = (int32_t)ranges.size();
  view.rangesMinMax = {-std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity()};
  view.ranges       = nullptr;
  if (!ranges.empty()) {
    view.rangesMinMax = {std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity()};
    for (const Range1f &r : ranges) {
      view.rangesMinMax.lower = r.lower < view.rangesMinMax.lower ? r.lower : view.rangesMinMax.lower;
      view.rangesMinMax.upper = r.upper > view.rangesMinMax.upper ? r.upper : view.rangesMinMax.upper;
    }
    rangesBuf.allocate(ranges.size() * sizeof(Range1f));
    rangesBuf.upload(ranges.data(), ranges.size() * sizeof(Range1f), device->stream);
    view.ranges = rangesBuf.as<Range1f>();
  }
  view.numValues = (int32_t)values.size();
  view.values    = nullptr;
  if (!values.empty()) {
    valuesBuf.allocate(values.size() * sizeof(float));
    valuesBuf.upload(values.data(), values.size() * sizeof(float), device->stream);
    view.values = valuesBuf.as<float>();
  }
  deviceView.allocate(sizeof(IteratorContextView));
  deviceView.upload(&view, sizeof(IteratorContextView), device->stream);
  committed = true;
}

}  // namespace vklhip

// ---- error plumbing ------------------------------------------------------------------

namespace {

void report(Device *d, VKLError e, const char *msg)
{
  if (d)
    d->handleError(e, msg);
  else
    std::fprintf(stderr, "OPENVKL ERROR (no device): %s\n", msg);
}

#define VKL_TRY(devExpr)   \
  Device *dev_ = nullptr;  \
  try {                    \
    dev_ = (devExpr);
#define VKL_CATCH(ret)                                                      \
  }                                                                         \
  catch (const VklError &e)                                                 \
  {                                                                         \
    report(dev_, e.code, e.what());                                         \
    return ret;                                                             \
  }                                                                         \
  catch (const std::bad_alloc &)                                            \
  {                                                                         \
    report(dev_, VKL_OUT_OF_MEMORY, "Open VKL was unable to allocate memory"); \
    return ret;                                                             \
  }                                                                         \
  catch (const std::exception &e)                                           \
  {                                                                         \
    report(dev_, VKL_UNKNOWN_ERROR, e.what());                              \
    return ret;                                                             \
  }

ManagedObject *objectOf(VKLObject o)
{
  if (!o.host) throw VklError("null object handle");
  return static_cast<ManagedObject *>(o.host);
}

Device *deviceOf(VKLObject o) { return objectOf(o)->device; }

template <typename T>
T *as(VKLObject o, const char *what)
{
  T *p = dynamic_cast<T *>(objectOf(o));
  if (!p) throw VklError(std::string("object is not a ") + what);
  return p;
}

VKLObject handle(ManagedObject *o, void *devicePtr = nullptr)
{
  VKLObject h;
  h.host   = o;
  h.device = devicePtr;
  return h;
}

const float kNaN = std::numeric_limits<float>::quiet_NaN();

// Device-side view of a caller pointer: device memory is used in place; host
// memory is staged through a stream-ordered allocation.
struct Staged
{
  void *dev        = nullptr;
  bool owned       = false;
  const void *host = nullptr;
  size_t bytes     = 0;
  hipStream_t st   = nullptr;

  Staged(const void *p, size_t n, hipStream_t s, bool copyIn) : host(p), bytes(n), st(s)
  {
    if (!p || n == 0) return;
    if (isDevicePointer(p)) {
      dev = const_cast<void *>(p);
      return;
    }
    checkHip(hipMallocAsync(&dev, n, st), "hipMallocAsync");
    owned = true;
    if (copyIn) checkHip(hipMemcpyAsync(dev, p, n, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
  }
  void copyOut()
  {
    if (owned) checkHip(hipMemcpyAsync(const_cast<void *>(host), dev, bytes, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
  }
  ~Staged()
  {
    if (owned) (void)hipFreeAsync(dev, st);
  }
  template <typename T>
  T *as() const { return static_cast<T *>(dev); }
};

void syncDevice(Device *d) { checkHip(hipStreamSynchronize(d->stream), "hipStreamSynchronize"); }

// Compute units of the device (cached per HIP ordinal).
int computeUnits(Device *d)
{
  static int cached[64] = {0};
  int &c = cached[d->hipDevice & 63];
  if (!c) {
    int v = 256;
    (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, d->hipDevice);
    c = v > 0 ? v : 256;
  }
  return c;
}

Sampler *samplerOf(const VKLSampler *s)
{
  if (!s || !s->host) throw VklError("null sampler");
  Sampler *sp = static_cast<Sampler *>(s->host);
  if (!sp->committed) throw VklError("sampler is not committed");
  return sp;
}

// Binned sampling pays off once the volume is much larger than the L2 caches
// and the batch is large enough to amortise the binning pass.
constexpr uint64_t kBinnedMinPoints      = 1ull << 20;
constexpr uint64_t kBinnedMinVolumeBytes = 64ull << 20;
constexpr uint32_t kBinnedMaxBins        = 16384;  // LDS histogram of the binning kernel

int maxLdsPerBlock(int device)
{
  static int cached[64] = {0};
  int &c = cached[device & 63];
  if (!c) {
    int v = 65536;
    (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    c = v;
  }
  return c;
}

// Geometry of the binned path for this call, or false for the direct kernel.
bool binnedGeometry(Sampler *s, const AttrView &a, uint64_t n, BinOp op, BinGeometry &geo)
{
  const GridView &g = s->view.grid;
  // Temporal volumes sample through the time-interpolating source only.
  if (s->binnedMode == 0 || !g.dense || g.temporalFormat) return false;
  const uint32_t es = (uint32_t)sizeOfType((VKLDataType)a.type);
  if (s->binnedMode < 0) {
    // One record load per stencil already: the direct kernel is the better fit.
    if (g.layout.cornerPacked) return false;
    // Tricubic with cubic records: the stencil-cooperative kernel.
    if (a.cubic && (op == kBinSampleCubic || op == kBinGradientCubic)) return false;
    const uint64_t bytes = (uint64_t)g.activeSize[0] * g.activeSize[1] * g.activeSize[2] * es;
    if (n < kBinnedMinPoints || bytes < kBinnedMinVolumeBytes) return false;
  }
  // Half the LDS of a CU per brick (two bins resident per CU), at most the
  // per-workgroup limit.
  const size_t lds = std::min<size_t>((size_t)maxLdsPerBlock(s->device->hipDevice), 80u << 10);
  return chooseBinGeometry(g.activeSize, es, op, lds, kBinnedMaxBins, geo);
}

BinOp sampleOp(int filter)
{
  return filter == kFilterNearest ? kBinSampleNearest : filter == kFilterCubic ? kBinSampleCubic : kBinSampleLinear;
}

// Byte layout of a binned call's scratch at base: records, segment counters,
// overflow list, overflow count. Returns the bytes it needs (base may be null
// to only size it).
uint64_t binScratchLayout(const BinGeometry &geo, uint64_t n, uint8_t *base, BinScratch &sc)
{
  const uint64_t segs = (uint64_t)geo.numBins * 8;
  const uint64_t cap  = (n + segs - 1) / segs * 5 / 4 + 64;
  auto align          = [](uint64_t x) { return (x + 255) & ~255ull; };
  const uint64_t recBytes = align(segs * cap * 16), cntBytes = align(segs * 4), ovBytes = align(n * 4);
  sc.records       = reinterpret_cast<float4 *>(base);
  sc.segCount      = reinterpret_cast<uint32_t *>(base + recBytes);
  sc.overflow      = reinterpret_cast<uint32_t *>(base + recBytes + cntBytes);
  sc.overflowCount = reinterpret_cast<uint32_t *>(base + recBytes + cntBytes + ovBytes);
  sc.capacity      = (uint32_t)cap;
  return recBytes + cntBytes + ovBytes + 256;
}

// Scratch of a binned call (sampler-owned, grown on demand).
BinScratch binnedScratch(Sampler *s, const BinGeometry &geo, uint64_t n)
{
  BinScratch sc;
  const uint64_t total = binScratchLayout(geo, n, nullptr, sc);
  if (s->binBuf.bytes < total) s->binBuf.allocate(total);
  binScratchLayout(geo, n, s->binBuf.as<uint8_t>(), sc);
  return sc;
}

// One attribute into out[i * stride]: binned or direct. The binned kernels
// use the sampler's scratch, or with fixedScratch the caller's buffer of
// fixedBytes (calls whose scratch does not fit are rejected).
void sampleInto(Sampler *s, uint32_t attr, uint64_t n, const float *xyz, float *out, uint32_t stride,
                const float *times, uint8_t *fixedScratch = nullptr, uint64_t fixedBytes = 0)
{
  Device *d            = s->device;
  const AttrView &a    = s->attr(attr);
  const SamplerView sv = s->viewCopy();
  BinGeometry geo;
  const BinOp op = sampleOp(sv.filter);
  if (binnedGeometry(s, a, n, op, geo)) {
    if (fixedScratch) {
      BinScratch sc;
      const uint64_t need = binScratchLayout(geo, n, nullptr, sc);
      if (need > fixedBytes)
        throw VklError("binned sampling of " + std::to_string(n) + " points over " + std::to_string(geo.numBins) +
                           " bins needs " + std::to_string(need) + " bytes of scratch; the fixed scratch holds " +
                           std::to_string(fixedBytes),
                       VKL_INVALID_ARGUMENT);
      binScratchLayout(geo, n, fixedScratch, sc);
      checkHip(launchBinned(sv, a, geo, op, n, xyz, out, stride, sc, d->stream), "binned sample kernels");
      syncDevice(d);
      return;
    }
    std::lock_guard<std::mutex> lock(s->binMutex);
    const BinScratch sc = binnedScratch(s, geo, n);
    checkHip(launchBinned(sv, a, geo, op, n, xyz, out, stride, sc, d->stream), "binned sample kernels");
    // The scratch is reused by the next call: finish before releasing it.
    syncDevice(d);
    return;
  }
  checkHip(launchSampleStrided(sv, a, attr, n, xyz, out, stride, d->stream, times), "sample kernel");
}

// Host-pointer calls of at least kPipelineMinPoints points on the direct
// kernels: the points (and times) go in and the results come out in chunks of
// kPipelineChunk points, double-buffered, with the copies in on one stream,
// the kernels on a second and the copies out on a third, ordered by events;
// so the copy in of one chunk overlaps the kernel of the previous one and
// the copy out of the one before (PCIe carries both directions at once).
// Pageable host arrays pass through the device's page-locked staging slots,
// filled and drained by parallel host copies while the GPU works on the other
// slot. The call stays synchronous, and every point gets the same kernel
// arithmetic as in one launch.
constexpr uint64_t kPipelineChunk     = 1ull << 22;
constexpr uint64_t kPipelineMinPoints = 2 * kPipelineChunk;

// memcpy split over a few host threads (large copies only).
void parallelCopy(void *dst, const void *src, size_t bytes)
{
  const unsigned hw = std::thread::hardware_concurrency();
  const unsigned T  = bytes >= (8u << 20) ? std::max(1u, std::min(8u, hw)) : 1u;
  if (T <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t part = ((bytes + T - 1) / T + 63) & ~size_t(63);
  std::vector<std::thread> th;
  for (unsigned t = 1; t < T && t * part < bytes; t++)
    th.emplace_back([=]() {
      std::memcpy(static_cast<uint8_t *>(dst) + t * part, static_cast<const uint8_t *>(src) + t * part,
                  std::min(part, bytes - t * part));
    });
  std::memcpy(dst, src, std::min(part, bytes));
  for (auto &x : th) x.join();
}

// launch(count, xyz, times, out, stream) runs the kernels of one chunk on
// device arrays; xyz / times / out may each be host or device memory.
template <typename Launch>
void pipelinedHost(Device *d, uint64_t n, const float *xyz, const float *times, float *out, uint32_t outFloats,
                   const char *what, Launch launch)
{
  const bool xyzHost = !isDevicePointer(xyz), timesHost = times && !isDevicePointer(times);
  const bool outHost = !isDevicePointer(out);
  // Pageable arrays go through the page-locked staging slots.
  const bool xyzStage = xyzHost && !isPinnedHost(xyz), timesStage = timesHost && !isPinnedHost(times);
  const bool outStage = outHost && !isPinnedHost(out);
  const size_t inBytes = kPipelineChunk * 12, tmBytes = times ? kPipelineChunk * 4 : 0;
  const size_t outBytes = kPipelineChunk * 4 * (size_t)outFloats, slotBytes = inBytes + tmBytes + outBytes;
  std::unique_lock<std::mutex> staging(d->stagingMutex, std::defer_lock);
  uint8_t *pin = nullptr;
  if (xyzStage || timesStage || outStage) {
    staging.lock();
    if (d->pinnedBytes < 2 * slotBytes) {
      if (d->pinned) (void)hipHostFree(d->pinned);
      d->pinned      = nullptr;
      d->pinnedBytes = 0;
      checkHip(hipHostMalloc(&d->pinned, 2 * slotBytes, hipHostMallocDefault), "hipHostMalloc");
      d->pinnedBytes = 2 * slotBytes;
    }
    pin = static_cast<uint8_t *>(d->pinned);
  }
  hipStream_t up = nullptr, comp = nullptr, down = nullptr;
  hipEvent_t ready = nullptr, upDone[2] = {}, kernelDone[2] = {}, downDone[2] = {};
  void *buf = nullptr;
  // Results of a chunk waiting in a staging slot: (first point, count).
  uint64_t pendFirst[2] = {0, 0}, pendCount[2] = {0, 0};
  auto drain = [&](int k) {
    if (!pendCount[k]) return;
    checkHip(hipEventSynchronize(downDone[k]), "hipEventSynchronize");
    parallelCopy(out + (uint64_t)outFloats * pendFirst[k], pin + k * slotBytes + inBytes + tmBytes,
                 pendCount[k] * 4 * (size_t)outFloats);
    pendCount[k] = 0;
  };
  auto finish = [&]() {
    for (hipStream_t s : {up, comp, down})
      if (s) (void)hipStreamSynchronize(s);
    if (buf) {
      (void)hipFreeAsync(buf, d->stream);
      (void)hipStreamSynchronize(d->stream);
    }
    for (hipStream_t s : {up, comp, down})
      if (s) (void)hipStreamDestroy(s);
    for (hipEvent_t e : {ready, upDone[0], upDone[1], kernelDone[0], kernelDone[1], downDone[0], downDone[1]})
      if (e) (void)hipEventDestroy(e);
    buf = nullptr;
    up = comp = down = nullptr;
    ready = nullptr;
  };
  try {
    // The buffers, and the stages, follow the work already queued on the
    // device stream.
    checkHip(hipMallocAsync(&buf, 2 * slotBytes, d->stream), "hipMallocAsync");
    for (hipEvent_t *e : {&ready, &upDone[0], &upDone[1], &kernelDone[0], &kernelDone[1], &downDone[0], &downDone[1]})
      checkHip(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
    checkHip(hipEventRecord(ready, d->stream), "hipEventRecord");
    for (hipStream_t *s : {&up, &comp, &down}) {
      checkHip(hipStreamCreateWithFlags(s, hipStreamNonBlocking), "hipStreamCreate");
      checkHip(hipStreamWaitEvent(*s, ready, 0), "hipStreamWaitEvent");
    }
    uint64_t chunk = 0;
    for (uint64_t first = 0; first < n; first += kPipelineChunk, chunk++) {
      const uint64_t m = std::min(kPipelineChunk, n - first);
      const int k      = (int)(chunk & 1);
      uint8_t *b       = static_cast<uint8_t *>(buf) + k * slotBytes;
      uint8_t *h       = pin ? pin + k * slotBytes : nullptr;
      const float *cx  = xyz + 3 * first;
      const float *ct  = times ? times + first : nullptr;
      if (xyzHost || timesHost) {
        // This slot's device inputs are free once the kernel two chunks back
        // is done, and its staged inputs once their copy in is.
        if (chunk >= 2) {
          checkHip(hipStreamWaitEvent(up, kernelDone[k], 0), "hipStreamWaitEvent");
          if (xyzStage || timesStage) checkHip(hipEventSynchronize(upDone[k]), "hipEventSynchronize");
        }
        if (xyzHost) {
          const void *src = cx;
          if (xyzStage) {
            parallelCopy(h, cx, m * 12);
            src = h;
          }
          checkHip(hipMemcpyAsync(b, src, m * 12, hipMemcpyHostToDevice, up), "hipMemcpyAsync");
          cx = reinterpret_cast<const float *>(b);
        }
        if (timesHost) {
          const void *src = ct;
          if (timesStage) {
            parallelCopy(h + inBytes, ct, m * 4);
            src = h + inBytes;
          }
          checkHip(hipMemcpyAsync(b + inBytes, src, m * 4, hipMemcpyHostToDevice, up), "hipMemcpyAsync");
          ct = reinterpret_cast<const float *>(b + inBytes);
        }
        checkHip(hipEventRecord(upDone[k], up), "hipEventRecord");
        checkHip(hipStreamWaitEvent(comp, upDone[k], 0), "hipStreamWaitEvent");
      }
      float *co = out + (uint64_t)outFloats * first;
      float *ko = co;
      if (outHost) {
        ko = reinterpret_cast<float *>(b + inBytes + tmBytes);
        // ... and its output once the copy out two chunks back is done.
        if (chunk >= 2) checkHip(hipStreamWaitEvent(comp, downDone[k], 0), "hipStreamWaitEvent");
      }
      checkHip(launch(m, cx, ct, ko, comp), what);
      checkHip(hipEventRecord(kernelDone[k], comp), "hipEventRecord");
      if (outHost) {
        // A staged slot still holding the results two chunks back hands them
        // to the caller first.
        if (outStage) drain(k);
        checkHip(hipStreamWaitEvent(down, kernelDone[k], 0), "hipStreamWaitEvent");
        void *dst = outStage ? static_cast<void *>(h + inBytes + tmBytes) : static_cast<void *>(co);
        checkHip(hipMemcpyAsync(dst, ko, m * 4 * (size_t)outFloats, hipMemcpyDeviceToHost, down), "hipMemcpyAsync");
        checkHip(hipEventRecord(downDone[k], down), "hipEventRecord");
        if (outStage) {
          pendFirst[k] = first;
          pendCount[k] = m;
        }
      }
    }
    if (outStage) {
      // Oldest first.
      drain((int)(chunk & 1));
      drain((int)((chunk + 1) & 1));
    }
  } catch (...) {
    finish();
    throw;
  }
  finish();
}

bool pipelines(uint64_t n, const void *xyz, const void *out)
{
  return n >= kPipelineMinPoints && (!isDevicePointer(xyz) || !isDevicePointer(out));
}

// times: n floats (host or device) or null (time 0); only temporal volumes
// read them.
void sampleN(Sampler *s, uint64_t n, const vkl_vec3f *oc, float *out, uint32_t attr, const float *times)
{
  Device *d = s->device;
  d->makeCurrent();
  (void)s->attr(attr);
  if (n == 0) return;
  const bool temporal = s->view.grid.temporalFormat != 0 && times;
  if (pipelines(n, oc, out)) {
    const AttrView &a    = s->attr(attr);
    const SamplerView sv = s->viewCopy();
    BinGeometry geo;
    if (!binnedGeometry(s, a, n, sampleOp(sv.filter), geo)) {
      pipelinedHost(d, n, reinterpret_cast<const float *>(oc), temporal ? times : nullptr, out, 1, "sample kernel",
                    [&](uint64_t m, const float *x, const float *t, float *o, hipStream_t st) {
                      return launchSampleStrided(sv, a, attr, m, x, o, 1, st, t);
                    });
      return;
    }
  }
  Staged in(oc, n * 12, d->stream, true);
  Staged tm(temporal ? times : nullptr, temporal ? n * 4 : 0, d->stream, true);
  Staged o(out, n * 4, d->stream, false);
  sampleInto(s, attr, n, in.as<float>(), o.as<float>(), 1, tm.as<float>());
  o.copyOut();
  syncDevice(d);
}

void gradientN(Sampler *s, uint64_t n, const vkl_vec3f *oc, vkl_vec3f *out, uint32_t attr, const float *times)
{
  Device *d = s->device;
  d->makeCurrent();
  const AttrView &a    = s->attr(attr);
  const SamplerView sv = s->viewCopy();
  if (n == 0) return;
  const bool temporal = sv.grid.temporalFormat != 0 && times;
  BinGeometry geo;
  const int gf   = sv.gradientFilter;
  const BinOp op = gf == kFilterCubic ? kBinGradientCubic : kBinGradientLinear;
  const bool binned = gf != kFilterNearest && binnedGeometry(s, a, n, op, geo);
  if (!binned && pipelines(n, oc, out)) {
    pipelinedHost(d, n, reinterpret_cast<const float *>(oc), temporal ? times : nullptr,
                  reinterpret_cast<float *>(out), 3, "gradient kernel",
                  [&](uint64_t m, const float *x, const float *t, float *o, hipStream_t st) {
                    return launchGradient(sv, a, attr, m, x, o, st, t);
                  });
    return;
  }
  Staged in(oc, n * 12, d->stream, true);
  Staged tm(temporal ? times : nullptr, temporal ? n * 4 : 0, d->stream, true);
  Staged o(out, n * 12, d->stream, false);
  if (binned) {
    std::lock_guard<std::mutex> lock(s->binMutex);
    const BinScratch sc = binnedScratch(s, geo, n);
    checkHip(launchBinned(sv, a, geo, op, n, in.as<float>(), o.as<float>(), 1, sc, d->stream),
             "binned gradient kernels");
    // The scratch is reused by the next call: finish before releasing it.
    syncDevice(d);
  } else {
    checkHip(launchGradient(sv, a, attr, n, in.as<float>(), o.as<float>(), d->stream, tm.as<float>()),
             "gradient kernel");
  }
  o.copyOut();
  syncDevice(d);
}

void sampleMN(Sampler *s, uint64_t n, const vkl_vec3f *oc, float *out, uint32_t m, const unsigned int *attrs,
              const float *times)
{
  Device *d = s->device;
  d->makeCurrent();
  if (n == 0 || m == 0) return;
  for (uint32_t k = 0; k < m; k++) (void)s->attr(attrs[k]);
  const bool temporal = s->view.grid.temporalFormat != 0 && times;
  if (pipelines(n, oc, out)) {
    const SamplerView sv = s->viewCopy();
    bool binned          = false;
    for (uint32_t k = 0; k < m && !binned; k++) {
      BinGeometry geo;
      binned = binnedGeometry(s, s->attr(attrs[k]), n, sampleOp(sv.filter), geo);
    }
    if (!binned) {
      pipelinedHost(d, n, reinterpret_cast<const float *>(oc), temporal ? times : nullptr, out, m, "sample kernel",
                    [&](uint64_t c, const float *x, const float *t, float *o, hipStream_t st) {
                      for (uint32_t k = 0; k < m; k++) {
                        const hipError_t e = launchSampleStrided(sv, s->attr(attrs[k]), attrs[k], c, x, o + k, m, st, t);
                        if (e != hipSuccess) return e;
                      }
                      return hipSuccess;
                    });
      return;
    }
  }
  Staged in(oc, n * 12, d->stream, true);
  Staged tm(temporal ? times : nullptr, temporal ? n * 4 : 0, d->stream, true);
  Staged o(out, n * m * 4, d->stream, false);
  // Sampler.h:226-244 layout samples[i * M + k]: one specialised launch per
  // attribute (attributes may differ in element type), strided by M.
  for (uint32_t k = 0; k < m; k++) sampleInto(s, attrs[k], n, in.as<float>(), o.as<float>() + k, m, tm.as<float>());
  o.copyOut();
  syncDevice(d);
}

// SoA (width W) -> AoS for the valid lanes.
template <int W, typename VV>
std::vector<vkl_vec3f> gatherLanes(const int *valid, const VV *v, std::vector<int> &lanes)
{
  std::vector<vkl_vec3f> p;
  for (int i = 0; i < W; i++)
    if (valid[i]) {
      lanes.push_back(i);
      p.push_back({v->x[i], v->y[i], v->z[i]});
    }
  return p;
}

// Times of the valid lanes (empty when times is null).
std::vector<float> laneTimes(const float *times, const std::vector<int> &lanes)
{
  std::vector<float> t;
  if (times)
    for (int l : lanes) t.push_back(times[l]);
  return t;
}

template <int W, typename VV>
void sampleW(const int *valid, const VKLSampler *sampler, const VV *oc, float *samples, unsigned attr,
             const float *times)
{
  Sampler *s = samplerOf(sampler);
  std::vector<int> lanes;
  std::vector<vkl_vec3f> p = gatherLanes<W>(valid, oc, lanes);
  std::vector<float> r(p.size());
  const std::vector<float> t = laneTimes(times, lanes);
  sampleN(s, p.size(), p.data(), r.data(), attr, times ? t.data() : nullptr);
  for (size_t k = 0; k < lanes.size(); k++) samples[lanes[k]] = r[k];
}

template <int W, typename VV>
void sampleMW(const int *valid, const VKLSampler *sampler, const VV *oc, float *samples, unsigned m,
              const unsigned *attrs, const float *times)
{
  Sampler *s = samplerOf(sampler);
  std::vector<int> lanes;
  std::vector<vkl_vec3f> p = gatherLanes<W>(valid, oc, lanes);
  std::vector<float> r(p.size() * m);
  const std::vector<float> t = laneTimes(times, lanes);
  sampleMN(s, p.size(), p.data(), r.data(), m, attrs, times ? t.data() : nullptr);
  // Sampler.h:226-244: samples[a * W + lane] for the vector calls.
  for (size_t k = 0; k < lanes.size(); k++)
    for (unsigned a = 0; a < m; a++) samples[a * W + lanes[k]] = r[k * m + a];
}

template <int W, typename VV>
void gradientW(const int *valid, const VKLSampler *sampler, const VV *oc, VV *out, unsigned attr,
               const float *times)
{
  Sampler *s = samplerOf(sampler);
  std::vector<int> lanes;
  std::vector<vkl_vec3f> p = gatherLanes<W>(valid, oc, lanes);
  std::vector<vkl_vec3f> r(p.size());
  const std::vector<float> t = laneTimes(times, lanes);
  gradientN(s, p.size(), p.data(), r.data(), attr, times ? t.data() : nullptr);
  for (size_t k = 0; k < lanes.size(); k++) {
    out->x[lanes[k]] = r[k].x;
    out->y[lanes[k]] = r[k].y;
    out->z[lanes[k]] = r[k].z;
  }
}

// ---- scalar iterators: state in the caller's buffer, stepped on the GPU ----

constexpr uint32_t kIntervalMagic = 0x564b4c49u;  // "VKLI"
constexpr uint32_t kHitMagic      = 0x564b4c48u;  // "VKLH"

struct IntervalBuffer
{
  IteratorContext *ctx;
  uint32_t magic;
  uint32_t pad;
  IntervalState state;
};

struct HitBuffer
{
  IteratorContext *ctx;
  uint32_t magic;
  uint32_t pad;
  HitState state;
};

static_assert(sizeof(IntervalBuffer) <= VKL_MAX_INTERVAL_ITERATOR_SIZE, "interval iterator too large");
static_assert(sizeof(HitBuffer) <= VKL_MAX_HIT_ITERATOR_SIZE, "hit iterator too large");

// Per-thread device scratch for single-ray steps (4 KiB, reused).
struct StepScratch
{
  int device = -1;
  void *ptr  = nullptr;
};
thread_local StepScratch t_scratch;

uint8_t *stepScratch(Device *d)
{
  if (t_scratch.device != d->hipDevice || !t_scratch.ptr) {
    checkHip(hipMalloc(&t_scratch.ptr, 4096), "hipMalloc");
    t_scratch.device = d->hipDevice;
  }
  return static_cast<uint8_t *>(t_scratch.ptr);
}

IteratorContext *contextOf(const VKLObject *c, bool hit)
{
  if (!c || !c->host) throw VklError("null iterator context");
  IteratorContext *ctx = static_cast<IteratorContext *>(c->host);
  if (ctx->isHit != hit) throw VklError("wrong iterator context type");
  if (!ctx->committed) throw VklError("iterator context is not committed");
  return ctx;
}

void initInterval(IteratorContext *ctx, const vkl_vec3f *o, const vkl_vec3f *dir, const vkl_range1f *t, float time,
                  IntervalBuffer *buf)
{
  Device *d = ctx->device;
  d->makeCurrent();
  uint8_t *s  = stepScratch(d);
  float ray[9] = {o->x, o->y, o->z, dir->x, dir->y, dir->z, t ? t->lower : 0.f,
                  t ? t->upper : std::numeric_limits<float>::infinity(), time};
  checkHip(hipMemcpyAsync(s, ray, sizeof(ray), hipMemcpyHostToDevice, d->stream), "hipMemcpyAsync");
  (void)ctx->refreshSampler();
  checkHip(launchIntervalInit(ctx->deviceView.as<IteratorContextView>(), reinterpret_cast<IntervalState *>(s + 64),
                              reinterpret_cast<float *>(s),
                              d->stream),
           "interval init");
  checkHip(hipMemcpyAsync(&buf->state, s + 64, sizeof(IntervalState), hipMemcpyDeviceToHost, d->stream), "copy");
  syncDevice(d);
  buf->ctx   = ctx;
  buf->magic = kIntervalMagic;
}

int iterateInterval(IntervalBuffer *buf, VKLInterval *out)
{
  if (buf->magic != kIntervalMagic) throw VklError("invalid interval iterator");
  IteratorContext *ctx         = buf->ctx;
  (void)ctx->refreshSampler();
  Device *d = ctx->device;
  d->makeCurrent();
  uint8_t *s = stepScratch(d);
  // Layout: [0,64) result, [64, ...) state, [1024, ...) interval.
  checkHip(hipMemcpyAsync(s + 64, &buf->state, sizeof(IntervalState), hipMemcpyHostToDevice, d->stream), "copy");
  checkHip(launchIntervalNext(ctx->deviceView.as<IteratorContextView>(), reinterpret_cast<IntervalState *>(s + 64),
                              reinterpret_cast<IntervalOut *>(s + 1024), reinterpret_cast<int *>(s), d->stream),
           "interval step");
  struct
  {
    int result;
    IntervalOut out;
  } r;
  checkHip(hipMemcpyAsync(&buf->state, s + 64, sizeof(IntervalState), hipMemcpyDeviceToHost, d->stream), "copy");
  checkHip(hipMemcpyAsync(&r.result, s, 4, hipMemcpyDeviceToHost, d->stream), "copy");
  checkHip(hipMemcpyAsync(&r.out, s + 1024, sizeof(IntervalOut), hipMemcpyDeviceToHost, d->stream), "copy");
  syncDevice(d);
  if (r.result) std::memcpy(out, &r.out, sizeof(VKLInterval));
  return r.result;
}

void initHit(IteratorContext *ctx, const vkl_vec3f *o, const vkl_vec3f *dir, const vkl_range1f *t, float time,
             HitBuffer *buf)
{
  Device *d = ctx->device;
  d->makeCurrent();
  uint8_t *s   = stepScratch(d);
  float ray[9] = {o->x, o->y, o->z, dir->x, dir->y, dir->z, t ? t->lower : 0.f,
                  t ? t->upper : std::numeric_limits<float>::infinity(), time};
  checkHip(hipMemcpyAsync(s, ray, sizeof(ray), hipMemcpyHostToDevice, d->stream), "hipMemcpyAsync");
  (void)ctx->refreshSampler();
  checkHip(launchHitInit(ctx->deviceView.as<IteratorContextView>(), reinterpret_cast<HitState *>(s + 64), reinterpret_cast<float *>(s), d->stream),
           "hit init");
  checkHip(hipMemcpyAsync(&buf->state, s + 64, sizeof(HitState), hipMemcpyDeviceToHost, d->stream), "copy");
  syncDevice(d);
  buf->ctx   = ctx;
  buf->magic = kHitMagic;
}

int iterateHit(HitBuffer *buf, VKLHit *out)
{
  if (buf->magic != kHitMagic) throw VklError("invalid hit iterator");
  IteratorContext *ctx         = buf->ctx;
  const IteratorContextView cv = ctx->refreshSampler();
  Device *d                    = ctx->device;
  d->makeCurrent();
  uint8_t *s = stepScratch(d);
  checkHip(hipMemcpyAsync(s + 64, &buf->state, sizeof(HitState), hipMemcpyHostToDevice, d->stream), "copy");
  checkHip(launchHitNext(cv, ctx->deviceView.as<IteratorContextView>(),
                         ctx->sampler->attr(cv.attributeIndex), reinterpret_cast<HitState *>(s + 64), reinterpret_cast<HitOut *>(s + 1024),
                         reinterpret_cast<int *>(s), d->stream),
           "hit step");
  int result = 0;
  HitOut h;
  checkHip(hipMemcpyAsync(&buf->state, s + 64, sizeof(HitState), hipMemcpyDeviceToHost, d->stream), "copy");
  checkHip(hipMemcpyAsync(&result, s, 4, hipMemcpyDeviceToHost, d->stream), "copy");
  checkHip(hipMemcpyAsync(&h, s + 1024, sizeof(HitOut), hipMemcpyDeviceToHost, d->stream), "copy");
  syncDevice(d);
  if (result) std::memcpy(out, &h, sizeof(VKLHit));
  return result;
}

template <int W, typename VV, typename VR>
void initIntervalW(const int *valid, const VKLIntervalIteratorContext *c, const VV *o, const VV *dir, const VR *t,
                   const float *times, void *buffer)
{
  IteratorContext *ctx = contextOf(c, false);
  IntervalBuffer *b    = static_cast<IntervalBuffer *>(buffer);
  for (int i = 0; i < W; i++) {
    if (!valid[i]) {
      b[i].magic = 0;
      continue;
    }
    const vkl_vec3f oi = {o->x[i], o->y[i], o->z[i]}, di = {dir->x[i], dir->y[i], dir->z[i]};
    const vkl_range1f ti = {t->lower[i], t->upper[i]};
    initInterval(ctx, &oi, &di, &ti, times ? times[i] : 0.f, &b[i]);
  }
}

template <int W, typename VI>
void iterateIntervalW(const int *valid, void *it, VI *out, int *result)
{
  IntervalBuffer *b = static_cast<IntervalBuffer *>(it);
  for (int i = 0; i < W; i++) {
    result[i] = 0;
    if (!valid[i] || b[i].magic != kIntervalMagic) continue;
    VKLInterval iv;
    result[i] = iterateInterval(&b[i], &iv);
    if (result[i]) {
      out->tRange.lower[i]     = iv.tRange.lower;
      out->tRange.upper[i]     = iv.tRange.upper;
      out->valueRange.lower[i] = iv.valueRange.lower;
      out->valueRange.upper[i] = iv.valueRange.upper;
      out->nominalDeltaT[i]    = iv.nominalDeltaT;
    }
  }
}

template <int W, typename VV, typename VR>
void initHitW(const int *valid, const VKLHitIteratorContext *c, const VV *o, const VV *dir, const VR *t,
              const float *times, void *buffer)
{
  IteratorContext *ctx = contextOf(c, true);
  HitBuffer *b         = static_cast<HitBuffer *>(buffer);
  for (int i = 0; i < W; i++) {
    if (!valid[i]) {
      b[i].magic = 0;
      continue;
    }
    const vkl_vec3f oi = {o->x[i], o->y[i], o->z[i]}, di = {dir->x[i], dir->y[i], dir->z[i]};
    const vkl_range1f ti = {t->lower[i], t->upper[i]};
    initHit(ctx, &oi, &di, &ti, times ? times[i] : 0.f, &b[i]);
  }
}

template <int W, typename VH>
void iterateHitW(const int *valid, void *it, VH *out, int *result)
{
  HitBuffer *b = static_cast<HitBuffer *>(it);
  for (int i = 0; i < W; i++) {
    result[i] = 0;
    if (!valid[i] || b[i].magic != kHitMagic) continue;
    VKLHit h;
    result[i] = iterateHit(&b[i], &h);
    if (result[i]) {
      out->t[i]       = h.t;
      out->sample[i]  = h.sample;
      out->epsilon[i] = h.epsilon;
    }
  }
}

}  // namespace

// =====================================================================================
// C ABI
// =====================================================================================

extern "C" {

// ---- device -------------------------------------------------------------------------

OPENVKL_INTERFACE void vklInit(void) {}

OPENVKL_INTERFACE VKLDevice vklNewDevice(const char *name)
{
  VKL_TRY(nullptr)
  if (!name) throw VklError("null deviceName provided to vklNewDevice");
  const std::string n(name);
  if (n != "hip" && n != "default")
    throw VklError("unknown device '" + n + "' (this library provides the \"hip\" device)");
  return reinterpret_cast<VKLDevice>(new Device());
  VKL_CATCH(nullptr)
}

#define DEV(d) reinterpret_cast<Device *>(d)

OPENVKL_INTERFACE void vklDeviceSetLogCallback(VKLDevice d, VKLLogCallback cb, void *user)
{
  if (!d) return;
  DEV(d)->logCallback = cb;
  DEV(d)->logUserData = user;
}

OPENVKL_INTERFACE void vklDeviceSetErrorCallback(VKLDevice d, VKLErrorCallback cb, void *user)
{
  if (!d) return;
  DEV(d)->errorCallback = cb;
  DEV(d)->errorUserData = user;
}

OPENVKL_INTERFACE void vklDeviceSetInt(VKLDevice d, const char *name, int x)
{
  if (!d || !name) return;
  ManagedObject::Param p;
  p.type = VKL_INT;
  p.bytes.assign(reinterpret_cast<uint8_t *>(&x), reinterpret_cast<uint8_t *>(&x) + 4);
  DEV(d)->params[name] = p;
}

OPENVKL_INTERFACE void vklDeviceSetString(VKLDevice d, const char *name, const char *s)
{
  if (!d || !name || !s) return;
  ManagedObject::Param p;
  p.type = VKL_STRING;
  p.str  = s;
  DEV(d)->params[name] = p;
}

OPENVKL_INTERFACE void vklDeviceSetVoidPtr(VKLDevice d, const char *name, const void *a)
{
  if (!d || !name) return;
  ManagedObject::Param p;
  p.type = VKL_VOID_PTR;
  p.bytes.assign(reinterpret_cast<const uint8_t *>(&a), reinterpret_cast<const uint8_t *>(&a) + sizeof(a));
  DEV(d)->params[name] = p;
}

OPENVKL_INTERFACE void vklCommitDevice(VKLDevice d)
{
  VKL_TRY(DEV(d))
  if (!dev_) throw VklError("null device");
  dev_->commit();
  VKL_CATCH()
}

OPENVKL_INTERFACE VKLError vklDeviceGetLastErrorCode(VKLDevice d) { return d ? DEV(d)->lastErrorCode : VKL_INVALID_ARGUMENT; }

OPENVKL_INTERFACE const char *vklDeviceGetLastErrorMsg(VKLDevice d)
{
  return d ? DEV(d)->lastErrorMessage.c_str() : "null device";
}

OPENVKL_INTERFACE int vklGetNativeSIMDWidth(VKLDevice d) { return d ? 64 : 0; }

OPENVKL_INTERFACE void vklReleaseDevice(VKLDevice d)
{
  if (d && --DEV(d)->refs == 0) delete DEV(d);
}

// ---- objects --------------------------------------------------------------------------

OPENVKL_INTERFACE void vklCommit(VKLObject o)
{
  VKL_TRY(deviceOf(o))
  objectOf(o)->commit();
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklRelease(VKLObject o)
{
  VKL_TRY(deviceOf(o))
  objectOf(o)->release();
  VKL_CATCH()
}

OPENVKL_INTERFACE VKLData vklNewData(VKLDevice d, size_t numItems, VKLDataType type, const void *source,
                                     VKLDataCreationFlags flags, size_t byteStride)
{
  VKL_TRY(DEV(d))
  if (!dev_) throw VklError("null device provided to vklNewData");
  if (!dev_->committed) throw VklError("You must commit the device before using it!");
  dev_->makeCurrent();
  Data *data = new Data(dev_, numItems, type, source, flags, byteStride);
  VKLData h;
  h.host   = data;
  h.device = nullptr;
  return h;
  VKL_CATCH(VKLData())
}

OPENVKL_INTERFACE VKLVolume vklNewVolume(VKLDevice d, const char *type)
{
  VKL_TRY(DEV(d))
  if (!dev_) throw VklError("null device provided to vklNewVolume");
  if (!dev_->committed) throw VklError("You must commit the device before using it!");
  if (!type) throw VklError("null type provided to vklNewVolume");
  const std::string t(type);
  Volume *v = nullptr;
  if (t == "structuredRegular")
    v = new DenseVolume(dev_);
  else if (t == "vdb")
    v = new VdbVolume(dev_);
  else
    throw VklError("could not create volume of type '" + t + "' (the hip device supports structuredRegular, vdb)");
  VKLVolume h;
  h.host   = v;
  h.device = nullptr;
  return h;
  VKL_CATCH(VKLVolume())
}

OPENVKL_INTERFACE vkl_box3f vklGetBoundingBox(VKLVolume v)
{
  vkl_box3f b;
  std::memset(&b, 0, sizeof(b));
  VKL_TRY(deviceOf(v))
  Volume *vol = as<Volume>(v, "volume");
  std::memcpy(&b, vol->bbox, sizeof(b));
  return b;
  VKL_CATCH(b)
}

OPENVKL_INTERFACE void vklGetBoundingBoxRef(const VKLVolume *v, vkl_box3f *b) { *b = vklGetBoundingBox(*v); }

OPENVKL_INTERFACE unsigned int vklGetNumAttributes(VKLVolume v)
{
  VKL_TRY(deviceOf(v))
  return as<Volume>(v, "volume")->numAttributes();
  VKL_CATCH(0)
}

OPENVKL_INTERFACE unsigned int vklGetNumAttributesRef(const VKLVolume *v) { return vklGetNumAttributes(*v); }

OPENVKL_INTERFACE vkl_range1f vklGetValueRange(VKLVolume v, unsigned int a)
{
  vkl_range1f r = {kNaN, kNaN};
  VKL_TRY(deviceOf(v))
  Volume *vol = as<Volume>(v, "volume");
  if (a >= vol->valueRanges.size()) throw VklError("illegal attributeIndex requested on volume");
  r.lower = vol->valueRanges[a].lower;
  r.upper = vol->valueRanges[a].upper;
  return r;
  VKL_CATCH(r)
}

OPENVKL_INTERFACE void vklGetValueRangeRef(const VKLVolume *v, unsigned int a, vkl_range1f *r)
{
  *r = vklGetValueRange(*v, a);
}

OPENVKL_INTERFACE VKLSampler vklNewSampler(VKLVolume v)
{
  VKL_TRY(deviceOf(v))
  Volume *vol = as<Volume>(v, "volume");
  Sampler *s  = new Sampler(dev_, vol);
  VKLSampler h;
  h.host   = s;
  h.device = s->deviceView.ptr;
  return h;
  VKL_CATCH(VKLSampler())
}

OPENVKL_INTERFACE VKLFeatureFlags vklGetFeatureFlags(VKLSampler) { return (VKLFeatureFlags)VKL_FEATURE_FLAGS_DEFAULT; }

// ---- observers (openvkl/api/API.cpp:126-186) -----------------------------------------

OPENVKL_INTERFACE VKLObserver vklNewVolumeObserver(VKLVolume v, const char *type)
{
  VKL_TRY(deviceOf(v))
  if (!type) throw VklError("null type provided to vklNewVolumeObserver");
  Volume *vol = as<Volume>(v, "volume");
  Observer *o = nullptr;
  if (dynamic_cast<VdbVolume *>(vol) && std::string(type) == "InnerNode") {
    if (!vol->committed) throw VklError("Trying to create an observer on a vdb volume that was not committed.");
    o = new InnerNodeObserver(dev_, vol);
  }
  if (!o) throw VklError(std::string("unsupported observer type: ") + type);
  VKLObserver h;
  h.host   = o;
  h.device = nullptr;
  return h;
  VKL_CATCH(VKLObserver())
}

OPENVKL_INTERFACE VKLObserver vklNewSamplerObserver(VKLSampler s, const char *type)
{
  VKL_TRY(deviceOf(s))
  if (!type) throw VklError("null type provided to vklNewSamplerObserver");
  Sampler *sp = as<Sampler>(s, "sampler");
  Observer *o = nullptr;
  if (dynamic_cast<VdbVolume *>(sp->volume) && std::string(type) == "LeafNodeAccess") o = new LeafAccessObserver(dev_, sp);
  if (!o) throw VklError(std::string("unsupported observer type: ") + type);
  VKLObserver h;
  h.host   = o;
  h.device = nullptr;
  return h;
  VKL_CATCH(VKLObserver())
}

OPENVKL_INTERFACE const void *vklMapObserver(VKLObserver o)
{
  VKL_TRY(deviceOf(o))
  return as<Observer>(o, "observer")->map();
  VKL_CATCH(nullptr)
}

OPENVKL_INTERFACE void vklUnmapObserver(VKLObserver o)
{
  VKL_TRY(deviceOf(o))
  as<Observer>(o, "observer")->unmap();
  VKL_CATCH()
}

OPENVKL_INTERFACE VKLDataType vklGetObserverElementType(VKLObserver o)
{
  VKL_TRY(deviceOf(o))
  return as<Observer>(o, "observer")->elementType();
  VKL_CATCH(VKL_UNKNOWN)
}

OPENVKL_INTERFACE size_t vklGetObserverElementSize(VKLObserver o)
{
  VKL_TRY(deviceOf(o))
  return as<Observer>(o, "observer")->elementSize();
  VKL_CATCH(0)
}

OPENVKL_INTERFACE size_t vklGetObserverNumElements(VKLObserver o)
{
  VKL_TRY(deviceOf(o))
  return as<Observer>(o, "observer")->numElements();
  VKL_CATCH(0)
}

OPENVKL_INTERFACE VKLIntervalIteratorContext vklNewIntervalIteratorContext(VKLSampler s)
{
  VKL_TRY(deviceOf(s))
  IteratorContext *c = new IteratorContext(dev_, as<Sampler>(s, "sampler"), false);
  VKLIntervalIteratorContext h;
  h.host   = c;
  h.device = nullptr;
  return h;
  VKL_CATCH(VKLIntervalIteratorContext())
}

OPENVKL_INTERFACE VKLHitIteratorContext vklNewHitIteratorContext(VKLSampler s)
{
  VKL_TRY(deviceOf(s))
  IteratorContext *c = new IteratorContext(dev_, as<Sampler>(s, "sampler"), true);
  VKLHitIteratorContext h;
  h.host   = c;
  h.device = nullptr;
  return h;
  VKL_CATCH(VKLHitIteratorContext())
}

// ---- parameters ------------------------------------------------------------------------

OPENVKL_INTERFACE void vklSetParam(VKLObject o, const char *name, VKLDataType type, const void *mem)
{
  VKL_TRY(deviceOf(o))
  if (!name) throw VklError("null parameter name");
  objectOf(o)->setParam(name, type, mem);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklSetBool(VKLObject o, const char *name, int b)
{
  const bool v = b != 0;
  vklSetParam(o, name, VKL_BOOL, &v);
}
OPENVKL_INTERFACE void vklSetFloat(VKLObject o, const char *name, float x) { vklSetParam(o, name, VKL_FLOAT, &x); }
OPENVKL_INTERFACE void vklSetInt(VKLObject o, const char *name, int x) { vklSetParam(o, name, VKL_INT, &x); }
OPENVKL_INTERFACE void vklSetVec3f(VKLObject o, const char *name, float x, float y, float z)
{
  const float v[3] = {x, y, z};
  vklSetParam(o, name, VKL_VEC3F, v);
}
OPENVKL_INTERFACE void vklSetVec3i(VKLObject o, const char *name, int x, int y, int z)
{
  const int v[3] = {x, y, z};
  vklSetParam(o, name, VKL_VEC3I, v);
}
OPENVKL_INTERFACE void vklSetData(VKLObject o, const char *name, VKLData d) { vklSetParam(o, name, VKL_DATA, &d); }
OPENVKL_INTERFACE void vklSetString(VKLObject o, const char *name, const char *s) { vklSetParam(o, name, VKL_STRING, s); }
OPENVKL_INTERFACE void vklSetVoidPtr(VKLObject o, const char *name, void *v) { vklSetParam(o, name, VKL_VOID_PTR, &v); }

// ---- sampling ----------------------------------------------------------------------------

OPENVKL_INTERFACE float vklComputeSample(const VKLSampler *s, const vkl_vec3f *oc, unsigned int a, float time)
{
  VKL_TRY(samplerOf(s)->device)
  float r = kNaN;
  sampleN(samplerOf(s), 1, oc, &r, a, &time);
  return r;
  VKL_CATCH(kNaN)
}

OPENVKL_INTERFACE void vklComputeSample4(const int *valid, const VKLSampler *s, const vkl_vvec3f4 *oc, float *out,
                                         unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  sampleW<4>(valid, s, oc, out, a, times);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeSample8(const int *valid, const VKLSampler *s, const vkl_vvec3f8 *oc, float *out,
                                         unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  sampleW<8>(valid, s, oc, out, a, times);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeSample16(const int *valid, const VKLSampler *s, const vkl_vvec3f16 *oc, float *out,
                                          unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  sampleW<16>(valid, s, oc, out, a, times);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklComputeSampleN(const VKLSampler *s, unsigned int N, const vkl_vec3f *oc, float *out,
                                         unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  sampleN(samplerOf(s), N, oc, out, a, times);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklComputeSampleM(const VKLSampler *s, const vkl_vec3f *oc, float *out, unsigned int M,
                                         const unsigned int *attrs, float time)
{
  VKL_TRY(samplerOf(s)->device)
  sampleMN(samplerOf(s), 1, oc, out, M, attrs, &time);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeSampleM4(const int *valid, const VKLSampler *s, const vkl_vvec3f4 *oc, float *out,
                                          unsigned int M, const unsigned int *attrs, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  sampleMW<4>(valid, s, oc, out, M, attrs, times);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeSampleM8(const int *valid, const VKLSampler *s, const vkl_vvec3f8 *oc, float *out,
                                          unsigned int M, const unsigned int *attrs, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  sampleMW<8>(valid, s, oc, out, M, attrs, times);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeSampleM16(const int *valid, const VKLSampler *s, const vkl_vvec3f16 *oc, float *out,
                                           unsigned int M, const unsigned int *attrs, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  sampleMW<16>(valid, s, oc, out, M, attrs, times);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeSampleMN(const VKLSampler *s, unsigned int N, const vkl_vec3f *oc, float *out,
                                          unsigned int M, const unsigned int *attrs, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  sampleMN(samplerOf(s), N, oc, out, M, attrs, times);
  VKL_CATCH()
}

OPENVKL_INTERFACE vkl_vec3f vklComputeGradient(const VKLSampler *s, const vkl_vec3f *oc, unsigned int a, float time)
{
  vkl_vec3f r = {kNaN, kNaN, kNaN};
  VKL_TRY(samplerOf(s)->device)
  gradientN(samplerOf(s), 1, oc, &r, a, &time);
  return r;
  VKL_CATCH(r)
}
OPENVKL_INTERFACE void vklComputeGradient4(const int *valid, const VKLSampler *s, const vkl_vvec3f4 *oc,
                                           vkl_vvec3f4 *g, unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  gradientW<4>(valid, s, oc, g, a, times);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeGradient8(const int *valid, const VKLSampler *s, const vkl_vvec3f8 *oc,
                                           vkl_vvec3f8 *g, unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  gradientW<8>(valid, s, oc, g, a, times);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeGradient16(const int *valid, const VKLSampler *s, const vkl_vvec3f16 *oc,
                                            vkl_vvec3f16 *g, unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  gradientW<16>(valid, s, oc, g, a, times);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklComputeGradientN(const VKLSampler *s, unsigned int N, const vkl_vec3f *oc, vkl_vec3f *g,
                                           unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  gradientN(samplerOf(s), N, oc, g, a, times);
  VKL_CATCH()
}

// ---- interval iterators ---------------------------------------------------------------------

OPENVKL_INTERFACE size_t vklGetIntervalIteratorSize(const VKLIntervalIteratorContext *) { return sizeof(IntervalBuffer); }
OPENVKL_INTERFACE size_t vklGetIntervalIteratorSize4(const VKLIntervalIteratorContext *) { return 4 * sizeof(IntervalBuffer); }
OPENVKL_INTERFACE size_t vklGetIntervalIteratorSize8(const VKLIntervalIteratorContext *) { return 8 * sizeof(IntervalBuffer); }
OPENVKL_INTERFACE size_t vklGetIntervalIteratorSize16(const VKLIntervalIteratorContext *) { return 16 * sizeof(IntervalBuffer); }

OPENVKL_INTERFACE VKLIntervalIterator vklInitIntervalIterator(const VKLIntervalIteratorContext *c, const vkl_vec3f *o,
                                                              const vkl_vec3f *d, const vkl_range1f *t, float time,
                                                              void *buffer)
{
  VKL_TRY(contextOf(c, false)->device)
  if (!buffer) throw VklError("null iterator buffer");
  initInterval(contextOf(c, false), o, d, t, time, static_cast<IntervalBuffer *>(buffer));
  return reinterpret_cast<VKLIntervalIterator>(buffer);
  VKL_CATCH(nullptr)
}

OPENVKL_INTERFACE VKLIntervalIterator4 vklInitIntervalIterator4(const int *valid, const VKLIntervalIteratorContext *c,
                                                                const vkl_vvec3f4 *o, const vkl_vvec3f4 *d,
                                                                const vkl_vrange1f4 *t, const float *times, void *buffer)
{
  VKL_TRY(contextOf(c, false)->device)
  initIntervalW<4>(valid, c, o, d, t, times, buffer);
  return reinterpret_cast<VKLIntervalIterator4>(buffer);
  VKL_CATCH(nullptr)
}
OPENVKL_INTERFACE VKLIntervalIterator8 vklInitIntervalIterator8(const int *valid, const VKLIntervalIteratorContext *c,
                                                                const vkl_vvec3f8 *o, const vkl_vvec3f8 *d,
                                                                const vkl_vrange1f8 *t, const float *times, void *buffer)
{
  VKL_TRY(contextOf(c, false)->device)
  initIntervalW<8>(valid, c, o, d, t, times, buffer);
  return reinterpret_cast<VKLIntervalIterator8>(buffer);
  VKL_CATCH(nullptr)
}
OPENVKL_INTERFACE VKLIntervalIterator16 vklInitIntervalIterator16(const int *valid, const VKLIntervalIteratorContext *c,
                                                                  const vkl_vvec3f16 *o, const vkl_vvec3f16 *d,
                                                                  const vkl_vrange1f16 *t, const float *times,
                                                                  void *buffer)
{
  VKL_TRY(contextOf(c, false)->device)
  initIntervalW<16>(valid, c, o, d, t, times, buffer);
  return reinterpret_cast<VKLIntervalIterator16>(buffer);
  VKL_CATCH(nullptr)
}

OPENVKL_INTERFACE int vklIterateInterval(VKLIntervalIterator it, VKLInterval *out)
{
  IntervalBuffer *b = reinterpret_cast<IntervalBuffer *>(it);
  VKL_TRY(b && b->magic == kIntervalMagic ? b->ctx->device : nullptr)
  if (!dev_) throw VklError("invalid interval iterator");
  return iterateInterval(b, out);
  VKL_CATCH(0)
}
OPENVKL_INTERFACE void vklIterateInterval4(const int *valid, VKLIntervalIterator4 it, VKLInterval4 *out, int *result)
{
  VKL_TRY(nullptr)
  iterateIntervalW<4>(valid, it, out, result);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklIterateInterval8(const int *valid, VKLIntervalIterator8 it, VKLInterval8 *out, int *result)
{
  VKL_TRY(nullptr)
  iterateIntervalW<8>(valid, it, out, result);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklIterateInterval16(const int *valid, VKLIntervalIterator16 it, VKLInterval16 *out, int *result)
{
  VKL_TRY(nullptr)
  iterateIntervalW<16>(valid, it, out, result);
  VKL_CATCH()
}

// ---- hit iterators ---------------------------------------------------------------------------

OPENVKL_INTERFACE size_t vklGetHitIteratorSize(const VKLHitIteratorContext *) { return sizeof(HitBuffer); }
OPENVKL_INTERFACE size_t vklGetHitIteratorSize4(const VKLHitIteratorContext *) { return 4 * sizeof(HitBuffer); }
OPENVKL_INTERFACE size_t vklGetHitIteratorSize8(const VKLHitIteratorContext *) { return 8 * sizeof(HitBuffer); }
OPENVKL_INTERFACE size_t vklGetHitIteratorSize16(const VKLHitIteratorContext *) { return 16 * sizeof(HitBuffer); }

OPENVKL_INTERFACE VKLHitIterator vklInitHitIterator(const VKLHitIteratorContext *c, const vkl_vec3f *o,
                                                    const vkl_vec3f *d, const vkl_range1f *t, float time, void *buffer)
{
  VKL_TRY(contextOf(c, true)->device)
  if (!buffer) throw VklError("null iterator buffer");
  initHit(contextOf(c, true), o, d, t, time, static_cast<HitBuffer *>(buffer));
  return reinterpret_cast<VKLHitIterator>(buffer);
  VKL_CATCH(nullptr)
}
OPENVKL_INTERFACE VKLHitIterator4 vklInitHitIterator4(const int *valid, const VKLHitIteratorContext *c,
                                                      const vkl_vvec3f4 *o, const vkl_vvec3f4 *d,
                                                      const vkl_vrange1f4 *t, const float *times, void *buffer)
{
  VKL_TRY(contextOf(c, true)->device)
  initHitW<4>(valid, c, o, d, t, times, buffer);
  return reinterpret_cast<VKLHitIterator4>(buffer);
  VKL_CATCH(nullptr)
}
OPENVKL_INTERFACE VKLHitIterator8 vklInitHitIterator8(const int *valid, const VKLHitIteratorContext *c,
                                                      const vkl_vvec3f8 *o, const vkl_vvec3f8 *d,
                                                      const vkl_vrange1f8 *t, const float *times, void *buffer)
{
  VKL_TRY(contextOf(c, true)->device)
  initHitW<8>(valid, c, o, d, t, times, buffer);
  return reinterpret_cast<VKLHitIterator8>(buffer);
  VKL_CATCH(nullptr)
}
OPENVKL_INTERFACE VKLHitIterator16 vklInitHitIterator16(const int *valid, const VKLHitIteratorContext *c,
                                                        const vkl_vvec3f16 *o, const vkl_vvec3f16 *d,
                                                        const vkl_vrange1f16 *t, const float *times, void *buffer)
{
  VKL_TRY(contextOf(c, true)->device)
  initHitW<16>(valid, c, o, d, t, times, buffer);
  return reinterpret_cast<VKLHitIterator16>(buffer);
  VKL_CATCH(nullptr)
}

OPENVKL_INTERFACE int vklIterateHit(VKLHitIterator it, VKLHit *out)
{
  HitBuffer *b = reinterpret_cast<HitBuffer *>(it);
  VKL_TRY(b && b->magic == kHitMagic ? b->ctx->device : nullptr)
  if (!dev_) throw VklError("invalid hit iterator");
  return iterateHit(b, out);
  VKL_CATCH(0)
}
OPENVKL_INTERFACE void vklIterateHit4(const int *valid, VKLHitIterator4 it, VKLHit4 *out, int *result)
{
  VKL_TRY(nullptr)
  iterateHitW<4>(valid, it, out, result);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklIterateHit8(const int *valid, VKLHitIterator8 it, VKLHit8 *out, int *result)
{
  VKL_TRY(nullptr)
  iterateHitW<8>(valid, it, out, result);
  VKL_CATCH()
}
OPENVKL_INTERFACE void vklIterateHit16(const int *valid, VKLHitIterator16 it, VKLHit16 *out, int *result)
{
  VKL_TRY(nullptr)
  iterateHitW<16>(valid, it, out, result);
  VKL_CATCH()
}

// ---- HIP-device extensions (openvkl/openvkl_hip.h) ----------------------------------------------

OPENVKL_INTERFACE void vklComputeSampleNAsync(const VKLSampler *s, unsigned int N, const vkl_vec3f *oc, float *out,
                                              unsigned int a, void *stream)
{
  VKL_TRY(samplerOf(s)->device)
  Sampler *sp = samplerOf(s);
  dev_->makeCurrent();
  checkHip(launchSample(sp->viewCopy(), sp->attr(a), a, N, reinterpret_cast<const float *>(oc), out,
                        static_cast<hipStream_t>(stream)),
           "sample kernel");
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklComputeSampleNNoAlloc(const VKLSampler *s, unsigned int N, const vkl_vec3f *oc,
                                                float *out, unsigned int a, const float *times)
{
  VKL_TRY(samplerOf(s)->device)
  Sampler *sp = samplerOf(s);
  dev_->makeCurrent();
  (void)sp->attr(a);
  if (N == 0) return;
  const bool temporal = sp->viewCopy().grid.temporalFormat != 0 && times;
  if (!isDevicePointer(oc) || !isDevicePointer(out) || (temporal && !isDevicePointer(times)))
    throw VklError("vklComputeSampleNNoAlloc takes device pointers only (host arrays would need staging buffers)",
                   VKL_INVALID_ARGUMENT);
  const NoAllocScratch fixed = noAllocScratch(dev_->hipDevice);
  std::lock_guard<std::mutex> lock(*fixed.lock);
  sampleInto(sp, a, N, reinterpret_cast<const float *>(oc), out, 1, temporal ? times : nullptr, fixed.ptr,
             fixed.bytes);
  syncDevice(dev_);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklComputeRayMarchN(const VKLSampler *s, unsigned int numRays, const vkl_vec3f *origins,
                                           const vkl_vec3f *directions, const vkl_range1f *tRanges,
                                           unsigned int numSteps, float *integrals, unsigned int a)
{
  VKL_TRY(samplerOf(s)->device)
  Sampler *sp = samplerOf(s);
  dev_->makeCurrent();
  const AttrView &av = sp->attr(a);
  if (numRays == 0) return;
  if (numSteps == 0) throw VklError("numSteps must be positive", VKL_INVALID_ARGUMENT);
  if (!origins || !directions || !tRanges || !integrals) throw VklError("null ray array", VKL_INVALID_ARGUMENT);
  hipStream_t st = dev_->stream;
  Staged o(origins, numRays * 12ull, st, true), d(directions, numRays * 12ull, st, true);
  Staged t(tRanges, numRays * 8ull, st, true);
  Staged out(integrals, numRays * 4ull, st, false);
  checkHip(launchRayMarch(sp->viewCopy(), av, a, numRays, o.as<float>(), d.as<float>(), t.as<float>(), numSteps,
                          out.as<float>(), st),
           "ray march kernel");
  out.copyOut();
  syncDevice(dev_);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklComputeWoodcockN(const VKLSampler *s, unsigned int numRays, const vkl_vec3f *origins,
                                           const vkl_vec3f *directions, const vkl_range1f *tRanges,
                                           unsigned int *rngStates, float sigmaMax, const float *colorsAndOpacities,
                                           unsigned int numTfValues, vkl_range1f tfValueRange, float time,
                                           float shutter, unsigned int maxCollisions, unsigned int a, float *eventT,
                                           float *eventSamples, unsigned int *numCollisions)
{
  VKL_TRY(samplerOf(s)->device)
  Sampler *sp = samplerOf(s);
  dev_->makeCurrent();
  const AttrView &av = sp->attr(a);
  if (numRays == 0) return;
  if (!origins || !directions || !tRanges || !rngStates || !eventT || !eventSamples)
    throw VklError("null ray array", VKL_INVALID_ARGUMENT);
  if (numTfValues && !colorsAndOpacities) throw VklError("null transfer function", VKL_INVALID_ARGUMENT);
  hipStream_t st = dev_->stream;
  Staged o(origins, numRays * 12ull, st, true), d(directions, numRays * 12ull, st, true);
  Staged t(tRanges, numRays * 8ull, st, true), tf(colorsAndOpacities, numTfValues * 16ull, st, true);
  Staged rng(rngStates, numRays * 8ull, st, true);
  Staged outT(eventT, numRays * 4ull, st, false), outS(eventSamples, numRays * 4ull, st, false);
  Staged cnt(numCollisions, numCollisions ? numRays * 4ull : 0, st, false);
  WoodcockArgs w{};
  w.numRays       = numRays;
  w.maxCollisions = maxCollisions;
  w.tfSize        = numTfValues;
  w.sigmaMax      = sigmaMax;
  w.time          = time;
  w.shutter       = shutter;
  w.tfLo          = tfValueRange.lower;
  w.tfHi          = tfValueRange.upper;
  w.org           = o.as<float>();
  w.dir           = d.as<float>();
  w.tRange        = t.as<float>();
  w.tf            = tf.as<float>();
  w.rng           = rng.as<uint32_t>();
  w.eventT        = outT.as<float>();
  w.samples       = outS.as<float>();
  w.count         = cnt.as<uint32_t>();
  checkHip(launchWoodcock(sp->viewCopy(), av, a, w, st), "woodcock kernel");
  rng.copyOut();
  outT.copyOut();
  outS.copyOut();
  cnt.copyOut();
  syncDevice(dev_);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklComputeGradientNAsync(const VKLSampler *s, unsigned int N, const vkl_vec3f *oc,
                                                vkl_vec3f *out, unsigned int a, void *stream)
{
  VKL_TRY(samplerOf(s)->device)
  Sampler *sp = samplerOf(s);
  dev_->makeCurrent();
  checkHip(launchGradient(sp->viewCopy(), sp->attr(a), a, N, reinterpret_cast<const float *>(oc),
                          reinterpret_cast<float *>(out), static_cast<hipStream_t>(stream)),
           "gradient kernel");
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklIterateIntervalsBatch(const VKLIntervalIteratorContext *c, unsigned int numRays,
                                                const vkl_vec3f *origins, const vkl_vec3f *directions,
                                                const vkl_range1f *tRanges, const float *times, VKLInterval *intervals,
                                                unsigned int maxIntervalsPerRay, unsigned int *numIntervals)
{
  VKL_TRY(contextOf(c, false)->device)
  (void)times;
  IteratorContext *ctx = contextOf(c, false);
  dev_->makeCurrent();
  if (numRays == 0) return;
  hipStream_t st = dev_->stream;
  Staged o(origins, numRays * 12ull, st, true), d(directions, numRays * 12ull, st, true);
  Staged t(tRanges, tRanges ? numRays * 8ull : 0, st, true);
  Staged out(intervals, (uint64_t)numRays * maxIntervalsPerRay * sizeof(VKLInterval), st, false);
  Staged cnt(numIntervals, numRays * 4ull, st, false);
  const IteratorContextView hv = ctx->refreshSampler();
  checkHip(launchIntervalBatch(hv.maxIteratorDepth, denseQueryOf(hv.sampler.grid), ctx->deviceView.as<IteratorContextView>(), numRays, o.as<float>(), d.as<float>(), t.as<float>(),
                               out.as<IntervalOut>(), maxIntervalsPerRay, cnt.as<uint32_t>(), st),
           "interval batch");
  out.copyOut();
  cnt.copyOut();
  syncDevice(dev_);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklIterateHitsBatch(const VKLHitIteratorContext *c, unsigned int numRays,
                                           const vkl_vec3f *origins, const vkl_vec3f *directions,
                                           const vkl_range1f *tRanges, const float *times, VKLHit *hits,
                                           unsigned int maxHitsPerRay, unsigned int *numHits)
{
  VKL_TRY(contextOf(c, true)->device)
  IteratorContext *ctx = contextOf(c, true);
  dev_->makeCurrent();
  if (numRays == 0) return;
  hipStream_t st = dev_->stream;
  Staged o(origins, numRays * 12ull, st, true), d(directions, numRays * 12ull, st, true);
  Staged t(tRanges, tRanges ? numRays * 8ull : 0, st, true);
  Staged tm(times, times ? numRays * 4ull : 0, st, true);
  Staged out(hits, (uint64_t)numRays * maxHitsPerRay * sizeof(VKLHit), st, false);
  Staged cnt(numHits, numRays * 4ull, st, false);
  const IteratorContextView cv = ctx->refreshSampler();
  std::lock_guard<std::mutex> lock(ctx->batchMutex);
  const uint32_t lanes = batchLaneCount(numRays, computeUnits(dev_));
  if (!ctx->batchWork.ptr) ctx->batchWork.allocate(sizeof(uint32_t));
  checkHip(launchSetWork(ctx->batchWork.as<uint32_t>(), lanes, st), "batch work counter");
  checkHip(launchHitBatch(cv, ctx->deviceView.as<IteratorContextView>(), ctx->sampler->attr(cv.attributeIndex), numRays,
                          o.as<float>(), d.as<float>(), t.as<float>(), tm.as<float>(), out.as<HitOut>(), maxHitsPerRay,
                          cnt.as<uint32_t>(), ctx->batchWork.as<uint32_t>(), lanes, st),
           "hit batch");
  out.copyOut();
  cnt.copyOut();
  syncDevice(dev_);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklComputeRayMarchIntervalsN(const VKLIntervalIteratorContext *c, unsigned int numRays,
                                                    const vkl_vec3f *origins, const vkl_vec3f *directions,
                                                    const vkl_range1f *tRanges, float time, float samplingRate,
                                                    const float *colorsAndOpacities, unsigned int numTfValues,
                                                    vkl_range1f tfValueRange, unsigned int maxSamplesPerRay,
                                                    float *rgba, unsigned int *numSamples)
{
  VKL_TRY(contextOf(c, false)->device)
  IteratorContext *ctx = contextOf(c, false);
  dev_->makeCurrent();
  if (numRays == 0) return;
  if (!origins || !directions || !rgba) throw VklError("null ray array", VKL_INVALID_ARGUMENT);
  if (numTfValues && !colorsAndOpacities) throw VklError("null transfer function", VKL_INVALID_ARGUMENT);
  hipStream_t st = dev_->stream;
  Staged o(origins, numRays * 12ull, st, true), d(directions, numRays * 12ull, st, true);
  Staged t(tRanges, tRanges ? numRays * 8ull : 0, st, true);
  Staged tf(colorsAndOpacities, numTfValues * 16ull, st, true);
  Staged out(rgba, numRays * 16ull, st, false);
  Staged cnt(numSamples, numSamples ? numRays * 4ull : 0, st, false);
  RayMarchIntervalsArgs m{};
  m.numRays      = numRays;
  m.tfSize       = numTfValues;
  m.maxSamples   = maxSamplesPerRay;
  m.time         = time;
  m.samplingRate = samplingRate;
  m.tfLo         = tfValueRange.lower;
  m.tfHi         = tfValueRange.upper;
  m.org          = o.as<float>();
  m.dir          = d.as<float>();
  m.tRanges      = t.as<float>();
  m.tf           = tf.as<float>();
  m.rgba         = out.as<float>();
  m.count        = cnt.as<uint32_t>();
  const IteratorContextView cv = ctx->refreshSampler();
  checkHip(launchRayMarchIntervals(cv, ctx->deviceView.as<IteratorContextView>(), ctx->sampler->attr(cv.attributeIndex),
                                   m, st),
           "interval ray march");
  out.copyOut();
  cnt.copyOut();
  syncDevice(dev_);
  VKL_CATCH()
}

OPENVKL_INTERFACE void vklComputeHitRenderN(const VKLHitIteratorContext *c, unsigned int numRays,
                                            const vkl_vec3f *origins, const vkl_vec3f *directions,
                                            const vkl_range1f *tRanges, float time, const float *colorsAndOpacities,
                                            unsigned int numTfValues, vkl_range1f tfValueRange,
                                            unsigned int maxHitsPerRay, float *rgba, unsigned int *numHits)
{
  VKL_TRY(contextOf(c, true)->device)
  IteratorContext *ctx = contextOf(c, true);
  dev_->makeCurrent();
  if (numRays == 0) return;
  if (!origins || !directions || !rgba) throw VklError("null ray array", VKL_INVALID_ARGUMENT);
  if (numTfValues && !colorsAndOpacities) throw VklError("null transfer function", VKL_INVALID_ARGUMENT);
  hipStream_t st = dev_->stream;
  Staged o(origins, numRays * 12ull, st, true), d(directions, numRays * 12ull, st, true);
  Staged t(tRanges, tRanges ? numRays * 8ull : 0, st, true);
  Staged tf(colorsAndOpacities, numTfValues * 16ull, st, true);
  Staged out(rgba, numRays * 16ull, st, false);
  Staged cnt(numHits, numHits ? numRays * 4ull : 0, st, false);
  HitRenderArgs m{};
  m.numRays = numRays;
  m.tfSize  = numTfValues;
  m.maxHits = maxHitsPerRay;
  m.time    = time;
  m.tfLo    = tfValueRange.lower;
  m.tfHi    = tfValueRange.upper;
  m.org     = o.as<float>();
  m.dir     = d.as<float>();
  m.tRanges = t.as<float>();
  m.tf      = tf.as<float>();
  m.rgba    = out.as<float>();
  m.count   = cnt.as<uint32_t>();
  const IteratorContextView cv = ctx->refreshSampler();
  checkHip(launchHitRender(cv, ctx->deviceView.as<IteratorContextView>(), ctx->sampler->attr(cv.attributeIndex), m,
                           st),
           "hit render");
  out.copyOut();
  cnt.copyOut();
  syncDevice(dev_);
  VKL_CATCH()
}

OPENVKL_INTERFACE int vklHipGetVolumeLevel(VKLVolume v, unsigned int level, uint64_t *numNodes,
                                           const uint64_t **voxels, const vkl_range1f **valueRanges)
{
  VKL_TRY(deviceOf(v))
  Volume *vol = as<Volume>(v, "volume");
  if (level > 2) throw VklError("level must be 0, 1 or 2");
  if (!vol->committed) throw VklError("volume is not committed");
  if (numNodes) *numNodes = vol->grid.levels[level].numNodes;
  if (voxels) *voxels = vol->hostVoxels[level].data();
  if (valueRanges) *valueRanges = reinterpret_cast<const vkl_range1f *>(vol->hostRanges[level].data());
  return 1;
  VKL_CATCH(0)
}

}  // extern "C"
