// Batched sampling: many volumes with ragged point batches. The batches of a
// plan are grouped per filter: all structuredRegular batches with 32-bit
// voxel offsets in one group of any element types, every other voxel source
// type in a group of its own. Each group is one persistent launch over its
// flattened point space, split into tiles of `tile` points that workgroups
// stride over; a workgroup finds the range of batches of a tile once (binary
// search over the group's batch offsets, staged in LDS), each work-item its
// batch within that range. Results equal one
// vklComputeSampleN per batch: the same per-point device functions run.

#include "batch.h"
#include "kernels_dispatch.hh"
#include "vkl_device.hh"

namespace vklhip {

constexpr uint32_t kMaxLdsItems = 2048;

__device__ __forceinline__ uint32_t findItem(const uint64_t *begins, uint32_t n, uint64_t p)
{
  // last i with begins[i] <= p
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (begins[mid] <= p)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Sample i of one batch.
template <typename Src, int kFilter>
VKL_DEV void samplePoint(const BatchItem &it, uint64_t i)
{
  const PointSampler<Src, kFilter> smp(it.sampler.grid, it.attr, it.attrIndex);
  const F3 oc = {__builtin_nontemporal_load(it.xyz + 3 * i), __builtin_nontemporal_load(it.xyz + 3 * i + 1),
                 __builtin_nontemporal_load(it.xyz + 3 * i + 2)};
  __builtin_nontemporal_store(smp(oc), it.out + i);
}

// The workgroup's loop over tiles of `tile` points of the flattened point
// space. The range of batches a tile overlaps is found once per tile; each
// work-item then finds its point's batch within that range. A wave samples
// its points one batch at a time (the batch of its first active lane, then
// the next), so the batch descriptor is uniform: scalar loads and registers.
// A wave spans one or two batches unless the batches hold fewer points than a
// wave. Point(item, i) samples point i of a batch; it has one call site, so
// the sampler is inlined once.
template <typename Point>
VKL_DEV void forEachTilePoint(const BatchItem *__restrict__ items, const uint64_t *__restrict__ begins,
                              uint32_t numItems, uint64_t total, uint32_t tile, const Point &point)
{
  __shared__ uint64_t sBegins[kMaxLdsItems];
  const bool lds = numItems <= kMaxLdsItems;
  if (lds) {
    for (uint32_t i = threadIdx.x; i < numItems; i += blockDim.x) sBegins[i] = begins[i];
    __syncthreads();
  }
  const uint64_t *b = lds ? sBegins : begins;
  for (uint64_t base = (uint64_t)blockIdx.x * tile; base < total; base += (uint64_t)gridDim.x * tile) {
    const uint64_t end = base + tile < total ? base + tile : total;
    const uint32_t k0  = findItem(b, numItems, base);
    const uint32_t n   = findItem(b, numItems, end - 1) - k0 + 1;
    for (uint64_t p = base + threadIdx.x; p < end; p += blockDim.x) {
      const uint32_t k = k0 + findItem(b + k0, n, p);
      for (;;) {
        const uint32_t ku = __builtin_amdgcn_readfirstlane(k);
        if (ku == k) {
          const BatchItem &it = items[ku];
          point(it, p - it.begin);
          break;
        }
      }
    }
  }
}

// One specialisation group (one voxel source type and filter).
template <typename Src, int kFilter>
__global__ void kSampleBatch(const BatchItem *__restrict__ items, const uint64_t *__restrict__ begins,
                             uint32_t numItems, uint64_t total, uint32_t tile)
{
  forEachTilePoint(items, begins, numItems, total, tile,
                   [](const BatchItem &it, uint64_t i) { samplePoint<Src, kFilter>(it, i); });
}

// Every structuredRegular batch with 32-bit voxel offsets and this filter, of
// any element type, in one launch: the element type is chosen per batch.
template <int kFilter>
__global__ void kSampleBatchDense(const BatchItem *__restrict__ items, const uint64_t *__restrict__ begins,
                                  uint32_t numItems, uint64_t total, uint32_t tile)
{
  forEachTilePoint(items, begins, numItems, total, tile, [](const BatchItem &it, uint64_t i) {
    switch (it.attr.type) {
    case kTypeUchar: samplePoint<DenseSource<uint8_t, false>, kFilter>(it, i); break;
    case kTypeShort: samplePoint<DenseSource<int16_t, false>, kFilter>(it, i); break;
    case kTypeUshort: samplePoint<DenseSource<uint16_t, false>, kFilter>(it, i); break;
    case kTypeHalf: samplePoint<DenseSource<__half, false>, kFilter>(it, i); break;
    case kTypeDouble: samplePoint<DenseSource<double, false>, kFilter>(it, i); break;
    default: samplePoint<DenseSource<float, false>, kFilter>(it, i); break;
    }
  });
}

namespace {
struct BatchGroupLaunch
{
  const BatchGroup &grp;
  const BatchLaunch &cfg;
  hipStream_t st;
  template <typename Src, int kFilter>
  hipError_t operator()() const
  {
    hipLaunchKernelGGL((kSampleBatch<Src, kFilter>), dim3(cfg.blocks), dim3(cfg.threads), 0, st, grp.items,
                       grp.begins, grp.numItems, grp.total, cfg.tile);
    return hipGetLastError();
  }
};
}  // namespace

hipError_t launchSampleBatch(const BatchGroup *groups, uint32_t numGroups, const BatchLaunch &cfg,
                             hipStream_t stream)
{
  for (uint32_t g = 0; g < numGroups; g++) {
    const BatchGroup &grp = groups[g];
    if (grp.total == 0 || grp.numItems == 0) continue;
    hipError_t e = hipSuccess;
    if (grp.mergedDense) {
      const dim3 grid(cfg.blocks), block(cfg.threads);
      if (grp.filter == kFilterNearest)
        hipLaunchKernelGGL(kSampleBatchDense<kFilterNearest>, grid, block, 0, stream, grp.items, grp.begins,
                           grp.numItems, grp.total, cfg.tile);
      else if (grp.filter == kFilterCubic)
        hipLaunchKernelGGL(kSampleBatchDense<kFilterCubic>, grid, block, 0, stream, grp.items, grp.begins,
                           grp.numItems, grp.total, cfg.tile);
      else
        hipLaunchKernelGGL(kSampleBatchDense<kFilterLinear>, grid, block, 0, stream, grp.items, grp.begins,
                           grp.numItems, grp.total, cfg.tile);
      e = hipGetLastError();
    } else {
      e = dispatchSource(grp.grid, grp.attr, grp.filter, BatchGroupLaunch{grp, cfg, stream});
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace vklhip
