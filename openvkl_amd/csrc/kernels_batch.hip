// Batched sampling: many volumes with ragged point batches. The batches of a
// plan are grouped by sampler specialisation (voxel source type and filter);
// each group is one persistent launch over its flattened point space, split
// into tiles of `tile` points that workgroups stride over; each work-item
// finds its batch by binary search over the group's batch offsets (staged in
// LDS). Results equal one vklComputeSampleN per batch: the same per-point
// device functions run.

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

template <typename Src, int kFilter>
__global__ void kSampleBatch(const BatchItem *__restrict__ items, const uint64_t *__restrict__ begins,
                             uint32_t numItems, uint64_t total, uint32_t tile)
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
    for (uint64_t p = base + threadIdx.x; p < end; p += blockDim.x) {
      const uint32_t k    = findItem(b, numItems, p);
      const BatchItem &it = items[k];
      const uint64_t i    = p - it.begin;
      const F3 oc         = {it.xyz[3 * i], it.xyz[3 * i + 1], it.xyz[3 * i + 2]};
      const PointSampler<Src, kFilter> smp(it.sampler.grid, it.attr, it.attrIndex);
      it.out[i] = smp(oc);
    }
  }
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
    const hipError_t e = dispatchSource(grp.grid, grp.attr, grp.filter, BatchGroupLaunch{grp, cfg, stream});
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace vklhip
