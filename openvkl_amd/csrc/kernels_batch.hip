// Batched sampling: many volumes with ragged point batches in one persistent
// launch. The flattened point space [0, total) is split into tiles of
// `tile` points; workgroups stride over tiles; each work-item finds its batch
// by binary search over the batch prefix offsets (staged in LDS).
//
// Results equal one vklComputeSampleN per batch: the same per-point device
// functions run (dense linear float / half inline, everything else through
// the runtime-dispatched sampler).

#include "batch.h"
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

template <typename T>
__device__ __forceinline__ float denseLinear(const BatchItem &it, F3 oc)
{
  const F3 p = xfmPoint(it.sampler.grid.o2i, oc);
  const DenseSource<T, false> src(it.sampler.grid, it.attr, it.attrIndex);
  return sampleLinear(src, p);
}

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
      const uint32_t k   = findItem(b, numItems, p);
      const BatchItem &it = items[k];
      const uint64_t i   = p - it.begin;
      const F3 oc        = {it.xyz[3 * i], it.xyz[3 * i + 1], it.xyz[3 * i + 2]};
      float r;
      if (it.fastPath == kBatchDenseLinearF32)
        r = denseLinear<float>(it, oc);
      else if (it.fastPath == kBatchDenseLinearF16)
        r = denseLinear<__half>(it, oc);
      else
        r = sampleObject(it.sampler, it.attrIndex, oc);
      it.out[i] = r;
    }
  }
}

hipError_t launchSampleBatch(const BatchItem *items, const uint64_t *begins, uint32_t numItems, uint64_t total,
                             const BatchLaunch &cfg, hipStream_t stream)
{
  if (total == 0 || numItems == 0) return hipSuccess;
  hipLaunchKernelGGL(kSampleBatch, dim3(cfg.blocks), dim3(cfg.threads), 0, stream, items, begins, numItems, total,
                     cfg.tile);
  return hipGetLastError();
}

}  // namespace vklhip
