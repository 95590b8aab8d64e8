// Hit-iteration kernels (one specialisation per voxel source type and
// filter) and the functors that launch them; instantiated by kernels_iter.hip
// (integer structuredRegular voxels), kernels_hit_dense.hip (half / float /
// double), kernels_hit_other.hip (temporal and vdb sources) and
// kernels_hit_observed.hip (observed vdb samplers), so that the four sets
// compile in parallel.
#pragma once

#include "kernels_dispatch.hh"
#include "vkl_device.hh"

namespace vklhip {

constexpr int kIterBlock = 64;

VKL_DEV Range1f rayRange(const float *tRanges, uint32_t r)
{
  if (!tRanges) return {0.f, vinf()};
  return {tRanges[2 * r], tRanges[2 * r + 1]};
}

// The hit kernels read their context from device memory and sample through
// a PointSampler specialised on the volume's source type and filter (chosen
// on the host from the same context); temporal sources sample at the ray time.
template <typename Src, int kFilter>
__global__ void kHitNext(const IteratorContextView *__restrict__ cp, AttrView a, HitState *s, HitOut *out,
                         int *result)
{
  const IteratorContextView &c = *cp;
  PointSampler<Src, kFilter> smp(c.sampler.grid, a, c.attributeIndex);
  HitState st = *s;
  smp.setTime(st.time);
  HitOut o;
  const bool r = hitNext(c, smp, st, o);
  *s           = st;
  if (r) *out = o;
  *result = r ? 1 : 0;
}

template <typename Src, int kFilter>
__global__ __launch_bounds__(kIterBlock) void kHitBatch(const IteratorContextView *__restrict__ cp, AttrView a,
                                                         uint32_t numRays, const float *__restrict__ org,
                                                         const float *__restrict__ dir,
                                                         const float *__restrict__ tRanges,
                                                         const float *__restrict__ times, HitOut *__restrict__ out,
                                                         uint32_t maxOut, uint32_t *__restrict__ count)
{
  const uint32_t r = blockIdx.x * kIterBlock + threadIdx.x;
  if (r >= numRays) return;
  const IteratorContextView &c = *cp;
  PointSampler<Src, kFilter> smp(c.sampler.grid, a, c.attributeIndex);
  HitState st;
  hitInit(c, st, F3{org[3 * r], org[3 * r + 1], org[3 * r + 2]}, F3{dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]},
          rayRange(tRanges, r), times ? times[r] : 0.f);
  smp.setTime(st.time);
  uint32_t n = 0;
  HitOut o;
  HitOut *dst = out + (uint64_t)r * maxOut;
  while (n < maxOut && hitNext(c, smp, st, o)) dst[n++] = o;
  count[r] = n;
}

struct HitNextLaunch
{
  const IteratorContextView *c;
  AttrView a;
  HitState *state;
  HitOut *out;
  int *result;
  hipStream_t st;
  template <typename Src, int kFilter>
  hipError_t operator()() const
  {
    hipLaunchKernelGGL((kHitNext<Src, kFilter>), dim3(1), dim3(1), 0, st, c, a, state, out, result);
    return hipGetLastError();
  }
};

struct HitBatchLaunch
{
  const IteratorContextView *c;
  AttrView a;
  uint32_t numRays;
  const float *org, *dir, *tRanges, *times;
  HitOut *out;
  uint32_t maxOut;
  uint32_t *count;
  hipStream_t st;
  template <typename Src, int kFilter>
  hipError_t operator()() const
  {
    hipLaunchKernelGGL((kHitBatch<Src, kFilter>), dim3((numRays + kIterBlock - 1) / kIterBlock), dim3(kIterBlock), 0,
                       st, c, a, numRays, org, dir, tRanges, times, out, maxOut, count);
    return hipGetLastError();
  }
};

}  // namespace vklhip
