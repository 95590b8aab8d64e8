// Hit iteration over vdb volumes: the kernels of kernels_hit.hh instantiated
// for unobserved vdb sources, in a translation unit of their own so that they
// compile beside the others; temporal sources go to kernels_hit_temporal.hip.

#include "kernels.h"
#include "kernels_hit.hh"

namespace vklhip {

hipError_t launchHitNextOther(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                              HitState *state, HitOut *out, int *result, hipStream_t st)
{
  if (hostView.sampler.grid.temporalFormat) return launchHitNextTemporal(hostView, c, a, state, out, result, st);
  return dispatchSource<kSparseKinds>(hostView.sampler.grid, a, hostView.sampler.filter,
                                      HitNextLaunch{c, a, state, out, result, st});
}

hipError_t launchHitBatchOther(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                               uint32_t numRays, const float *org, const float *dir, const float *tRanges,
                               const float *times, HitOut *out, uint32_t maxOut, uint32_t *count, uint32_t *work, uint32_t lanes, hipStream_t st)
{
  if (numRays == 0) return hipSuccess;
  if (hostView.sampler.grid.temporalFormat)
    return launchHitBatchTemporal(hostView, c, a, numRays, org, dir, tRanges, times, out, maxOut, count, work, lanes, st);
  return dispatchSource<kSparseKinds>(hostView.sampler.grid, a, hostView.sampler.filter,
                                      HitBatchLaunch{c, a, numRays, org, dir, tRanges, times, out, maxOut, count, work, lanes,
                                                     st});
}

}  // namespace vklhip
