// Interval ray-march compositing kernels for vdb and temporal sources.
#include "kernels_rmi.hh"

namespace vklhip {

hipError_t launchRayMarchIntervalsSparse(const IteratorContextView &hostView, const IteratorContextView *c,
                                         const AttrView &a, const RayMarchIntervalsArgs &m, hipStream_t st)
{
  return dispatchSource<kSparseKinds | kTemporalKinds | kSparseObservedKinds>(
      hostView.sampler.grid, a, hostView.sampler.filter, RayMarchIntervalsLaunch{c, a, m, hostView.maxIteratorDepth, st, denseQueryOf(hostView.sampler.grid)});
}

}  // namespace vklhip
