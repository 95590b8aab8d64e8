// Interval ray-march compositing kernels for structuredRegular sources.
#include "kernels_rmi.hh"

namespace vklhip {

hipError_t launchRayMarchIntervals(const IteratorContextView &hostView, const IteratorContextView *c,
                                   const AttrView &a, const RayMarchIntervalsArgs &m, hipStream_t st)
{
  if (m.numRays == 0) return hipSuccess;
  const GridView &g = hostView.sampler.grid;
  if (!g.dense || g.temporalFormat) return launchRayMarchIntervalsSparse(hostView, c, a, m, st);
  return dispatchSource<kDenseKinds>(g, a, hostView.sampler.filter, RayMarchIntervalsLaunch{c, a, m, hostView.maxIteratorDepth, st, denseQueryOf(hostView.sampler.grid)});
}

}  // namespace vklhip
