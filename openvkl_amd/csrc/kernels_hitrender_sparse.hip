// Hit-iterator renderer kernels for vdb sources; temporal and observed vdb
// sources have units of their own, so the three compile in parallel.
#include "kernels_hitrender.hh"

namespace vklhip {

hipError_t launchHitRenderSparse(const IteratorContextView &hostView, const IteratorContextView *c,
                                 const AttrView &a, const HitRenderArgs &m, hipStream_t st)
{
  const GridView &g = hostView.sampler.grid;
  if (g.temporalFormat) return launchHitRenderTemporal(hostView, c, a, m, st);
  if (g.leafAccess) return launchHitRenderObserved(hostView, c, a, m, st);
  return dispatchSource<kSparseKinds>(g, a, hostView.sampler.filter, HitRenderLaunch{c, a, m, st, denseQueryOf(hostView.sampler.grid)});
}

}  // namespace vklhip
