// Hit-iterator renderer kernels for structuredRegular sources.
#include "kernels_hitrender.hh"

namespace vklhip {

hipError_t launchHitRender(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                           const HitRenderArgs &m, hipStream_t st)
{
  if (m.numRays == 0) return hipSuccess;
  const GridView &g = hostView.sampler.grid;
  if (!g.dense || g.temporalFormat) return launchHitRenderSparse(hostView, c, a, m, st);
  return dispatchSource<kDenseKinds>(g, a, hostView.sampler.filter, HitRenderLaunch{c, a, m, st, denseQueryOf(hostView.sampler.grid)});
}

}  // namespace vklhip
