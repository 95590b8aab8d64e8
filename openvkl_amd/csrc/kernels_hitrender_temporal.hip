// Hit-iterator renderer kernels for temporal (motion-blur) sources, dense and
// vdb.
#include "kernels_hitrender.hh"

namespace vklhip {

hipError_t launchHitRenderTemporal(const IteratorContextView &hostView, const IteratorContextView *c,
                                   const AttrView &a, const HitRenderArgs &m, hipStream_t st)
{
  return dispatchSource<kTemporalKinds>(hostView.sampler.grid, a, hostView.sampler.filter,
                                        HitRenderLaunch{c, a, m, st, denseQueryOf(hostView.sampler.grid)});
}

}  // namespace vklhip
