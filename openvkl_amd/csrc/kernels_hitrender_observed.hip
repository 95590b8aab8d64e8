// Hit-iterator renderer kernels for vdb samplers with a LeafNodeAccess
// observer.
#include "kernels_hitrender.hh"

namespace vklhip {

hipError_t launchHitRenderObserved(const IteratorContextView &hostView, const IteratorContextView *c,
                                   const AttrView &a, const HitRenderArgs &m, hipStream_t st)
{
  return dispatchSource<kSparseObservedKinds>(hostView.sampler.grid, a, hostView.sampler.filter,
                                              HitRenderLaunch{c, a, m, st, denseQueryOf(hostView.sampler.grid)});
}

}  // namespace vklhip
