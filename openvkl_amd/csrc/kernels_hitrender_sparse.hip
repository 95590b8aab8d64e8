// Hit-iterator renderer kernels for vdb and temporal sources.
#include "kernels_hitrender.hh"

namespace vklhip {

hipError_t launchHitRenderSparse(const IteratorContextView &hostView, const IteratorContextView *c,
                                 const AttrView &a, const HitRenderArgs &m, hipStream_t st)
{
  return dispatchSource<kSparseKinds | kTemporalKinds | kSparseObservedKinds>(
      hostView.sampler.grid, a, hostView.sampler.filter, HitRenderLaunch{c, a, m, st});
}

}  // namespace vklhip
