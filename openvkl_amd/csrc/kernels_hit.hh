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

// Persistent ray schedule of the batched iterators. Each lane of a
// wavefront iterates one ray at a time; when at least kRefill lanes of the
// wavefront have no ray, the wavefront takes that many new rays from the
// work counter with one atomic (ballot, then a rank per idle lane), so lanes
// whose rays end early do not wait for the longest ray of their wavefront.
// The counter starts at the number of launched lanes (lane i starts on ray
// i) and only grows; a wavefront ends once the counter has passed numRays
// and none of its lanes holds a ray. Every ray's results are those of the
// one-ray-per-lane schedule.
constexpr uint32_t kRefill = 16;

struct RaySchedule
{
  uint32_t r;
  bool has, fresh, drained;

  VKL_DEV explicit RaySchedule(uint32_t numRays)
      : r(blockIdx.x * kIterBlock + threadIdx.x), has(r < numRays), fresh(r < numRays), drained(false)
  {
  }

  // Called by every lane of the wavefront after each step; false once the
  // wavefront is done.
  VKL_DEV bool refill(uint32_t numRays, uint32_t *work)
  {
    const uint64_t idle   = __ballot(!has);
    const uint32_t nIdle  = (uint32_t)__popcll(idle);
    if (!drained && nIdle >= kRefill) {
      const int leader = __ffsll((unsigned long long)idle) - 1;
      uint32_t base    = 0;
      if ((int)__lane_id() == leader) base = atomicAdd(work, nIdle);
      base    = (uint32_t)__shfl((int)base, leader);
      drained = (uint64_t)base + nIdle >= numRays;
      if (!has) {
        const uint32_t k = base + (uint32_t)__popcll(idle & ((1ull << __lane_id()) - 1ull));
        if (k < numRays) {
          r     = k;
          has   = true;
          fresh = true;
        }
      }
    }
    return !(drained && __ballot(has) == 0);
  }
};

// Lanes a persistent batch launch starts with (a multiple of the
// wavefront), and the value its work counter starts at.
inline uint32_t batchLanes(uint32_t numRays, int cus)
{
  const uint64_t want = ((uint64_t)numRays + kIterBlock - 1) / kIterBlock;
  const uint64_t cap  = (uint64_t)(cus > 0 ? cus : 256) * 16;
  return (uint32_t)(want < cap ? want : cap) * kIterBlock;
}

// The hit kernels read their context from device memory and sample through
// a PointSampler specialised on the volume's source type and filter (chosen
// on the host from the same context); temporal sources sample at the ray time.
template <typename Src, int kFilter>
__global__ __launch_bounds__(1) void kHitNext(const IteratorContextView *__restrict__ cp, AttrView a, HitState *s, HitOut *out,
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

// Hit iteration waits on dependent loads, so occupancy pays: asking for 4
// waves per SIMD (128 VGPRs instead of ~155 at 3 waves) costs a ~30-byte
// spill and measured 13 % faster on the 4096^3 vdb. Cubic filters keep the
// compiler's choice.
template <typename Src, int kFilter, int kQuery>
__global__ __launch_bounds__(kIterBlock) __attribute__((amdgpu_waves_per_eu(kFilter == kFilterCubic ? 1 : 4)))
void kHitBatch(const IteratorContextView *__restrict__ cp, AttrView a,
                                                         uint32_t numRays, const float *__restrict__ org,
                                                         const float *__restrict__ dir,
                                                         const float *__restrict__ tRanges,
                                                         const float *__restrict__ times, HitOut *__restrict__ out,
                                                         uint32_t maxOut, uint32_t *__restrict__ count,
                                                         uint32_t *__restrict__ work)
{
  const IteratorContextView &c = *cp;
  PointSampler<Src, kFilter> smp(c.sampler.grid, a, c.attributeIndex);
  RaySchedule q(numRays);
  HitState st;
  uint32_t n = 0;
  do {
    if (q.fresh) {
      const uint32_t r = q.r;
      hitInit<hitDepthOf<Src>()>(c, st, F3{org[3 * r], org[3 * r + 1], org[3 * r + 2]},
                                 F3{dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]}, rayRange(tRanges, r),
                                 times ? times[r] : 0.f);
      smp.setTime(st.time);
      n       = 0;
      q.fresh = false;
    }
    if (q.has) {
      HitOut o;
      if (n < maxOut && hitNext<decltype(smp), kQuery>(c, smp, st, o)) {
        out[(uint64_t)q.r * maxOut + n] = o;
        n++;
      } else {
        count[q.r] = n;
        q.has      = false;
      }
    }
  } while (q.refill(numRays, work));
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

// work: the batch's work counter, already set to `lanes` (batchLanes).
struct HitBatchLaunch
{
  const IteratorContextView *c;
  AttrView a;
  uint32_t numRays;
  const float *org, *dir, *tRanges, *times;
  HitOut *out;
  uint32_t maxOut;
  uint32_t *count;
  uint32_t *work;
  uint32_t lanes;
  hipStream_t st;
  // vdb sources run the node-table query only (kQuery 0); structuredRegular
  // ones choose the query per context at run time.
  template <typename Src, int kFilter>
  hipError_t operator()() const
  {
    constexpr int kQuery = IsDenseSource<Src>::value ? -1 : 0;
    hipLaunchKernelGGL((kHitBatch<Src, kFilter, kQuery>), dim3(lanes / kIterBlock), dim3(kIterBlock), 0, st, c, a,
                       numRays, org, dir, tRanges, times, out, maxOut, count, work);
    return hipGetLastError();
  }
};

}  // namespace vklhip
