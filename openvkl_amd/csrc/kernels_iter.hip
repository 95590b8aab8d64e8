// Majorant build and iterator kernels.
//
//  * kLeafRanges: per-leaf value range over the filter-support window
//    (VdbSampler.ispc:460-557), the input of the majorant grid.
//  * interval / hit iterators: single-ray steps for the per-ray C API and
//    batched kernels (one ray per work-item) for vklIterate*Batch.

#include "kernels.h"
#include "kernels_dispatch.hh"
#include "kernels_hit.hh"
#include "vkl_device.hh"

namespace vklhip {

constexpr int kRangeBlock = 128;

// ---- majorants ----------------------------------------------------------------

// Value range of one neighbour node's voxel sub-window (VdbSampler.ispc:383-455).
__device__ Range1f nodeRange(const GridView &g, const AttrView &a, uint32_t attr, uint32_t ox, uint32_t oy,
                             uint32_t oz, const uint32_t *rx, const uint32_t *ry, const uint32_t *rz)
{
  Range1f r = {vinf(), -vinf()};
  const bool inDom = ox < g.activeSize[0] && oy < g.activeSize[1] && oz < g.activeSize[2];
  if (g.dense) {
    if (!inDom) return {a.background, a.background};
    for (uint32_t x = rx[0]; x < rx[1]; ++x)
      for (uint32_t y = ry[0]; y < ry[1]; ++y)
        for (uint32_t z = rz[0]; z < rz[1]; ++z) {
          const uint32_t px = ox + x, py = oy + y, pz = oz + z;
          if (!(px < g.activeSize[0] && py < g.activeSize[1] && pz < g.activeSize[2])) continue;
          if (g.temporalFormat) {
            // every time sample of the voxel (VdbSampler_dense.ih, temporal
            // value ranges)
            const uint64_t v = ((uint64_t)pz * g.activeSize[1] + py) * g.activeSize[0] + px;
            uint64_t b, e;
            if (g.temporalFormat == 1u) {
              b = v * g.numTimesteps;
              e = b + g.numTimesteps;
            } else if (g.tuIndex64) {
              b = ((const uint64_t *)g.tuIndices)[v];
              e = ((const uint64_t *)g.tuIndices)[v + 1];
            } else {
              b = ((const uint32_t *)g.tuIndices)[v];
              e = ((const uint32_t *)g.tuIndices)[v + 1];
            }
            for (uint64_t k = b; k < e; ++k) {
              const float tv = fetchAny(a.data, a.type, k);
              r.lower        = vmin(r.lower, tv);
              r.upper        = vmax(r.upper, tv);
            }
            continue;
          }
          const float v = fetchAny(a.data, a.type, denseOffset(g.layout, px, py, pz));
          r.lower            = vmin(r.lower, v);
          r.upper            = vmax(r.upper, v);
        }
    return r;
  }
  uint64_t voxel = 0;
  const Range1f *er = nullptr;
  uint32_t eo       = 0;
  if (inDom) {
    SparseSource<float> src(g, a, attr);
    src.maxDepth = 3;
    voxel        = src.traverse(ox & ~7u, oy & ~7u, oz & ~7u, er, eo);
  }
  if ((voxel & 3u) != 3u) {
    if (!inDom) return {a.background, a.background};
    return {0.f, 0.f};
  }
  if (g.temporalFormat) {
    // temporal leaves: every time sample of the window's voxels (of the one
    // voxel of a tile)
    const uint64_t li   = voxel >> 6;
    const uint32_t tf   = (uint32_t)(voxel >> 4) & 3u;
    const uint64_t base = g.leafDataOffset[li];
    auto extend = [&](uint64_t vi) {
      uint64_t b = base + vi, e = b + 1;
      if (tf == 1u) {
        b = base + vi * g.leafSteps[li];
        e = b + g.leafSteps[li];
      } else if (tf == 2u) {
        const uint64_t *ix = (const uint64_t *)g.tuIndices + g.leafIdxOffset[li];
        b                  = base + ix[vi];
        e                  = base + ix[vi + 1];
      }
      for (uint64_t k = b; k < e; ++k) {
        const float v = fetchAny(a.data, a.type, k);
        r.lower       = vmin(r.lower, v);
        r.upper       = vmax(r.upper, v);
      }
    };
    if (((voxel >> 2) & 3u) != 1u) {
      extend(0);
      return r;
    }
    for (uint32_t x = rx[0]; x < rx[1]; ++x)
      for (uint32_t y = ry[0]; y < ry[1]; ++y)
        for (uint32_t z = rz[0]; z < rz[1]; ++z) extend((x << 6) | (y << 3) | z);
    return r;
  }
  const uint64_t leaf = g.leafSlot ? g.leafSlot[voxel >> 6] : (voxel >> 6);
  if (((voxel >> 2) & 3u) == 0u) {
    const float v = fetchAny(a.tile, a.type, leaf);
    return {v, v};
  }
  for (uint32_t x = rx[0]; x < rx[1]; ++x)
    for (uint32_t y = ry[0]; y < ry[1]; ++y)
      for (uint32_t z = rz[0]; z < rz[1]; ++z) {
        const float v = fetchAny(a.data, a.type, leaf * 512u + ((x << 6) | (y << 3) | z));
        r.lower       = vmin(r.lower, v);
        r.upper       = vmax(r.upper, v);
      }
  return r;
}

// 3x3x3 leaf neighbourhood, windows [7,8), [0,8), [0,2) per axis; NaN ranges
// (e.g. an undefined background) are skipped.
__global__ __launch_bounds__(kRangeBlock) void kLeafRanges(GridView g, AttrView a, uint32_t attr,
                                                            uint64_t numLeaves, const uint32_t *__restrict__ offs,
                                                            Range1f *__restrict__ out)
{
  const uint64_t i = (uint64_t)blockIdx.x * kRangeBlock + threadIdx.x;
  if (i >= numLeaves) return;
  const uint32_t cx = offs[3 * i], cy = offs[3 * i + 1], cz = offs[3 * i + 2];
  const uint32_t win[3][2] = {{7, 8}, {0, 8}, {0, 2}};
  Range1f range = {vinf(), -vinf()};
  for (int ix = 0; ix < 3; ++ix)
    for (int iy = 0; iy < 3; ++iy)
      for (int iz = 0; iz < 3; ++iz) {
        const uint32_t ox = cx + (uint32_t)((ix - 1) * 8), oy = cy + (uint32_t)((iy - 1) * 8),
                       oz = cz + (uint32_t)((iz - 1) * 8);
        const Range1f nr = nodeRange(g, a, attr, ox, oy, oz, win[ix], win[iy], win[iz]);
        if (!isnan(nr.lower)) range.lower = vmin(range.lower, nr.lower);
        if (!isnan(nr.upper)) range.upper = vmax(range.upper, nr.upper);
      }
  out[i] = range;
}

hipError_t launchLeafRanges(const GridView &g, uint32_t attr, uint64_t numLeaves, const uint32_t *leafOffsets,
                            Range1f *out, hipStream_t stream)
{
  if (numLeaves == 0) return hipSuccess;
  AttrView a;
  hipError_t e = hipMemcpyAsync(&a, g.attrs + attr, sizeof(a), hipMemcpyDeviceToHost, stream);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return e;
  const unsigned blocks = (unsigned)((numLeaves + kRangeBlock - 1) / kRangeBlock);
  hipLaunchKernelGGL(kLeafRanges, dim3(blocks), dim3(kRangeBlock), 0, stream, g, a, attr, numLeaves, leafOffsets,
                     out);
  return hipGetLastError();
}

// ---- single-ray steps ----------------------------------------------------------

// The iterator kernels read their context through a pointer to device
// memory: a context passed by value is copied to scratch memory as soon as
// any field is indexed at run time.
__global__ __launch_bounds__(1) void kIntervalInit(const IteratorContextView *__restrict__ cp, IntervalState *s, const float *ray)
{
  const IteratorContextView &c = *cp;
  intervalInit(c, *s, F3{ray[0], ray[1], ray[2]}, F3{ray[3], ray[4], ray[5]}, Range1f{ray[6], ray[7]});
}

__global__ __launch_bounds__(1) void kIntervalNext(const IteratorContextView *__restrict__ cp, IntervalState *s, IntervalOut *out,
                              int *result)
{
  const IteratorContextView &c = *cp;
  IntervalState st = *s;
  IntervalOut o;
  const bool r = intervalNext(c, st, o);
  *s           = st;
  if (r) *out = o;
  *result = r ? 1 : 0;
}

__global__ __launch_bounds__(1) void kHitInit(const IteratorContextView *__restrict__ cp, HitState *s, const float *ray)
{
  const IteratorContextView &c = *cp;
  hitInit(c, *s, F3{ray[0], ray[1], ray[2]}, F3{ray[3], ray[4], ray[5]}, Range1f{ray[6], ray[7]}, ray[8]);
}


hipError_t launchIntervalInit(const IteratorContextView *c, IntervalState *state, const float *ray, hipStream_t st)
{
  hipLaunchKernelGGL(kIntervalInit, dim3(1), dim3(1), 0, st, c, state, ray);
  return hipGetLastError();
}
hipError_t launchIntervalNext(const IteratorContextView *c, IntervalState *state, IntervalOut *out, int *result,
                              hipStream_t st)
{
  hipLaunchKernelGGL(kIntervalNext, dim3(1), dim3(1), 0, st, c, state, out, result);
  return hipGetLastError();
}
hipError_t launchHitInit(const IteratorContextView *c, HitState *state, const float *ray, hipStream_t st)
{
  hipLaunchKernelGGL(kHitInit, dim3(1), dim3(1), 0, st, c, state, ray);
  return hipGetLastError();
}


hipError_t launchHitNext(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                         HitState *state, HitOut *out, int *result, hipStream_t st)
{
  const GridView &g = hostView.sampler.grid;
  if (!g.dense && !g.temporalFormat && g.leafAccess)
    return launchHitNextObserved(hostView, c, a, state, out, result, st);
  if (g.temporalFormat || !g.dense) return launchHitNextOther(hostView, c, a, state, out, result, st);
  if (a.type == kTypeHalf || a.type == kTypeFloat || a.type == kTypeDouble)
    return launchHitNextDenseFloat(hostView, c, a, state, out, result, st);
  return dispatchSource<kDenseIntKinds>(hostView.sampler.grid, a, hostView.sampler.filter,
                                        HitNextLaunch{c, a, state, out, result, st});
}

// ---- batches ---------------------------------------------------------------------


// Interval batches keep one ray per lane: a persistent schedule (as the hit
// batches use) measured 12 % slower here, where ray lengths vary less than
// the cost of the schedule's per-step bookkeeping. The kernel is specialised
// on the context's iterator depth, so the DDA keeps only the levels it walks.
// The iteration is bound by dependent-load latency, so occupancy pays: at
// depth 2 (82 VGPRs, 5 waves per SIMD) asking for 6 waves costs a 12-byte
// spill and measured 11 % faster on the 4096^3 vdb; at depth 3 (100 VGPRs,
// 4 waves) 5 waves cost a 20-byte spill and measured 6 % faster.
template <int kDepth, int kQuery>
__global__ __launch_bounds__(kIterBlock) __attribute__((amdgpu_waves_per_eu(kDepth <= 2 ? 6 : 5)))
void kIntervalBatch(const IteratorContextView *__restrict__ cp, uint32_t numRays,
                                                              const float *__restrict__ org,
                                                              const float *__restrict__ dir,
                                                              const float *__restrict__ tRanges,
                                                              IntervalOut *__restrict__ out, uint32_t maxOut,
                                                              uint32_t *__restrict__ count)
{
  const uint32_t r = blockIdx.x * kIterBlock + threadIdx.x;
  if (r >= numRays) return;
  const IteratorContextView &c = *cp;
  IntervalState st;
  intervalInit<kDepth>(c, st, F3{org[3 * r], org[3 * r + 1], org[3 * r + 2]},
                       F3{dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]}, rayRange(tRanges, r));
  uint32_t n = 0;
  IntervalOut o;
  IntervalOut *dst = out + (uint64_t)r * maxOut;
  while (n < maxOut && intervalNext<false, kDepth, kQuery>(c, st, o)) dst[n++] = o;
  count[r] = n;
}

__global__ void kSetWork(uint32_t *work, uint32_t value) { *work = value; }

hipError_t launchSetWork(uint32_t *work, uint32_t value, hipStream_t st)
{
  hipLaunchKernelGGL(kSetWork, dim3(1), dim3(1), 0, st, work, value);
  return hipGetLastError();
}

uint32_t batchLaneCount(uint32_t numRays, int cus) { return batchLanes(numRays, cus); }

namespace {

template <int kQuery>
void intervalBatchDepth(uint32_t depth, dim3 grid, dim3 block, hipStream_t st, const IteratorContextView *c,
                        uint32_t numRays, const float *org, const float *dir, const float *tRanges, IntervalOut *out,
                        uint32_t maxOut, uint32_t *count)
{
  switch (depth) {
  case 0:
    hipLaunchKernelGGL((kIntervalBatch<0, kQuery>), grid, block, 0, st, c, numRays, org, dir, tRanges, out, maxOut,
                       count);
    break;
  case 1:
    hipLaunchKernelGGL((kIntervalBatch<1, kQuery>), grid, block, 0, st, c, numRays, org, dir, tRanges, out, maxOut,
                       count);
    break;
  case 2:
    hipLaunchKernelGGL((kIntervalBatch<2, kQuery>), grid, block, 0, st, c, numRays, org, dir, tRanges, out, maxOut,
                       count);
    break;
  default:
    hipLaunchKernelGGL((kIntervalBatch<3, kQuery>), grid, block, 0, st, c, numRays, org, dir, tRanges, out, maxOut,
                       count);
    break;
  }
}

}  // namespace

hipError_t launchIntervalBatch(uint32_t depth, bool denseQuery, const IteratorContextView *c, uint32_t numRays,
                               const float *org, const float *dir, const float *tRanges, IntervalOut *out,
                               uint32_t maxOut, uint32_t *count, hipStream_t st)
{
  if (numRays == 0) return hipSuccess;
  const dim3 grid((numRays + kIterBlock - 1) / kIterBlock), block(kIterBlock);
  if (denseQuery)
    intervalBatchDepth<1>(depth, grid, block, st, c, numRays, org, dir, tRanges, out, maxOut, count);
  else
    intervalBatchDepth<0>(depth, grid, block, st, c, numRays, org, dir, tRanges, out, maxOut, count);
  return hipGetLastError();
}

hipError_t launchHitBatch(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                          uint32_t numRays, const float *org, const float *dir, const float *tRanges,
                          const float *times, HitOut *out, uint32_t maxOut, uint32_t *count, uint32_t *work, uint32_t lanes, hipStream_t st)
{
  if (numRays == 0) return hipSuccess;
  const GridView &g = hostView.sampler.grid;
  if (!g.dense && !g.temporalFormat && g.leafAccess)
    return launchHitBatchObserved(hostView, c, a, numRays, org, dir, tRanges, times, out, maxOut, count, work, lanes, st);
  if (g.temporalFormat || !g.dense)
    return launchHitBatchOther(hostView, c, a, numRays, org, dir, tRanges, times, out, maxOut, count, work, lanes, st);
  if (a.type == kTypeHalf || a.type == kTypeFloat || a.type == kTypeDouble)
    return launchHitBatchDenseFloat(hostView, c, a, numRays, org, dir, tRanges, times, out, maxOut, count, work, lanes, st);
  return dispatchSource<kDenseIntKinds>(hostView.sampler.grid, a, hostView.sampler.filter,
                                        HitBatchLaunch{c, a, numRays, org, dir, tRanges, times, out, maxOut, count, work, lanes,
                                                       st});
}

}  // namespace vklhip
