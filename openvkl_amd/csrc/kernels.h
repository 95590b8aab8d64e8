// Host-side launchers of the HIP kernels (implemented in *.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "vkl_views.h"

namespace vklhip {

// vklComputeSampleN: out[i] = sample(xyz[i]) of attribute `attr`, whose view
// (a host copy of s.grid.attrs[attr]) selects the kernel specialisation.
// times (device, n floats) or null (time 0) for temporal volumes.
hipError_t launchSample(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                        float *out, hipStream_t stream, const float *times = nullptr);
// vklComputeGradientN: out[3i..3i+2] = object-space gradient.
hipError_t launchGradient(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                          float *out, hipStream_t stream, const float *times = nullptr);
// out[i * outStride] = sample(xyz[i]) of attribute `attr` (vklComputeSampleMN
// issues one launch per requested attribute with outStride = M).
hipError_t launchSampleStrided(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                               float *out, uint32_t outStride, hipStream_t stream, const float *times = nullptr);

// Fixed-step ray marching fused with sampling: out[r] = sum over s < steps of
// sample(org[r] + t_s dir[r]), t_s = t0 + (s + 0.5) (t1 - t0) / steps with
// [t0, t1] = tRange[r] (2 floats per ray), summed in step order.
hipError_t launchRayMarch(const SamplerView &s, const AttrView &a, uint32_t attr, uint32_t numRays, const float *org,
                          const float *dir, const float *tRange, uint32_t steps, float *out, hipStream_t stream);

// Woodcock (delta) tracking fused with sampling (kernels_woodcock.hip). All
// pointers are device memory; org / dir 3 floats, tRange 2 floats and rng 2
// uint32 (TEA state, updated) per ray; tf holds tfSize (r, g, b, opacity)
// entries over the value range [tfLo, tfHi]; count may be null.
struct WoodcockArgs
{
  uint32_t numRays, maxCollisions, tfSize;
  float sigmaMax, time, shutter, tfLo, tfHi;
  const float *org, *dir, *tRange, *tf;
  uint32_t *rng;
  float *eventT, *samples;
  uint32_t *count;
};
hipError_t launchWoodcock(const SamplerView &s, const AttrView &a, uint32_t attr, const WoodcockArgs &w,
                          hipStream_t stream);

// Interval-driven ray marching with transfer-function compositing
// (kernels_rmi_*.hip). org / dir 3 floats and tRanges 2 floats (or null:
// [0, inf)) per ray; tf holds tfSize (r, g, b, opacity) entries over
// [tfLo, tfHi]; rgba 4 floats per ray; count (may be null) = samples taken;
// at most maxSamples samples per ray. Device pointers.
struct RayMarchIntervalsArgs
{
  uint32_t numRays, tfSize, maxSamples;
  float time, samplingRate, tfLo, tfHi;
  const float *org, *dir, *tRanges, *tf;
  float *rgba;
  uint32_t *count;
};
hipError_t launchRayMarchIntervals(const IteratorContextView &hostView, const IteratorContextView *c,
                                   const AttrView &a, const RayMarchIntervalsArgs &m, hipStream_t stream);
hipError_t launchRayMarchIntervalsSparse(const IteratorContextView &hostView, const IteratorContextView *c,
                                         const AttrView &a, const RayMarchIntervalsArgs &m, hipStream_t stream);

// The hit-iterator renderer (kernels_hitrender_*.hip): per ray the hits of the
// context's iso values, shaded (gradient normal, ambient + two lights with
// shadow hit queries, transfer-function albedo) and composited with surface
// opacity 0.6 until 0.99; at most maxHits hits. Layout as RayMarchIntervalsArgs.
struct HitRenderArgs
{
  uint32_t numRays, tfSize, maxHits;
  float time, tfLo, tfHi;
  const float *org, *dir, *tRanges, *tf;
  float *rgba;
  uint32_t *count;
};
hipError_t launchHitRender(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                           const HitRenderArgs &m, hipStream_t stream);
hipError_t launchHitRenderSparse(const IteratorContextView &hostView, const IteratorContextView *c,
                                 const AttrView &a, const HitRenderArgs &m, hipStream_t stream);
hipError_t launchHitRenderTemporal(const IteratorContextView &hostView, const IteratorContextView *c,
                                   const AttrView &a, const HitRenderArgs &m, hipStream_t stream);
hipError_t launchHitRenderObserved(const IteratorContextView &hostView, const IteratorContextView *c,
                                   const AttrView &a, const HitRenderArgs &m, hipStream_t stream);

// Copy an x-fastest dims[0] x dims[1] x dims[2] array of elemSize-byte
// elements into brick layout L (total = padded element count).
hipError_t launchSwizzle(const void *src, void *dst, uint32_t elemSize, const uint32_t dims[3],
                         const DenseLayout &L, uint64_t total, hipStream_t stream);

// Corner-packed copy (8 elements per cell, see DenseLayout) of an x-fastest
// dims[0] x dims[1] x dims[2] array; dst holds 8 * cells elements.
hipError_t launchPackCorners(const void *src, void *dst, uint32_t elemSize, const uint32_t dims[3],
                             hipStream_t stream);

// Cubic records (16 elements per voxel, see AttrView::cubic; dst holds
// 16 * voxels elements) or, with stencils, whole stencils (64 elements per
// cell; 64 * voxels) of an x-fastest dims[0] x dims[1] x dims[2] array.
hipError_t launchPackCubicRecords(const void *src, void *dst, uint32_t elemSize, const uint32_t dims[3],
                                  bool stencils, hipStream_t stream);

// Tricubic samples (out[i * stride]) or gradients (out[3 i .. 3 i + 2]) of n
// points of a structuredRegular volume whose attribute has cubic records.
hipError_t launchCubicRecords(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                              float *out, uint32_t stride, bool gradient, hipStream_t stream);

// vdb commit: the leaf index grid over g.leafDims (filled with kLeafEmpty,
// then the slot of each dense leaf -- origins absolute, 3 int32 per leaf,
// packed-slot order -- then kLeafTraverse over every cell a tile node of
// input level 1, 2 or 3 covers: tileOrigins[level - 1], numTiles[level - 1]);
// corner-packed volumes: the cell records (8 floats per cell, 512 cells per
// dense leaf) of one attribute.
hipError_t launchBuildLeafIndex(const GridView &g, const int32_t *origins, uint64_t numLeaves,
                                const int32_t *const tileOrigins[3], const uint64_t numTiles[3], uint32_t *index,
                                hipStream_t stream);
hipError_t launchPackLeafCells(const GridView &g, const AttrView &a, uint32_t attr, const int32_t *origins,
                               uint64_t numLeaves, float *cells, hipStream_t stream);

// Per-leaf value ranges over the [-1, +10) voxel window (majorant build).
// leafOffsets: 3 u32 root-relative leaf origins per leaf.
hipError_t launchLeafRanges(const GridView &g, uint32_t attr, uint64_t numLeaves, const uint32_t *leafOffsets,
                            Range1f *out, hipStream_t stream);

// Single-ray iterator steps for the scalar API (state in device memory).
// ray: origin (3), direction (3), tRange (2), time (1) in device memory.
// Interval steps and batches take the context in device memory
// (IteratorContext::deviceView).
hipError_t launchIntervalInit(const IteratorContextView *c, IntervalState *state, const float *ray, hipStream_t stream);
hipError_t launchIntervalNext(const IteratorContextView *c, IntervalState *state, IntervalOut *out, int *result,
                              hipStream_t stream);
// Hit steps take the context in device memory (IteratorContext::deviceView);
// hostView (the same context on the host) and the attribute's view select
// the kernel specialisation.
hipError_t launchHitInit(const IteratorContextView *c, HitState *state, const float *ray, hipStream_t stream);
hipError_t launchHitNext(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                         HitState *state, HitOut *out, int *result, hipStream_t stream);

// Batched iteration: one ray per thread; ray r writes up to maxOut results at
// out[r*maxOut], count[r] = number written. org / dir: 3 floats per ray;
// tRanges: 2 floats per ray (null = [0, inf)); times may be null.
// denseQuery: the context's queries use the dense node numbering
// (denseQueryOf); the kernel is specialised on it.
hipError_t launchIntervalBatch(uint32_t depth, bool denseQuery, const IteratorContextView *c, uint32_t numRays,
                               const float *org, const float *dir, const float *tRanges, IntervalOut *out,
                               uint32_t maxOut, uint32_t *count, hipStream_t stream);
// Whether iterator queries on g use the dense node numbering (a
// structuredRegular volume within one level-1 node, VdbQueryVoxelDense.ih:48-57).
inline bool denseQueryOf(const GridView &g)
{
  return g.dense != 0 && g.activeSize[0] <= kRes[1] && g.activeSize[1] <= kRes[1] && g.activeSize[2] <= kRes[1];
}
// Hit batches run a persistent schedule: `lanes` (batchLaneCount) lanes take
// rays from the work counter `work`, which launchSetWork must set to `lanes`
// on the same stream before each batch launch.
uint32_t batchLaneCount(uint32_t numRays, int cus);
hipError_t launchSetWork(uint32_t *work, uint32_t value, hipStream_t stream);
hipError_t launchHitBatch(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                          uint32_t numRays, const float *org, const float *dir, const float *tRanges,
                          const float *times, HitOut *out, uint32_t maxOut, uint32_t *count, uint32_t *work, uint32_t lanes, hipStream_t stream);
// The same for structuredRegular volumes of half / float / double voxels
// (kernels_hit_dense.hip), for temporal structuredRegular and vdb volumes
// (kernels_hit_temporal.hip), for vdb volumes (kernels_hit_other.hip) and for
// observed vdb samplers (kernels_hit_observed.hip); launchHitNext /
// launchHitBatch forward to them.
hipError_t launchHitNextObserved(const IteratorContextView &hostView, const IteratorContextView *c,
                                 const AttrView &a, HitState *state, HitOut *out, int *result, hipStream_t stream);
hipError_t launchHitBatchObserved(const IteratorContextView &hostView, const IteratorContextView *c,
                                  const AttrView &a, uint32_t numRays, const float *org, const float *dir,
                                  const float *tRanges, const float *times, HitOut *out, uint32_t maxOut,
                                  uint32_t *count, uint32_t *work, uint32_t lanes, hipStream_t stream);
hipError_t launchHitNextDenseFloat(const IteratorContextView &hostView, const IteratorContextView *c,
                                   const AttrView &a, HitState *state, HitOut *out, int *result, hipStream_t stream);
hipError_t launchHitBatchDenseFloat(const IteratorContextView &hostView, const IteratorContextView *c,
                                    const AttrView &a, uint32_t numRays, const float *org, const float *dir,
                                    const float *tRanges, const float *times, HitOut *out, uint32_t maxOut,
                                    uint32_t *count, uint32_t *work, uint32_t lanes, hipStream_t stream);
hipError_t launchHitNextTemporal(const IteratorContextView &hostView, const IteratorContextView *c,
                                 const AttrView &a, HitState *state, HitOut *out, int *result, hipStream_t stream);
hipError_t launchHitBatchTemporal(const IteratorContextView &hostView, const IteratorContextView *c,
                                  const AttrView &a, uint32_t numRays, const float *org, const float *dir,
                                  const float *tRanges, const float *times, HitOut *out, uint32_t maxOut,
                                  uint32_t *count, uint32_t *work, uint32_t lanes, hipStream_t stream);
hipError_t launchHitNextOther(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                              HitState *state, HitOut *out, int *result, hipStream_t stream);
hipError_t launchHitBatchOther(const IteratorContextView &hostView, const IteratorContextView *c, const AttrView &a,
                               uint32_t numRays, const float *org, const float *dir, const float *tRanges,
                               const float *times, HitOut *out, uint32_t maxOut, uint32_t *count, uint32_t *work, uint32_t lanes, hipStream_t stream);

}  // namespace vklhip
