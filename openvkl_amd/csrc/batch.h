// Batched multi-volume sampling: descriptor table entries, specialisation
// groups and launch shape.
#pragma once

#include <hip/hip_runtime_api.h>

#include "vkl_views.h"

namespace vklhip {

// One (sampler, points, outputs) batch; lives in device memory.
struct BatchItem
{
  SamplerView sampler;
  AttrView attr;
  const float *xyz;  // count * 3 floats (device)
  float *out;        // count floats (device)
  uint64_t begin;    // offset of this batch in its group's flattened point space
  uint64_t count;
  uint32_t attrIndex;
  uint32_t pad;
};

// Batches sharing one kernel specialisation. grid / attr / filter (host
// copies from the group's first batch) select the instantiation; items and
// begins are device arrays.
struct BatchGroup
{
  GridView grid;
  AttrView attr;
  int32_t filter;
  uint32_t mergedDense;  // structuredRegular, 32-bit offsets, element type per batch
  uint32_t numItems;
  uint64_t total;
  const BatchItem *items;
  const uint64_t *begins;
};

// Launch shape and grouping chosen by the autotuner (the same shape for
// every group).
struct BatchLaunch
{
  uint32_t blocks;
  uint32_t threads;
  uint32_t tile;    // points per workgroup per iteration
  uint32_t merged;  // 1: structuredRegular batches of a filter share a launch
};

hipError_t launchSampleBatch(const BatchGroup *groups, uint32_t numGroups, const BatchLaunch &cfg,
                             hipStream_t stream);

// Whether a batch joins the filter's merged structuredRegular group (32-bit
// voxel offsets, not temporal; any element type).
inline bool mergesDense(const GridView &g, const AttrView &a)
{
  if (!g.dense || g.temporalFormat) return false;
  if (a.type != kTypeUchar && a.type != kTypeShort && a.type != kTypeUshort && a.type != kTypeHalf &&
      a.type != kTypeFloat && a.type != kTypeDouble)
    return false;
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  return nbz * g.layout.strideZ < (1ull << 32);
}


// Specialisation key of (grid, attribute, filter): batches with equal keys
// run through the same kernel instantiation (dense offsets 32 / 64 bit or
// temporal, sparse observed or not).
inline uint32_t sourceKey(const GridView &g, const AttrView &a, int filter)
{
  uint32_t variant = 0;
  if (g.dense && g.temporalFormat) {
    variant = 1u << 11;
  } else if (g.dense) {
    const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
    variant            = nbz * g.layout.strideZ >= (1ull << 32) ? 1u : 0u;
  } else {
    variant = g.temporalFormat ? (1u << 11) : (g.leafAccess ? 1u : 0u);
  }
  return (a.type << 12) | ((uint32_t)filter << 2) | (g.dense ? 2u : 0u) | variant;
}

// Group key of (grid, attribute, filter): batches with equal keys run in one
// launch -- the merged structuredRegular group of a filter, or one kernel
// instantiation (dense offsets 32 / 64 bit or temporal, sparse observed or
// not).
inline uint32_t groupKey(const GridView &g, const AttrView &a, int filter)
{
  if (mergesDense(g, a)) return (1u << 31) | (uint32_t)filter;
  return sourceKey(g, a, filter);
}

}  // namespace vklhip
