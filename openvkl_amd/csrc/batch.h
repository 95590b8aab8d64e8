// Batched multi-volume sampling: descriptor table entries and launch shape.
#pragma once

#include <hip/hip_runtime_api.h>

#include "vkl_views.h"

namespace vklhip {

enum : int32_t { kBatchGeneric = 0, kBatchDenseLinearF32 = 1, kBatchDenseLinearF16 = 2 };

// One (sampler, points, outputs) batch; lives in device memory.
struct BatchItem
{
  SamplerView sampler;
  AttrView attr;
  const float *xyz;  // count * 3 floats (device)
  float *out;        // count floats (device)
  uint64_t begin;    // offset of this batch in the flattened point space
  uint64_t count;
  uint32_t attrIndex;
  int32_t fastPath;
};

// Launch shape chosen by the autotuner.
struct BatchLaunch
{
  uint32_t blocks;
  uint32_t threads;
  uint32_t tile;  // points per workgroup per iteration
};

hipError_t launchSampleBatch(const BatchItem *items, const uint64_t *begins, uint32_t numItems, uint64_t total,
                             const BatchLaunch &cfg, hipStream_t stream);

}  // namespace vklhip
