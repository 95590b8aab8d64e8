// Spatially binned stream sampling of large dense volumes: bin geometry,
// scratch layout and launcher (kernels_binned.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "vkl_views.h"

namespace vklhip {

// Bins are brick regions of 2^shift voxels per axis; count = bins per axis.
struct BinGeometry
{
  uint32_t shift[3];
  uint32_t count[3];
  uint32_t numBins;
};

// Device scratch of one binned call. records: numBins x 8 segments of
// `capacity` 16-byte records {p.x, p.y, p.z, point index}.
struct BinScratch
{
  float4 *records;
  uint32_t *segCount;       // numBins * 8
  uint32_t *overflow;       // point indices that did not fit their segment (n)
  uint32_t *overflowCount;  // 1
  uint32_t capacity;
};

// Largest brick of at most maxBrickBytes (one-voxel halo included) whose bin
// count over `dims` is at most maxBins; false if there is none.
bool chooseBinGeometry(const uint32_t dims[3], uint32_t elemSize, size_t maxBrickBytes, uint32_t maxBins,
                       BinGeometry &out);
size_t binnedBrickBytes(const BinGeometry &b, uint32_t elemSize);

// Linear filter, dense volumes: out[i * outStride] = sample(xyz[i]), bit-equal
// to launchSample. The scratch must hold b.numBins * 8 * capacity records.
hipError_t launchSampleBinned(const SamplerView &s, const AttrView &a, const BinGeometry &b, uint64_t n,
                              const float *xyz, float *out, uint32_t outStride, const BinScratch &scratch,
                              hipStream_t stream);

}  // namespace vklhip
