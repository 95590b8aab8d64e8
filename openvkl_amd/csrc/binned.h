// Spatially binned stream sampling of large dense volumes: bin geometry,
// scratch layout and launcher (kernels_binned.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "vkl_views.h"

namespace vklhip {

// Operation of a binned launch (the filter's stencil sets the brick halo).
enum BinOp : int {
  kBinSampleNearest  = 0,
  kBinSampleLinear   = 1,
  kBinSampleCubic    = 2,
  kBinGradientLinear = 3,
  kBinGradientCubic  = 4,
};

// Bins are brick regions of 2^shift voxels per axis (count = bins per axis).
// A bin's LDS brick adds haloLo voxels below and haloHi above the region on
// every axis: the filter's taps for any point whose floor lies in the region.
struct BinGeometry
{
  uint32_t shift[3];
  uint32_t count[3];
  uint32_t numBins;
  uint32_t haloLo, haloHi;
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

// Largest brick of at most maxBrickBytes (halo of `op` included) whose bin
// count over `dims` is at most maxBins; false if there is none.
bool chooseBinGeometry(const uint32_t dims[3], uint32_t elemSize, BinOp op, size_t maxBrickBytes, uint32_t maxBins,
                       BinGeometry &out);
size_t binnedBrickBytes(const BinGeometry &b, uint32_t elemSize);

// Dense volumes. Samples: out[i * outStride] = sample(xyz[i]); gradients:
// out[3i .. 3i+2] (outStride ignored). Bit-equal to launchSample /
// launchGradient. The scratch must hold b.numBins * 8 * capacity records.
hipError_t launchBinned(const SamplerView &s, const AttrView &a, const BinGeometry &b, BinOp op, uint64_t n,
                        const float *xyz, float *out, uint32_t outStride, const BinScratch &scratch,
                        hipStream_t stream);

}  // namespace vklhip
