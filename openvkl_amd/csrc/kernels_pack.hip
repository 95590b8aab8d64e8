// Commit-time builders of the corner-packed vdb representation: the leaf
// index grid (dense-leaf slot per 8^3 leaf cell of the domain) and the cell
// records (the 8 values the exact sampler reads at each cell's corners).

#include "kernels.h"
#include "vkl_device.hh"

namespace vklhip {

constexpr int kPackBlock = 256;

__global__ __launch_bounds__(kPackBlock) void kFillU32(uint32_t *__restrict__ p, uint64_t n, uint32_t v)
{
  const uint64_t i = (uint64_t)blockIdx.x * kPackBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// Leaf origins are absolute index coordinates, multiples of 8 from the root.
__global__ __launch_bounds__(kPackBlock) void kSetLeafIndex(GridView g, const int32_t *__restrict__ origins,
                                                             uint64_t numLeaves, uint32_t *__restrict__ index)
{
  const uint64_t l = (uint64_t)blockIdx.x * kPackBlock + threadIdx.x;
  if (l >= numLeaves) return;
  const uint32_t x = (uint32_t)(origins[3 * l] - g.rootOrigin[0]) >> 3;
  const uint32_t y = (uint32_t)(origins[3 * l + 1] - g.rootOrigin[1]) >> 3;
  const uint32_t z = (uint32_t)(origins[3 * l + 2] - g.rootOrigin[2]) >> 3;
  if (x < g.leafDims[0] && y < g.leafDims[1] && z < g.leafDims[2])
    index[((uint64_t)z * g.leafDims[1] + y) * g.leafDims[0] + x] = (uint32_t)l;
}

// One work-item per (leaf, cell): record[k] = at(corner k), exactly what the
// traversal path returns for that corner.
template <typename T>
__global__ __launch_bounds__(kPackBlock) void kPackLeafCells(GridView g, AttrView a, uint32_t attr,
                                                              const int32_t *__restrict__ origins,
                                                              uint64_t numLeaves, float *__restrict__ cells)
{
  const uint64_t t = (uint64_t)blockIdx.x * kPackBlock + threadIdx.x;
  if (t >= numLeaves * 512u) return;
  const uint64_t l    = t >> 9;
  const uint32_t cell = (uint32_t)(t & 511u);
  const int32_t x = origins[3 * l] + (int32_t)(cell >> 6), y = origins[3 * l + 1] + (int32_t)((cell >> 3) & 7u),
                z = origins[3 * l + 2] + (int32_t)(cell & 7u);
  AttrView base = a;
  base.cells    = nullptr;  // evaluate through the node table
  GridView gt   = g;
  gt.leafIndex  = nullptr;  // by traversal, not through the flattened index
  const SparseSource<T> src(gt, base, attr);
  CellRecord<float> r;
#pragma unroll
  for (int k = 0; k < 8; k++) r.v[k] = src.at(x + ((k >> 2) & 1), y + ((k >> 1) & 1), z + (k & 1));
  reinterpret_cast<CellRecord<float> *>(cells)[t] = r;
}

// Tile nodes of one level, each covering cells^3 leaf cells from its origin:
// those cells are marked kLeafTraverse (the node table decides there, also
// where a dense leaf overlaps the tile).
__global__ __launch_bounds__(kPackBlock) void kMarkTiles(GridView g, const int32_t *__restrict__ origins,
                                                          uint64_t numTiles, uint32_t cells, uint32_t *__restrict__ index)
{
  const uint64_t t     = (uint64_t)blockIdx.x * kPackBlock + threadIdx.x;
  const uint64_t per   = (uint64_t)cells * cells * cells;
  if (t >= numTiles * per) return;
  const uint64_t k = t / per, c = t % per;
  const uint32_t x = ((uint32_t)(origins[3 * k] - g.rootOrigin[0]) >> 3) + (uint32_t)(c % cells);
  const uint32_t y = ((uint32_t)(origins[3 * k + 1] - g.rootOrigin[1]) >> 3) + (uint32_t)((c / cells) % cells);
  const uint32_t z = ((uint32_t)(origins[3 * k + 2] - g.rootOrigin[2]) >> 3) + (uint32_t)(c / cells / cells);
  if (x < g.leafDims[0] && y < g.leafDims[1] && z < g.leafDims[2])
    index[((uint64_t)z * g.leafDims[1] + y) * g.leafDims[0] + x] = kLeafTraverse;
}

hipError_t launchBuildLeafIndex(const GridView &g, const int32_t *origins, uint64_t numLeaves,
                                const int32_t *const tileOrigins[3], const uint64_t numTiles[3], uint32_t *index,
                                hipStream_t stream)
{
  const uint64_t n = (uint64_t)g.leafDims[0] * g.leafDims[1] * g.leafDims[2];
  if (n) hipLaunchKernelGGL(kFillU32, dim3((unsigned)((n + kPackBlock - 1) / kPackBlock)), dim3(kPackBlock), 0, stream,
                            index, n, kLeafEmpty);
  if (numLeaves)
    hipLaunchKernelGGL(kSetLeafIndex, dim3((unsigned)((numLeaves + kPackBlock - 1) / kPackBlock)), dim3(kPackBlock), 0,
                       stream, g, origins, numLeaves, index);
  // Tiles of input levels 1, 2, 3 span kRes[level] voxels: 512, 16, 1 leaf
  // cells per axis.
  for (int l = 0; l < 3; l++) {
    if (!numTiles[l]) continue;
    const uint32_t cells = kRes[l + 1] / 8 ? kRes[l + 1] / 8 : 1;
    const uint64_t total = numTiles[l] * cells * cells * cells;
    hipLaunchKernelGGL(kMarkTiles, dim3((unsigned)((total + kPackBlock - 1) / kPackBlock)), dim3(kPackBlock), 0, stream,
                       g, tileOrigins[l], numTiles[l], cells, index);
  }
  return hipGetLastError();
}

hipError_t launchPackLeafCells(const GridView &g, const AttrView &a, uint32_t attr, const int32_t *origins,
                               uint64_t numLeaves, float *cells, hipStream_t stream)
{
  const uint64_t n = numLeaves * 512u;
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)((n + kPackBlock - 1) / kPackBlock)), block(kPackBlock);
  if (a.type == kTypeHalf)
    hipLaunchKernelGGL(kPackLeafCells<__half>, grid, block, 0, stream, g, a, attr, origins, numLeaves, cells);
  else
    hipLaunchKernelGGL(kPackLeafCells<float>, grid, block, 0, stream, g, a, attr, origins, numLeaves, cells);
  return hipGetLastError();
}

}  // namespace vklhip
