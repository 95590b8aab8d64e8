// Device-resident views of committed objects. Plain structs shared by the
// C++ object layer (which fills them at commit) and the HIP kernels (which
// take them by value as kernel arguments).
#pragma once

#include <stdint.h>

namespace vklhip {

// VDB topology 6/5/4/3 (openvkl/CMakeLists.txt:37-49).
constexpr int kNumLevels = 4;
constexpr uint32_t kLogRes[5]      = {6, 5, 4, 3, 0};
constexpr uint32_t kTotalLogRes[5] = {18, 12, 7, 3, 0};
constexpr uint32_t kRes[5]         = {262144, 4096, 128, 8, 1};
constexpr uint64_t kNumVoxels[4]   = {262144, 32768, 4096, 512};

// VKLDataType codes used for voxel data.
enum : uint32_t {
  kTypeUchar  = 2500,
  kTypeShort  = 3000,
  kTypeUshort = 3500,
  kTypeHalf   = 5800,
  kTypeFloat  = 6000,
  kTypeDouble = 7000,
};

enum : int { kFilterNearest = 0, kFilterLinear = 100, kFilterCubic = 200 };

// Leaf index grid entries that are not dense-leaf slots.
constexpr uint32_t kLeafEmpty    = 0xffffffffu;
constexpr uint32_t kLeafTraverse = 0xfffffffeu;

struct Range1f
{
  float lower, upper;
};

// One attribute of a volume.
struct AttrView
{
  const void *data;   // dense: x-fastest voxels; vdb: packed dense leaves (512 each)
  const void *tile;   // vdb: packed tile values (one per tile leaf), may be null
  uint32_t type;      // VKLDataType of the elements
  float background;   // value outside the domain
  // vdb with "cornerPacked": per dense leaf (packed-slot order) 512 cell
  // records of 8 floats, the values the sampler reads at the cell's 8
  // trilinear corners (record order 4 dx + 2 dy + dz); null otherwise.
  const float *cells;
  // structuredRegular with "cubicRecords": per voxel (x, y, z), x fastest, a
  // record of 16 elements e[4 j + k] = v(x, y - 1 + j, z - 1 + k), y and z
  // clamped into the domain. The 4x4x4 tricubic stencil of cell (x, y, z) is
  // the 4 consecutive records of x - 1 .. x + 2 (64 elements in a row).
  // With "cubicStencils" (cubicStencils = 1) instead: per cell (x, y, z), x
  // fastest, its whole stencil e[16 i + 4 j + k] = v(x - 1 + i, y - 1 + j,
  // z - 1 + k), coordinates clamped into the domain. Null otherwise.
  const void *cubic;
  uint32_t cubicStencils;
};

// One inner level of the node table (levels 0..2).
struct LevelView
{
  const uint64_t *voxels;       // numNodes * res^3 voxel codes
  const Range1f *valueRange;    // numNodes * res^3 * numAttributes
  uint64_t numNodes;
};

// Dense voxel addressing: element offset of root-relative voxel (x, y, z) is
// X(x) + Y(y) + Z(z) with
//   X(x) = ((x >> sx) << S) | (x & mx)
//   Y(y) = (y >> sy) * strideY + ((y & my) << sx)
//   Z(z) = (z >> sz) * strideZ + ((z & mz) << (sx + sy))
// for bricks of 2^sx x 2^sy x 2^sz voxels (S = sx + sy + sz, bricks x-fastest).
// sx = sy = sz = 0 is the reference's x-fastest array (strideY = nx,
// strideZ = nx * ny).
//
// Corner-packed layout (cornerPacked = 1, with sx = sy = sz = 0): cell
// (x, y, z) holds its 8 trilinear corners v(x + dx, y + dy, z + dz) as one
// record of 8 elements in the order 4 dx + 2 dy + dz (corners past the last
// voxel repeat it); X(x) = 8 x, strideY = 8 nx, strideZ = 8 nx ny. Voxel
// (x, y, z) is element 0 of its cell, so every voxel-addressed read works
// unchanged, and a trilinear stencil is one aligned 8-element load.
struct DenseLayout
{
  uint32_t sx, sy, sz;
  uint32_t cornerPacked;
  uint64_t strideY, strideZ;
};

struct GridView
{
  float o2i[12];          // object -> index, row-major 3x4
  int32_t rootOrigin[3];  // index space
  uint32_t activeSize[3]; // domain size in voxels
  float domainBox[6];     // domain-space (root-relative) bounding box
  int32_t dense;
  int32_t constantCellData;
  uint32_t numAttributes;
  uint32_t maxSamplingDepth;
  DenseLayout layout;         // dense volumes: voxel addressing
  const AttrView *attrs;      // device array [numAttributes]
  const uint64_t *leafSlot;   // vdb with per-leaf node.data: leaf index -> packed slot
  LevelView levels[3];
  // vdb leaf index grid: per 8^3 leaf cell of the domain (x fastest,
  // leafDims = ceil(activeSize / 8)) the dense-leaf slot, kLeafEmpty where
  // no node covers the cell, kLeafTraverse where a tile node does (the node
  // table decides there). With "cornerPacked" the cell records hold values
  // for maxSamplingDepth == packedDepth. Null when not built.
  const uint32_t *leafIndex;
  uint32_t leafDims[3];
  uint32_t packedDepth;
  // vdb samplers with a LeafNodeAccess observer: traversals ending on a leaf
  // pointer store accessStamp at leafAccess[leaf index] (null: unobserved).
  uint32_t *leafAccess;
  uint32_t accessStamp;
  // structuredRegular temporal data (VKLTemporalFormat): 0 constant; 1
  // structured, numTimesteps values per voxel (time fastest); 2 unstructured,
  // voxel v owns positions [tuIndices[v], tuIndices[v + 1]) of the data and of
  // tuTimes (indices uint32, or uint64 when tuIndex64).
  uint32_t temporalFormat;
  uint32_t numTimesteps;
  const void *tuIndices;
  uint32_t tuIndex64;
  const float *tuTimes;
  // vdb with temporal leaves (temporalFormat != 0 on a sparse grid): every
  // leaf's elements in one array per attribute, leaf i at leafDataOffset[i];
  // the leaf's temporal format is in its voxel code; structured leaves have
  // leafSteps[i] samples per voxel; unstructured leaves index tuIndices
  // (uint64) from leafIdxOffset[i] and tuTimes from leafTimesOffset[i].
  const uint64_t *leafDataOffset;
  const uint32_t *leafSteps;
  const uint64_t *leafIdxOffset;
  const uint64_t *leafTimesOffset;
};

// Sampler = grid + filters.
struct SamplerView
{
  GridView grid;
  int32_t filter;
  int32_t gradientFilter;
};

// ---- iterator state (lives in caller-owned buffers; also used in batches) ----

// Dda.ih:17-36 (4 levels x 3 axes, index axis * 4 + level). The reference's
// per-level tDelta = res * |1 / dir| and idxDelta = sign(dir) * res are not
// stored: the steps recompute them from absRcp (= |1 / dir| as ddaInit
// computes it) and the sign of rayDir, which gives the same values, and the
// state stays small enough to live in registers.
struct DdaState
{
  uint32_t level;
  float t;
  float rayOrg[3];
  float rayDir[3];
  float absRcp[3];
  int32_t idx[12];
  float tNext[12];
};

struct IntervalState
{
  float tMax;
  float nominalDeltaT;
  DdaState dda;
  // vdb node-table cache: the level-2 node that the last query reached for
  // the 128^3 block ck (ck[0] == INT32_MIN: empty). A query in the same block
  // reaches the same node through the same level-0 / level-1 voxels.
  int32_t ck[3];
  uint32_t cnode;
};

struct HitState
{
  IntervalState iv;
  float origin[3];
  float direction[3];
  float time;
  float curLower, curUpper;  // current interval t range
  float curNominalDeltaT;
  float lastHitOffsetT;
  uint32_t pad;
};

// Interval / hit iterator contexts.
struct IteratorContextView
{
  SamplerView sampler;
  uint32_t attributeIndex;
  uint32_t maxIteratorDepth;
  int32_t numRanges;       // value-range selector (0 = everything)
  Range1f rangesMinMax;
  const Range1f *ranges;   // device array [numRanges]
  int32_t numValues;       // hit iterator iso values
  const float *values;     // device array [numValues]
};

struct IntervalOut
{
  float tLower, tUpper, vLower, vUpper, nominalDeltaT;
};

struct HitOut
{
  float t, sample, epsilon;
};

}  // namespace vklhip
