// CPU parity oracle: C++ restatement of the Open VKL 2.0.2 CPU-device hot path
// (VDB / structuredRegular sampling, majorant build, DDA interval iteration,
// bisection hit iteration).
//
// TEST INFRASTRUCTURE ONLY -- never linked into the shipped library.
// Built with -ffp-contract=off: binary32 arithmetic, no FMA contraction, in
// the operation order of the cited reference lines. Citations are to files
// under openvkl/devices/cpu/ unless stated otherwise.

#include "vkl_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

// --- constants ---------------------------------------------------------------

constexpr uint32_t kUchar = 2500, kShort = 3000, kUshort = 3500, kHalf = 5800,
                   kFloat = 6000, kDouble = 7000;
constexpr int kFilterNearest = 0, kFilterLinear = 100, kFilterCubic = 200;
constexpr uint32_t kFormatTile = 0, kFormatDense = 1;

// Topology 6/5/4/3 (openvkl/CMakeLists.txt:37-49, vdb_topology.h.in:23-27).
constexpr uint32_t kLogRes[5]      = {6, 5, 4, 3, 0};
constexpr uint32_t kTotalLogRes[5] = {18, 12, 7, 3, 0};
constexpr uint32_t kRes[5]         = {262144, 4096, 128, 8, 1};
constexpr uint64_t kNumVoxels[4]   = {262144, 32768, 4096, 512};
constexpr int kNumLevels           = 4;

const float kInf    = std::numeric_limits<float>::infinity();
const float kFltMin = std::numeric_limits<float>::min();

std::string g_lastError;

// VKL_BACKGROUND_UNDEFINED: quiet NaN 0xFFC068B5 (VKLBackgroundUndefined.h:30).
float vkl_bg()
{
  const uint32_t bits = 0xFFC068B5u;
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

struct V3f { float x, y, z; };
struct V3i { int32_t x, y, z; };
struct V3u { uint32_t x, y, z; };
struct R1f { float lower, upper; };

inline R1f emptyRange() { return {kInf, -kInf}; }

// ISPC / rkcommon min and max: (a < b) ? a : b.
inline float fmin_(float a, float b) { return a < b ? a : b; }
inline float fmax_(float a, float b) { return a > b ? a : b; }

// rkcommon lerp: (1 - t) * a + t * b.
inline float lerp(float t, float a, float b) { return (1.f - t) * a + t * b; }

// float -> int32 with the x86 cvttss2si convention (NaN / out of range ->
// INT32_MIN), applied after floor.
inline int32_t toInt(float f)
{
  if (!(f >= -2147483648.f && f < 2147483648.f)) return INT32_MIN;
  return (int32_t)f;
}

// rkcommon divide_safe: 1 / x with |x| < FLT_MIN clamped to +-FLT_MIN.
inline float divide_safe(float x)
{
  const float d = (std::fabs(x) < kFltMin) ? (x >= 0.f ? kFltMin : -kFltMin) : x;
  return 1.f / d;
}
// rkcommon rcp_safe (IEEE reciprocal; ISPC's rcp() may be approximate,
// covered by the 1e-6 tolerance on nominalDeltaT).
inline float rcp_safe(float x) { return divide_safe(x); }

float halfToFloat(uint16_t h)
{
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t exp  = (h >> 10) & 0x1fu;
  uint32_t mant       = h & 0x3ffu;
  uint32_t bits;
  if (exp == 0) {
    if (mant == 0) {
      bits = sign;
    } else {  // subnormal half -> normal float
      int e = -1;
      do {
        e++;
        mant <<= 1;
      } while ((mant & 0x400u) == 0);
      mant &= 0x3ffu;
      bits = sign | ((uint32_t)(127 - 15 - e) << 23) | (mant << 13);
    }
  } else if (exp == 31) {
    bits = sign | 0x7f800000u | (mant << 13);
  } else {
    bits = sign | ((exp - 15 + 127) << 23) | (mant << 13);
  }
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

// Typed element fetch, Data.ih:66-226 (plain conversion, no normalisation).
inline float fetch(const void *base, uint32_t type, uint64_t i)
{
  switch (type) {
  case kUchar: return (float)((const uint8_t *)base)[i];
  case kShort: return (float)((const int16_t *)base)[i];
  case kUshort: return (float)((const uint16_t *)base)[i];
  case kHalf: return halfToFloat(((const uint16_t *)base)[i]);
  case kFloat: return ((const float *)base)[i];
  case kDouble: return (float)((const double *)base)[i];
  }
  return 0.f;
}

size_t typeSize(uint32_t type)
{
  switch (type) {
  case kUchar: return 1;
  case kShort: case kUshort: case kHalf: return 2;
  case kFloat: return 4;
  case kDouble: return 8;
  }
  throw std::runtime_error("unsupported data type");
}

// Voxel codes, VdbGrid.h:91-203.
inline bool voxelIsEmpty(uint64_t v) { return (v & 3u) == 0u; }
inline bool voxelIsError(uint64_t v) { return (v & 3u) == 1u; }
inline bool voxelIsChild(uint64_t v) { return (v & 3u) == 2u; }
inline bool voxelIsLeaf(uint64_t v) { return (v & 3u) == 3u; }
inline uint64_t makeChild(uint64_t idx) { return (idx << 2) | 2u; }
inline uint64_t makeLeaf(uint64_t idx, uint32_t fmt, uint32_t tfmt)
{
  return (idx << 6) + (((uint64_t)tfmt & 3u) << 4) + (((uint64_t)fmt & 3u) << 2) + 3u;
}
inline uint64_t makeError(uint32_t level, uint32_t voxelOffset)
{
  return ((uint64_t)voxelOffset << 32) + ((uint64_t)level << 8) + 1u;
}
inline uint64_t leafIndexOf(uint64_t v) { return v >> 6; }
inline uint32_t leafFormatOf(uint64_t v) { return (uint32_t)((v >> 2) & 3u); }

// vdb_topology.h.in:96-170: linear voxel index of a root-relative offset
// inside a level-l node, z fastest.
inline uint64_t offsetToLinear(uint32_t l, uint32_t x, uint32_t y, uint32_t z)
{
  const uint32_t m = kRes[l] - 1;
  const uint64_t vx = (uint64_t)(x & m) >> kTotalLogRes[l + 1];
  const uint64_t vy = (uint64_t)(y & m) >> kTotalLogRes[l + 1];
  const uint64_t vz = (uint64_t)(z & m) >> kTotalLogRes[l + 1];
  return (vx << (2 * kLogRes[l])) + (vy << kLogRes[l]) + vz;
}

// --- tricubic matrix ----------------------------------------------------------

// The Lekien-Marsden matrix maps the 64 constraints (f, fx, fy, fz, fxy,
// fxz, fyz, fxyz at the 8 cell corners, corner index 4x+2y+z) to the 64
// coefficients a_ijk (index 16i+4j+k) of sum a_ijk x^i y^j z^k. It is the
// inverse of the constraint matrix, which has small integer entries; the
// inverse is integral too. Derived here instead of transcribed.
struct Tricubic
{
  int32_t A[64][64];
  Tricubic()
  {
    static const int dtype[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                    {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    auto dmono = [](int n, int d, int c) -> double {
      if (d == 0) return (n == 0) ? 1.0 : (c ? 1.0 : 0.0);
      if (n == 0) return 0.0;
      return n * ((n - 1 == 0) ? 1.0 : (c ? 1.0 : 0.0));
    };
    std::vector<std::vector<double>> M(64, std::vector<double>(128, 0.0));
    for (int d = 0; d < 8; d++)
      for (int q = 0; q < 8; q++) {
        const int row = 8 * d + q;
        const int cx = (q >> 2) & 1, cy = (q >> 1) & 1, cz = q & 1;
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++)
              M[row][16 * i + 4 * j + k] = dmono(i, dtype[d][0], cx) *
                                           dmono(j, dtype[d][1], cy) *
                                           dmono(k, dtype[d][2], cz);
        M[row][64 + row] = 1.0;
      }
    // Gauss-Jordan on [B | I] -> [I | B^-1].
    for (int c = 0; c < 64; c++) {
      int p = c;
      for (int r = c + 1; r < 64; r++)
        if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
      std::swap(M[c], M[p]);
      const double piv = M[c][c];
      for (int k = 0; k < 128; k++) M[c][k] /= piv;
      for (int r = 0; r < 64; r++) {
        if (r == c || M[r][c] == 0.0) continue;
        const double f = M[r][c];
        for (int k = 0; k < 128; k++) M[r][k] -= f * M[c][k];
      }
    }
    // B a = CTR  =>  a = B^-1 CTR; rows of B are constraints, columns are
    // coefficients, so B^-1 has rows = coefficients, columns = constraints.
    for (int r = 0; r < 64; r++)
      for (int c = 0; c < 64; c++) A[r][c] = (int32_t)std::lround(M[r][64 + c]);
  }
};
const Tricubic &tricubic()
{
  static Tricubic t;
  return t;
}

}  // namespace

// --- grid ------------------------------------------------------------------

struct OracleLevel
{
  uint64_t numNodes = 0;
  std::vector<V3u> origin;
  std::vector<uint64_t> voxels;
  std::vector<R1f> valueRange;  // numNodes * numVoxels * numAttributes
};

struct OracleGrid
{
  float o2i[12], i2o[12];
  V3i rootOrigin;
  V3u activeSize;
  float domainBox[6];
  bool constantCellData = false;
  bool dense            = false;

  uint32_t numAttributes = 1;
  std::vector<uint32_t> attrTypes;
  std::vector<float> background;

  // dense
  V3i dims{0, 0, 0};
  std::vector<std::vector<uint8_t>> denseData;

  // sparse
  uint64_t numLeaves = 0;
  bool packed        = false;
  std::vector<std::vector<uint8_t>> packedDense, packedTile;
  std::vector<std::vector<uint8_t>> leafData;  // [leaf * numAttr + a]

  OracleLevel levels[3];
  std::vector<R1f> volumeRange;
  float bbox[6];

  int filter = kFilterLinear, gradientFilter = kFilterLinear;
  uint32_t maxSamplingDepth = 3;
  bool cubicDouble = false;

  // LeafNodeAccess observation (VdbLeafAccessObserver.ih): when set, every
  // sampling traversal that ends on a leaf pointer sets entry [leaf index].
  uint32_t *leafAccess = nullptr;

  // structuredRegular temporal data: format 0 constant, 1 structured
  // (numTimesteps values per voxel), 2 unstructured (tuIndices / tuTimes);
  // curTime is the time of the sample being evaluated.
  uint32_t temporalFormat = 0;
  uint32_t numTimesteps   = 0;
  std::vector<uint64_t> tuIndices;
  std::vector<float> tuTimes;
  mutable float curTime = 0.f;

  // vdb temporal leaves (node.data only): per input node its temporal
  // format, structured time steps and unstructured index / time arrays.
  std::vector<uint32_t> leafTF, leafSteps;
  std::vector<std::vector<uint64_t>> leafIdx;
  std::vector<std::vector<float>> leafTimes;
};

namespace {

// VdbGrid.h:56-89.
inline V3f xfmVector(const float *M, V3f p)
{
  return {M[0] * p.x + M[1] * p.y + M[2] * p.z,
          M[3] * p.x + M[4] * p.y + M[5] * p.z,
          M[6] * p.x + M[7] * p.y + M[8] * p.z};
}
inline V3f xfmPoint(const float *M, V3f p)
{
  V3f r = xfmVector(M, p);
  return {r.x + M[9], r.y + M[10], r.z + M[11]};
}
inline V3f xfmNormal(const float *MInv, V3f p)
{
  return {MInv[0] * p.x + MInv[3] * p.y + MInv[6] * p.z,
          MInv[1] * p.x + MInv[4] * p.y + MInv[7] * p.z,
          MInv[2] * p.x + MInv[5] * p.y + MInv[8] * p.z};
}

// rkcommon AffineSpace3f inverse as used in DenseVdbVolume.cpp:41-64 /
// VdbVolume.cpp:400-404: o2i.l = i2o.l^-1, o2i.p = -(o2i.l * i2o.p).
// Matrix stored row-major: M[0..8] rows, M[9..11] translation. The linear
// inverse is the adjoint divided by the determinant (rkcommon
// LinearSpace3::inverse: adjoint / det, columns vx, vy, vz).
void invertAffine(const float *i2o, float *o2i)
{
  // Columns of the linear part (rkcommon stores vx, vy, vz as columns).
  const V3f vx{i2o[0], i2o[3], i2o[6]};
  const V3f vy{i2o[1], i2o[4], i2o[7]};
  const V3f vz{i2o[2], i2o[5], i2o[8]};
  auto cross = [](V3f a, V3f b) {
    return V3f{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  };
  auto dot = [](V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
  // adjoint rows = cross products of columns; det = dot(vx, cross(vy, vz)).
  const V3f r0 = cross(vy, vz), r1 = cross(vz, vx), r2 = cross(vx, vy);
  const float det = dot(vx, r0);
  // inverse = adjoint * rcp(det) (rkcommon LinearSpace3: adjoint() / det(),
  // operator/ multiplies by rcp = 1/x); row i of the inverse is r_i * (1/det).
  const float rdet = 1.f / det;
  const V3f q0{r0.x * rdet, r0.y * rdet, r0.z * rdet};
  const V3f q1{r1.x * rdet, r1.y * rdet, r1.z * rdet};
  const V3f q2{r2.x * rdet, r2.y * rdet, r2.z * rdet};
  o2i[0] = q0.x; o2i[1] = q0.y; o2i[2] = q0.z;
  o2i[3] = q1.x; o2i[4] = q1.y; o2i[5] = q1.z;
  o2i[6] = q2.x; o2i[7] = q2.y; o2i[8] = q2.z;
  const V3f p{i2o[9], i2o[10], i2o[11]};
  const V3f lp = xfmVector(o2i, p);
  o2i[9]  = -lp.x;
  o2i[10] = -lp.y;
  o2i[11] = -lp.z;
}

inline V3u toDomainOffset(V3i ic, V3i root)
{
  return {(uint32_t)(ic.x - root.x), (uint32_t)(ic.y - root.y), (uint32_t)(ic.z - root.z)};
}
inline bool isInDomain(V3u size, V3u o)
{
  return o.x < size.x && o.y < size.y && o.z < size.z;
}

// --- dense sampling (VdbSampler_traverseAndSample.ih:314-349,
//     VdbSampler_dense.ih:290-330) ---------------------------------------------

// interpolateTemporallyStructured, temporal_data_interpolation.ih:13-45:
// the voxel's samples start at baseIdx of data (element type ty).
float interpStructured(const uint8_t *d, uint32_t ty, uint32_t numTimesteps, uint64_t baseIdx, float time)
{
  const float timestepFloat = time * ((float)(numTimesteps - 1));
  const uint32_t timestep   = (uint32_t)timestepFloat;
  const float w             = timestepFloat - ((float)timestep);
  const uint32_t next       = (w == 0.f) ? timestep : timestep + 1;
  const float v0 = fetch(d, ty, baseIdx + timestep);
  const float v1 = fetch(d, ty, baseIdx + next);
  return (1.f - w) * v0 + w * v1;
}

// interpolateTemporallyUnstructured, temporal_data_interpolation.ih:82-152:
// voxel v owns positions [indices[v], indices[v + 1]) of data and times.
// rcp is taken as the correctly rounded 1 / x.
float interpUnstructured(const uint8_t *d, uint32_t ty, const uint64_t *indices, const float *times, uint64_t voxel,
                         float time)
{
  const uint64_t dataIdx = indices[voxel], dataNextIdx = indices[voxel + 1];
  const float t0 = times[dataIdx];
  if (time <= t0) return fetch(d, ty, dataIdx);
  const float t1 = times[dataNextIdx - 1];
  if (time >= t1) return fetch(d, ty, dataNextIdx - 1);
  const uint64_t numTimeSamples = dataNextIdx - dataIdx;
  if (numTimeSamples == 2) {
    const float w = (time - t0) * (1.f / (t1 - t0));
    return (1.f - w) * fetch(d, ty, dataIdx) + w * fetch(d, ty, dataNextIdx - 1);
  }
  uint64_t lower = 0, upper = numTimeSamples;
  while (lower <= upper) {
    const uint64_t mid   = (lower + upper) / 2;
    const float timesMid = times[dataIdx + mid];
    if (time >= timesMid && time <= times[dataIdx + mid + 1]) {
      lower = dataIdx + mid;
      upper = lower + 1;
      break;
    } else if (time < timesMid) {
      upper = mid;
    } else {
      lower = mid;
    }
  }
  const float timesLow = times[lower], timesHigh = times[upper];
  const float w        = (time - timesLow) * (1.f / (timesHigh - timesLow));
  return (1.f - w) * fetch(d, ty, lower) + w * fetch(d, ty, upper);
}

// VdbSampler_dense.ih:360-446: structured base index numTimesteps * voxel.
float temporallyStructured(const OracleGrid &g, uint32_t attr, uint64_t voxel, float time)
{
  return interpStructured(g.denseData[attr].data(), g.attrTypes[attr], g.numTimesteps,
                          (uint64_t)g.numTimesteps * voxel, time);
}

float temporallyUnstructured(const OracleGrid &g, uint32_t attr, uint64_t voxel, float time)
{
  return interpUnstructured(g.denseData[attr].data(), g.attrTypes[attr], g.tuIndices.data(), g.tuTimes.data(),
                            voxel, time);
}

inline float denseCorner(const OracleGrid &g, uint32_t attr, V3i ic)
{
  const V3u o = toDomainOffset(ic, g.rootOrigin);
  if (!isInDomain(g.activeSize, o)) return g.background[attr];
  const uint64_t idx = (uint64_t)o.z * g.activeSize.y * (uint64_t)g.activeSize.x +
                       (uint64_t)o.y * g.activeSize.x + o.x;
  if (g.temporalFormat == 1) return temporallyStructured(g, attr, idx, g.curTime);
  if (g.temporalFormat == 2) return temporallyUnstructured(g, attr, idx, g.curTime);
  return fetch(g.denseData[attr].data(), g.attrTypes[attr], idx);
}

// --- sparse traversal (VdbSampler_traverseAndSample.ih:85-202,
//     VdbSampleInner.ih.in:33-77, VdbSamplerDispatchInner.ih.in:34-113) -------

// Returns the voxel code found for a domain offset already known to be in
// the domain; error codes when a child / leaf lies below maxSamplingDepth.
uint64_t traverse(const OracleGrid &g, V3u o, uint32_t maxDepth)
{
  uint64_t nodeIndex = 0;
  for (uint32_t l = 0; l < 3; ++l) {
    const uint64_t vo = nodeIndex * kNumVoxels[l] + offsetToLinear(l, o.x, o.y, o.z);
    const uint32_t vo32 = (uint32_t)vo;
    const uint64_t v    = g.levels[l].voxels[vo32];
    if ((voxelIsLeaf(v) || voxelIsChild(v)) && (l + 1 > maxDepth))
      return makeError(l, vo32);
    if (voxelIsChild(v) && l + 1 < 3) {
      nodeIndex = v >> 2;
      continue;
    }
    return v;
  }
  return 0;
}

float leafValue(const OracleGrid &g, uint64_t voxel, V3u o, uint32_t attr)
{
  const uint64_t leaf = leafIndexOf(voxel);
  const uint32_t fmt  = leafFormatOf(voxel);
  const uint32_t t    = g.attrTypes[attr];
  const uint64_t lin  = ((uint64_t)(o.x & 7u) << 6) + ((o.y & 7u) << 3) + (o.z & 7u);
  if (g.packed) {  // VdbSampler_packed_denseZYX.ih:77-186, _packed_tile.ih:53-110
    if (fmt == kFormatDense) return fetch(g.packedDense[attr].data(), t, leaf * 512 + lin);
    return fetch(g.packedTile[attr].data(), t, leaf);
  }
  // VdbSampler_denseZYX.ih:124-183, VdbSampler_tile.ih:82-120
  const auto &d      = g.leafData[leaf * g.numAttributes + attr];
  const uint64_t vi  = fmt == kFormatDense ? lin : 0;
  const uint32_t tfm = (uint32_t)(voxel >> 4) & 3u;  // leaf temporal format
  if (tfm == 1)  // VdbSampler_denseZYX.ih:191-260: base numTimesteps * voxel
    return interpStructured(d.data(), t, g.leafSteps[leaf], (uint64_t)g.leafSteps[leaf] * vi, g.curTime);
  if (tfm == 2)
    return interpUnstructured(d.data(), t, g.leafIdx[leaf].data(), g.leafTimes[leaf].data(), vi, g.curTime);
  return fetch(d.data(), t, vi);
}

// Value range of one voxel of a temporal leaf: every time sample
// (VdbSampler_denseZYX.ih:45-118).
R1f temporalVoxelRange(const OracleGrid &g, uint64_t voxel, uint64_t vi, uint32_t attr)
{
  const uint64_t leaf = leafIndexOf(voxel);
  const uint32_t tfm  = (uint32_t)(voxel >> 4) & 3u;
  const auto &d       = g.leafData[leaf * g.numAttributes + attr];
  uint64_t b = vi, e = vi + 1;
  if (tfm == 1) {
    b = vi * g.leafSteps[leaf];
    e = b + g.leafSteps[leaf];
  } else if (tfm == 2) {
    b = g.leafIdx[leaf][vi];
    e = g.leafIdx[leaf][vi + 1];
  }
  R1f r = emptyRange();
  for (uint64_t k = b; k < e; ++k) {
    const float v = fetch(d.data(), g.attrTypes[attr], k);
    r.lower = fmin_(r.lower, v);
    r.upper = fmax_(r.upper, v);
  }
  return r;
}

// VdbSampler_sample_inner, VdbSampler_traverseAndSample.ih:144-197.
float sampleVoxel(const OracleGrid &g, uint64_t voxel, V3u o, uint32_t attr)
{
  if (voxelIsLeaf(voxel)) return leafValue(g, voxel, o, attr);
  if (voxelIsError(voxel)) {
    const uint32_t level = (uint8_t)(voxel >> 8);
    const uint32_t vo    = (uint32_t)(voxel >> 32);
    const R1f r = g.levels[level].valueRange[(uint64_t)vo * g.numAttributes + attr];
    return 0.5f * (r.lower + r.upper);
  }
  if (!isInDomain(g.activeSize, o)) return g.background[attr];
  return 0.f;
}

// VdbSampler_traverseAndSample (varying), :277-309.
float sparseCorner(const OracleGrid &g, uint32_t attr, V3i ic)
{
  const V3u o = toDomainOffset(ic, g.rootOrigin);
  if (!isInDomain(g.activeSize, o)) return g.background[attr];
  const uint64_t v = traverse(g, o, g.maxSamplingDepth);
  // VdbSampleInner.ih.in:69-74: a non-empty voxel that ends the traversal
  // (error voxels return earlier, :48-54) marks its leaf index.
  if (g.leafAccess && voxelIsLeaf(v)) __atomic_store_n(&g.leafAccess[leafIndexOf(v)], 1u, __ATOMIC_RELAXED);
  return sampleVoxel(g, v, o, attr);
}

inline float corner(const OracleGrid &g, uint32_t attr, V3i ic)
{
  return g.dense ? denseCorner(g, attr, ic) : sparseCorner(g, attr, ic);
}

inline V3i floorToInt(V3f p)
{
  return {toInt(std::floor(p.x)), toInt(std::floor(p.y)), toInt(std::floor(p.z))};
}

// VdbSampler_nearest.ih:25-57.
float sampleNearest(const OracleGrid &g, V3f p, uint32_t attr)
{
  const float off = g.constantCellData ? 0.f : 0.5f;
  const V3i ic = {toInt(std::floor(p.x + off)), toInt(std::floor(p.y + off)),
                  toInt(std::floor(p.z + off))};
  return corner(g, attr, ic);
}

// 8 corner values, s[4x + 2y + z] (VdbSampler_trilinear.ih:98-107).
inline void trilinearCorners(const OracleGrid &g, V3i ic, uint32_t attr, float *s)
{
  for (int i = 0; i < 8; i++)
    s[i] = corner(g, attr, {ic.x + ((i >> 2) & 1), ic.y + ((i >> 1) & 1), ic.z + (i & 1)});
}

// VdbSampler_trilinear.ih:355-399.
float sampleLinear(const OracleGrid &g, V3f p, uint32_t attr)
{
  const V3i ic = floorToInt(p);
  const V3f d  = {p.x - (float)ic.x, p.y - (float)ic.y, p.z - (float)ic.z};
  float s[8];
  trilinearCorners(g, ic, attr, s);
  return lerp(d.x, lerp(d.y, lerp(d.z, s[0], s[1]), lerp(d.z, s[2], s[3])),
              lerp(d.y, lerp(d.z, s[4], s[5]), lerp(d.z, s[6], s[7])));
}

// VdbSampler_trilinear.ih:720-780 (index space).
V3f gradientLinear(const OracleGrid &g, V3f p, uint32_t attr)
{
  const V3i ic = floorToInt(p);
  const V3f d  = {p.x - (float)ic.x, p.y - (float)ic.y, p.z - (float)ic.z};
  float s[8];
  trilinearCorners(g, ic, attr, s);
  V3f r;
  r.x = lerp(d.y, lerp(d.z, s[4] - s[0], s[5] - s[1]), lerp(d.z, s[6] - s[2], s[7] - s[3]));
  r.y = lerp(d.x, lerp(d.z, s[2] - s[0], s[3] - s[1]), lerp(d.z, s[6] - s[4], s[7] - s[5]));
  r.z = lerp(d.x, lerp(d.y, s[1] - s[0], s[3] - s[2]), lerp(d.y, s[5] - s[4], s[7] - s[6]));
  return r;
}

// Tricubic (VdbSampler_tricubic.ih:391-578): stencil [-1,2]^3, index
// (x+1)*16 + (y+1)*4 + (z+1); constraints by central differences; coefficients
// via the Lekien-Marsden matrix, summed in increasing constraint order.
template <typename R>
void tricubicCoefficients(const OracleGrid &g, V3i ic, uint32_t attr, R *coef)
{
  R S[64];
  for (int i = 0; i < 64; i++)
    S[i] = (R)corner(g, attr, {ic.x + (i >> 4) - 1, ic.y + ((i >> 2) & 3) - 1, ic.z + (i & 3) - 1});
  auto spl = [&](int x, int y, int z) { return S[(x + 1) * 16 + (y + 1) * 4 + (z + 1)]; };
  const R h = (R)0.5, q = (R)0.25, e = (R)0.125;
  R C[64];
  for (int c = 0; c < 8; c++) {
    const int x = (c >> 2) & 1, y = (c >> 1) & 1, z = c & 1;
    C[c]      = spl(x, y, z);
    C[8 + c]  = h * (-spl(x - 1, y, z) + spl(x + 1, y, z));
    C[16 + c] = h * (-spl(x, y - 1, z) + spl(x, y + 1, z));
    C[24 + c] = h * (-spl(x, y, z - 1) + spl(x, y, z + 1));
    C[32 + c] = q * (spl(x - 1, y - 1, z) - spl(x - 1, y + 1, z) - spl(x + 1, y - 1, z) + spl(x + 1, y + 1, z));
    C[40 + c] = q * (spl(x - 1, y, z - 1) - spl(x - 1, y, z + 1) - spl(x + 1, y, z - 1) + spl(x + 1, y, z + 1));
    C[48 + c] = q * (spl(x, y - 1, z - 1) - spl(x, y - 1, z + 1) - spl(x, y + 1, z - 1) + spl(x, y + 1, z + 1));
    C[56 + c] = e * (-spl(x - 1, y - 1, z - 1) + spl(x - 1, y - 1, z + 1) + spl(x - 1, y + 1, z - 1) -
                     spl(x - 1, y + 1, z + 1) + spl(x + 1, y - 1, z - 1) - spl(x + 1, y - 1, z + 1) -
                     spl(x + 1, y + 1, z - 1) + spl(x + 1, y + 1, z + 1));
  }
  const Tricubic &T = tricubic();
  for (int r = 0; r < 64; r++) {
    bool first = true;
    R acc      = 0;
    for (int c = 0; c < 64; c++) {
      const int32_t a = T.A[r][c];
      if (a == 0) continue;
      const R term = (R)a * C[c];
      acc   = first ? term : acc + term;
      first = false;
    }
    coef[r] = acc;
  }
}

template <typename R>
float sampleCubicT(const OracleGrid &g, V3f p, uint32_t attr)
{
  const V3i ic = floorToInt(p);
  R coef[64];
  tricubicCoefficients<R>(g, ic, attr, coef);
  const V3f d = {p.x - (float)ic.x, p.y - (float)ic.y, p.z - (float)ic.z};
  const R dx = d.x, dy = d.y, dz = d.z;
  const R X[4] = {1, dx, dx * dx, dx * dx * dx};
  const R Y[4] = {1, dy, dy * dy, dy * dy * dy};
  const R Z[4] = {1, dz, dz * dz, dz * dz * dz};
  R result = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) result += coef[16 * i + 4 * j + k] * X[i] * Y[j] * Z[k];
  return (float)result;
}

template <typename R>
V3f gradientCubicT(const OracleGrid &g, V3f p, uint32_t attr)
{
  const V3i ic = floorToInt(p);
  R coef[64];
  tricubicCoefficients<R>(g, ic, attr, coef);
  const V3f d = {p.x - (float)ic.x, p.y - (float)ic.y, p.z - (float)ic.z};
  const R dx = d.x, dy = d.y, dz = d.z;
  const R X[4] = {1, dx, dx * dx, dx * dx * dx};
  const R Y[4] = {1, dy, dy * dy, dy * dy * dy};
  const R Z[4] = {1, dz, dz * dz, dz * dz * dz};
  R rx = 0, ry = 0, rz = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        const R c = coef[16 * i + 4 * j + k];
        if (i > 0) rx += c * (R)i * X[i - 1] * Y[j] * Z[k];
        if (j > 0) ry += c * (R)j * X[i] * Y[j - 1] * Z[k];
        if (k > 0) rz += c * (R)k * X[i] * Y[j] * Z[k - 1];
      }
  return V3f{(float)rx, (float)ry, (float)rz};
}

// The literal reference form is binary32; the binary64 form evaluates the
// same algorithm (same matrix, same order) with a 1e-16 rounding unit, for
// bounding the binary32 form's own rounding error.
float sampleCubic(const OracleGrid &g, V3f p, uint32_t attr)
{
  return g.cubicDouble ? sampleCubicT<double>(g, p, attr) : sampleCubicT<float>(g, p, attr);
}

V3f gradientCubic(const OracleGrid &g, V3f p, uint32_t attr)
{
  return g.cubicDouble ? gradientCubicT<double>(g, p, attr) : gradientCubicT<float>(g, p, attr);
}

float sampleIndex(const OracleGrid &g, int filter, V3f p, uint32_t attr)
{
  switch (filter) {
  case kFilterNearest: return sampleNearest(g, p, attr);
  case kFilterCubic: return sampleCubic(g, p, attr);
  default: return sampleLinear(g, p, attr);
  }
}

// VdbSampler.ispc:97-129 (stream sample), :303-335 (stream gradient).
float sampleObject(const OracleGrid &g, V3f oc, uint32_t attr)
{
  return sampleIndex(g, g.filter, xfmPoint(g.o2i, oc), attr);
}

V3f gradientObject(const OracleGrid &g, V3f oc, uint32_t attr)
{
  const V3f p = xfmPoint(g.o2i, oc);
  V3f gi{0.f, 0.f, 0.f};
  if (g.gradientFilter == kFilterLinear)
    gi = gradientLinear(g, p, attr);
  else if (g.gradientFilter == kFilterCubic)
    gi = gradientCubic(g, p, attr);
  else
    return V3f{0.f, 0.f, 0.f};  // VdbSampler_nearest.ih:313-363
  return xfmNormal(g.o2i, gi);
}

// --- grid build (VdbVolume.cpp:71-392, 583-1062; DenseVdbVolume.cpp:20-119) --

inline V3u nodeOrigin(V3u o, uint32_t level)
{
  const uint32_t mask = ~(kRes[level] - 1);
  return {o.x & mask, o.y & mask, o.z & mask};
}

struct LeafInput
{
  std::vector<uint32_t> level, format;
  std::vector<V3i> origin;
};

// insertLeaves (VdbVolume.cpp:255-331): nodes numbered in first-encounter
// order of the leaves, binned by level; leaf codes carry the input index, or
// the packed dense / tile index when packed data is used.
void buildLevels(OracleGrid &g, const LeafInput &in, const std::vector<V3u> &offsets)
{
  std::vector<std::vector<uint64_t>> binned(kNumLevels);
  for (uint64_t i = 0; i < g.numLeaves; ++i) {
    if (in.level[i] == 0) throw std::runtime_error("there must not be any leaf nodes on level 0");
    if (in.level[i] >= (uint32_t)kNumLevels)
      throw std::runtime_error("invalid node level " + std::to_string(in.level[i]) +
                               " for this vdb configuration");
    binned[in.level[i]].push_back(i);
  }
  std::vector<uint64_t> packedIndex(g.numLeaves, 0);
  if (g.packed) {
    uint64_t nd = 0, nt = 0;
    for (uint64_t i = 0; i < g.numLeaves; ++i)
      packedIndex[i] = (in.format[i] == kFormatDense) ? nd++ : nt++;
  }
  const uint32_t nattr = g.numAttributes;
  auto addNode = [&](uint32_t l, V3u org) -> uint64_t {
    OracleLevel &L = g.levels[l];
    const uint64_t idx = L.numNodes++;
    L.origin.push_back(org);
    L.voxels.resize(L.numNodes * kNumVoxels[l], 0);
    L.valueRange.resize(L.numNodes * kNumVoxels[l] * nattr, emptyRange());
    return idx;
  };
  addNode(0, V3u{0, 0, 0});
  for (uint32_t ll = 0; ll < (uint32_t)kNumLevels; ++ll) {
    for (uint64_t idx : binned[ll]) {
      const V3u off = offsets[idx];
      uint64_t nodeIndex = 0;
      for (uint32_t l = 0; l < ll; ++l) {
        OracleLevel &L = g.levels[l];
        const uint64_t v = nodeIndex * kNumVoxels[l] + offsetToLinear(l, off.x, off.y, off.z);
        uint64_t &voxel  = L.voxels[v];
        if (voxelIsLeaf(voxel))
          throw std::runtime_error("Attempted to insert a leaf node into a leaf node");
        if (voxelIsEmpty(voxel)) {
          if (l + 1 < ll) {
            const uint64_t child = addNode(l + 1, nodeOrigin(off, l + 1));
            g.levels[l].voxels[v] = makeChild(child);  // re-index: addNode may reallocate
            nodeIndex = child;
          } else {
            g.levels[l].voxels[v] = makeLeaf(g.packed ? packedIndex[idx] : idx, in.format[idx],
                                             g.leafTF.empty() ? 0 : g.leafTF[idx]);
          }
        } else {
          nodeIndex = voxel >> 2;
        }
      }
    }
  }
}

// Neighbour-node value range, VdbSampler.ispc:383-455.
R1f nodeValueRange(const OracleGrid &g, uint64_t voxel, V3u o, const uint32_t rx[2],
                   const uint32_t ry[2], const uint32_t rz[2], uint32_t attr)
{
  const float bgOut = g.background[attr];
  R1f r = emptyRange();
  if (g.dense) {
    if (!isInDomain(g.activeSize, o)) return {bgOut, bgOut};
    // VdbSampler_dense.ih:35-90
    for (uint32_t x = rx[0]; x < rx[1]; ++x)
      for (uint32_t y = ry[0]; y < ry[1]; ++y)
        for (uint32_t z = rz[0]; z < rz[1]; ++z) {
          const V3u p{o.x + x, o.y + y, o.z + z};
          if (!(p.x < (uint32_t)g.dims.x && p.y < (uint32_t)g.dims.y && p.z < (uint32_t)g.dims.z))
            continue;
          const uint64_t idx = (uint64_t)p.z * g.activeSize.y * (uint64_t)g.activeSize.x +
                               (uint64_t)p.y * g.activeSize.x + p.x;
          // temporal data: every time sample of the voxel
          uint64_t b = idx, e = idx + 1;
          if (g.temporalFormat == 1) {
            b = idx * g.numTimesteps;
            e = b + g.numTimesteps;
          } else if (g.temporalFormat == 2) {
            b = g.tuIndices[idx];
            e = g.tuIndices[idx + 1];
          }
          for (uint64_t k = b; k < e; ++k) {
            const float v = fetch(g.denseData[attr].data(), g.attrTypes[attr], k);
            r.lower = fmin_(r.lower, v);
            r.upper = fmax_(r.upper, v);
          }
        }
    return r;
  }
  if (!voxelIsLeaf(voxel)) {
    if (!isInDomain(g.activeSize, o)) return {bgOut, bgOut};
    return {0.f, 0.f};
  }
  if (!g.packed && ((voxel >> 4) & 3u) != 0) {
    if (leafFormatOf(voxel) == kFormatTile) return temporalVoxelRange(g, voxel, 0, attr);
    for (uint32_t x = rx[0]; x < rx[1]; ++x)
      for (uint32_t y = ry[0]; y < ry[1]; ++y)
        for (uint32_t z = rz[0]; z < rz[1]; ++z) {
          const R1f v = temporalVoxelRange(g, voxel, ((uint64_t)x << 6) + (y << 3) + z, attr);
          r.lower = fmin_(r.lower, v.lower);
          r.upper = fmax_(r.upper, v.upper);
        }
    return r;
  }
  if (leafFormatOf(voxel) == kFormatTile) {
    const float v = leafValue(g, voxel, V3u{0, 0, 0}, attr);
    return {v, v};
  }
  for (uint32_t x = rx[0]; x < rx[1]; ++x)
    for (uint32_t y = ry[0]; y < ry[1]; ++y)
      for (uint32_t z = rz[0]; z < rz[1]; ++z) {
        const float v = leafValue(g, voxel, V3u{x, y, z}, attr);
        r.lower = fmin_(r.lower, v);
        r.upper = fmax_(r.upper, v);
      }
  return r;
}

// VdbSampler_computeValueRange, VdbSampler.ispc:460-557: 3x3x3 leaf-sized
// neighbourhood, voxel ranges [7,8), [0,8), [0,2) per axis; NaN skipped.
R1f leafValueRange(const OracleGrid &g, V3u center, uint32_t attr)
{
  static const uint32_t vr[3][2] = {{7, 8}, {0, 8}, {0, 2}};
  R1f range = emptyRange();
  for (int ix = 0; ix < 3; ++ix)
    for (int iy = 0; iy < 3; ++iy)
      for (int iz = 0; iz < 3; ++iz) {
        const V3i nodeOrg{g.rootOrigin.x + (int32_t)center.x + (ix - 1) * 8,
                          g.rootOrigin.y + (int32_t)center.y + (iy - 1) * 8,
                          g.rootOrigin.z + (int32_t)center.z + (iz - 1) * 8};
        const V3u o = toDomainOffset(nodeOrg, g.rootOrigin);
        uint64_t voxel = 0;
        if (!g.dense && isInDomain(g.activeSize, o))
          voxel = traverse(g, nodeOrigin(o, 3), 3);
        const R1f nr = nodeValueRange(g, voxel, o, vr[ix], vr[iy], vr[iz], attr);
        if (!std::isnan(nr.lower)) range.lower = fmin_(range.lower, nr.lower);
        if (!std::isnan(nr.upper)) range.upper = fmax_(range.upper, nr.upper);
      }
  return range;
}

// computeValueRanges, VdbVolume.cpp:338-392, then the volume range from
// level 0 (:1048-1057).
void buildValueRanges(OracleGrid &g, const LeafInput &in, const std::vector<V3u> &offsets)
{
  const uint32_t nattr = g.numAttributes;
  // The leaf ranges read the finished node table only; they are computed on
  // all host threads, then folded into the levels in leaf order.
  std::vector<R1f> leafRanges(g.numLeaves * nattr);
  {
    const uint64_t nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (uint64_t t = 0; t < nt; ++t)
      pool.emplace_back([&, t] {
        for (uint64_t idx = t; idx < g.numLeaves; idx += nt)
          for (uint32_t a = 0; a < nattr; ++a) leafRanges[idx * nattr + a] = leafValueRange(g, offsets[idx], a);
      });
    for (auto &th : pool) th.join();
  }
  for (uint64_t idx = 0; idx < g.numLeaves; ++idx) {
    const R1f *lr = &leafRanges[idx * nattr];
    uint64_t nodeIndex = 0;
    for (uint32_t l = 0; l < in.level[idx]; ++l) {
      OracleLevel &L = g.levels[l];
      const uint64_t v = nodeIndex * kNumVoxels[l] +
                         offsetToLinear(l, offsets[idx].x, offsets[idx].y, offsets[idx].z);
      for (uint32_t a = 0; a < nattr; ++a) {
        R1f &r  = L.valueRange[v * nattr + a];
        r.lower = fmin_(r.lower, lr[a].lower);
        r.upper = fmax_(r.upper, lr[a].upper);
      }
      const uint64_t voxel = L.voxels[v];
      if (voxelIsLeaf(voxel)) break;
      nodeIndex = voxel >> 2;
    }
  }
  g.volumeRange.assign(nattr, emptyRange());
  for (uint32_t a = 0; a < nattr; ++a)
    for (uint64_t i = 0; i < kNumVoxels[0]; ++i) {
      const R1f r = g.levels[0].valueRange[i * nattr + a];
      g.volumeRange[a].lower = fmin_(g.volumeRange[a].lower, r.lower);
      g.volumeRange[a].upper = fmax_(g.volumeRange[a].upper, r.upper);
    }
}

void finishBounds(OracleGrid &g, const float idxBox[6])
{
  for (int i = 0; i < 3; i++) {
    g.domainBox[i]     = idxBox[i] - (float)(&g.rootOrigin.x)[i];
    g.domainBox[3 + i] = idxBox[3 + i] - (float)(&g.rootOrigin.x)[i];
  }
  // VdbVolume.cpp:796-806: transform the 8 index-box corners.
  g.bbox[0] = g.bbox[1] = g.bbox[2] = kInf;
  g.bbox[3] = g.bbox[4] = g.bbox[5] = -kInf;
  for (int i = 0; i < 8; ++i) {
    const V3f v{(i & 1) ? idxBox[3] : idxBox[0], (i & 2) ? idxBox[4] : idxBox[1],
                (i & 4) ? idxBox[5] : idxBox[2]};
    const V3f w = xfmPoint(g.i2o, v);
    g.bbox[0] = std::min(g.bbox[0], w.x);
    g.bbox[1] = std::min(g.bbox[1], w.y);
    g.bbox[2] = std::min(g.bbox[2], w.z);
    g.bbox[3] = std::max(g.bbox[3], w.x);
    g.bbox[4] = std::max(g.bbox[4], w.y);
    g.bbox[5] = std::max(g.bbox[5], w.z);
  }
}

// --- interval iteration (VdbIterator.ispc, Dda.ih, HDDA.ih.in,
//     VdbQueryVoxel.ih.in, VdbQueryVoxelDense.ih) -------------------------------

struct Dda
{
  uint32_t level;
  float t;
  V3f rayOrg, rayDir;
  int32_t idx[12];
  float tNext[12], tDelta[12];
  int32_t idxDelta[12];
};
inline int OX(uint32_t l) { return (int)l; }
inline int OY(uint32_t l) { return 4 + (int)l; }
inline int OZ(uint32_t l) { return 8 + (int)l; }

inline int dirSign(float v) { return ((int)(0 < v)) - ((int)(v < 0)); }
inline float dirRcp(float v) { return (v == 0) ? kInf : divide_safe(v); }

// Dda.ih:66-84 (intersect_planes).
inline float planeT(float plane, float org, float inv)
{
  const float t = (plane - org) * inv;
  return (plane == org || std::isnan(t)) ? kInf : t;
}

// Dda.ih:127-140 (clampToCell).
inline int32_t clampCell(float f, int32_t res)
{
  const int32_t o = toInt(std::floor(fmax_(f, 0.f)));
  return o & ~(res - 1);
}

// Dda.ih:145-196.
void ddaInit(Dda &d, V3f org, V3f dir, float tMin, uint32_t level)
{
  const V3f oo = {org.x + tMin * dir.x, org.y + tMin * dir.y, org.z + tMin * dir.z};
  const int sx = dirSign(dir.x), sy = dirSign(dir.y), sz = dirSign(dir.z);
  const V3f iDir = {dirRcp(dir.x), dirRcp(dir.y), dirRcp(dir.z)};
  d.level  = std::min<uint32_t>(level, 3);
  d.t      = tMin;
  d.rayOrg = org;
  d.rayDir = dir;
  for (uint32_t l = 0; l <= d.level; ++l) {
    const int32_t res = (int32_t)kRes[l + 1];
    const int32_t ix = clampCell(oo.x, res), iy = clampCell(oo.y, res), iz = clampCell(oo.z, res);
    d.idx[OX(l)] = ix;
    d.idx[OY(l)] = iy;
    d.idx[OZ(l)] = iz;
    d.idxDelta[OX(l)] = sx * res;
    d.idxDelta[OY(l)] = sy * res;
    d.idxDelta[OZ(l)] = sz * res;
    d.tDelta[OX(l)] = (float)res * std::fabs(iDir.x);
    d.tDelta[OY(l)] = (float)res * std::fabs(iDir.y);
    d.tDelta[OZ(l)] = (float)res * std::fabs(iDir.z);
    d.tNext[OX(l)] = planeT((float)(ix + res * (sx >= 0)), org.x, iDir.x);
    d.tNext[OY(l)] = planeT((float)(iy + res * (sy >= 0)), org.y, iDir.y);
    d.tNext[OZ(l)] = planeT((float)(iz + res * (sz >= 0)), org.z, iDir.z);
  }
}

// Dda.ih:204-266.
void ddaReinitBelow(Dda &d, uint32_t level, int32_t dx, int32_t dy, int32_t dz)
{
  const V3f iDir = {dirRcp(d.rayDir.x), dirRcp(d.rayDir.y), dirRcp(d.rayDir.z)};
  const int sx = dirSign(d.rayDir.x), sy = dirSign(d.rayDir.y), sz = dirSign(d.rayDir.z);
  V3f oo = {d.rayOrg.x + d.t * d.rayDir.x, d.rayOrg.y + d.t * d.rayDir.y,
            d.rayOrg.z + d.t * d.rayDir.z};
  if (dx != 0)
    oo.x = oo.x + ((float)sx) * 1e-2f;
  else if (dy != 0)
    oo.y = oo.y + ((float)sy) * 1e-2f;
  else
    oo.z = oo.z + ((float)sz) * 1e-2f;
  for (uint32_t l = level + 1; l <= d.level; ++l) {
    const int32_t res = (int32_t)kRes[l + 1];
    d.idx[OX(l)] = clampCell(oo.x, res);
    d.idx[OY(l)] = clampCell(oo.y, res);
    d.idx[OZ(l)] = clampCell(oo.z, res);
    d.tNext[OX(l)] = planeT((float)(d.idx[OX(l)] + res * (sx >= 0)), d.rayOrg.x, iDir.x);
    d.tNext[OY(l)] = planeT((float)(d.idx[OY(l)] + res * (sy >= 0)), d.rayOrg.y, iDir.y);
    d.tNext[OZ(l)] = planeT((float)(d.idx[OZ(l)] + res * (sz >= 0)), d.rayOrg.z, iDir.z);
  }
}

// Dda.ih:271-309; ties broken x, then y, then z.
void ddaStep(Dda &d, uint32_t l)
{
  const float tx = d.tNext[OX(l)], ty = d.tNext[OY(l)], tz = d.tNext[OZ(l)];
  const float minT = fmin_(fmin_(tx, ty), tz);
  if (minT == tx) {
    d.t = tx;
    d.tNext[OX(l)] = tx + d.tDelta[OX(l)];
    d.idx[OX(l)] += d.idxDelta[OX(l)];
  } else if (minT == ty) {
    d.t = ty;
    d.tNext[OY(l)] = ty + d.tDelta[OY(l)];
    d.idx[OY(l)] += d.idxDelta[OY(l)];
  } else {
    d.t = tz;
    d.tNext[OZ(l)] = tz + d.tDelta[OZ(l)];
    d.idx[OZ(l)] += d.idxDelta[OZ(l)];
  }
}

// HDDA.ih.in:24-101, unrolled over levels 0..3.
bool hddaStepLevel(Dda &d, uint32_t descLevel, uint32_t L)
{
  const bool lowest = (descLevel == L);
  bool stepped      = false;
  if (L + 1 < (uint32_t)kNumLevels) {
    if (lowest) {
      const int32_t ox = d.idx[OX(L)], oy = d.idx[OY(L)], oz = d.idx[OZ(L)];
      ddaStep(d, L);
      stepped = true;
      if (L < d.level)
        ddaReinitBelow(d, L, d.idx[OX(L)] - ox, d.idx[OY(L)] - oy, d.idx[OZ(L)] - oz);
    } else if (hddaStepLevel(d, descLevel, L + 1)) {
      const int32_t res = (int32_t)kRes[L + 1];
      const uint32_t C = L + 1;
      const bool left = d.idx[OX(C)] < d.idx[OX(L)] || d.idx[OY(C)] < d.idx[OY(L)] ||
                        d.idx[OZ(C)] < d.idx[OZ(L)] || d.idx[OX(C)] >= d.idx[OX(L)] + res ||
                        d.idx[OY(C)] >= d.idx[OY(L)] + res || d.idx[OZ(C)] >= d.idx[OZ(L)] + res;
      if (left) {
        ddaStep(d, L);
        stepped = true;
      }
    }
  } else {
    ddaStep(d, L);
    stepped = true;
  }
  return stepped;
}

struct VoxelDesc
{
  uint32_t level;
  bool isEmpty;
  R1f valueRange;
};

struct Ranges
{
  std::vector<R1f> r;
  R1f minmax{-kInf, kInf};
};

// ValueRanges.ih:22-52, box_utility.ih:18-59.
inline bool overlaps(R1f a, R1f b) { return a.upper >= b.lower && a.lower <= b.upper; }
bool rangesOverlap(const Ranges &q, R1f r)
{
  if (q.r.empty()) return true;
  if (!overlaps(q.minmax, r)) return false;
  for (const R1f &x : q.r)
    if (overlaps(x, r)) return true;
  return false;
}

// VdbQueryVoxel.ih.in:42-116 (sparse).
void queryVoxel(const OracleGrid &g, int32_t x, int32_t y, int32_t z, uint32_t depth, uint32_t attr,
                const Ranges &q, VoxelDesc &desc)
{
  uint64_t nodeIndex = 0;
  for (uint32_t L = 0; L < 3; ++L) {
    const uint64_t vo = offsetToLinear(L, (uint32_t)x, (uint32_t)y, (uint32_t)z) + nodeIndex * kNumVoxels[L];
    const uint32_t vo32 = (uint32_t)vo;
    const uint64_t voxel = g.levels[L].voxels[vo32];
    desc.valueRange = g.levels[L].valueRange[(uint64_t)vo32 * g.numAttributes + attr];
    desc.level      = L;
    desc.isEmpty    = voxelIsEmpty(voxel) || !rangesOverlap(q, desc.valueRange);
    if (desc.isEmpty) return;
    const bool atDepth = (L + 1 > depth);
    const bool isLeaf  = voxelIsLeaf(voxel);
    if (isLeaf && !atDepth) desc.level = L + 1;
    if (L + 2 < (uint32_t)kNumLevels && !(atDepth || isLeaf)) {
      nodeIndex = voxel >> 2;
      continue;
    }
    return;
  }
}

// VdbQueryVoxelDense.ih:115-164 (node indices for x-fastest numbering).
void queryVoxelDense(const OracleGrid &g, int32_t x, int32_t y, int32_t z, uint32_t depth, uint32_t attr,
                     const Ranges &q, VoxelDesc &desc)
{
  desc.level = depth;
  const V3u o{(uint32_t)x, (uint32_t)y, (uint32_t)z};
  if (!isInDomain(g.activeSize, o)) {
    desc.isEmpty = true;
    return;
  }
  const uint64_t n1x = (g.activeSize.x + kRes[1] - 1) >> kTotalLogRes[1];
  const uint64_t n1y = (g.activeSize.y + kRes[1] - 1) >> kTotalLogRes[1];
  const uint64_t l1 = (uint64_t)(o.z >> kTotalLogRes[1]) * n1y * n1x +
                      (uint64_t)(o.y >> kTotalLogRes[1]) * n1x + (o.x >> kTotalLogRes[1]);
  uint64_t nodeIdx;
  int lvl;
  if (depth >= 2) {
    const uint64_t n2x = (g.activeSize.x + kRes[2] - 1) >> kTotalLogRes[2];
    const uint64_t n2y = (g.activeSize.y + kRes[2] - 1) >> kTotalLogRes[2];
    const uint64_t l2 = (uint64_t)(o.z >> kTotalLogRes[2]) * n2y * n2x +
                        (uint64_t)(o.y >> kTotalLogRes[2]) * n2x + (o.x >> kTotalLogRes[2]);
    nodeIdx = l1 * kNumVoxels[1] + l2 * kNumVoxels[2] + offsetToLinear(2, o.x, o.y, o.z);
    lvl     = 2;
  } else if (depth == 1) {
    nodeIdx = l1 * kNumVoxels[1] + offsetToLinear(1, o.x, o.y, o.z);
    lvl     = 1;
  } else {
    nodeIdx = offsetToLinear(0, o.x, o.y, o.z);
    lvl     = 0;
  }
  desc.valueRange = g.levels[lvl].valueRange[nodeIdx * g.numAttributes + attr];
  desc.isEmpty    = !rangesOverlap(q, desc.valueRange);
}

struct IntervalIt
{
  const OracleGrid *g;
  uint32_t attr;
  float tMax, nominalDeltaT;
  Dda dda;
};

// box_utility.ih:62-75.
R1f intersectBox(V3f org, V3f dir, const float *box, R1f limit)
{
  const V3f inv = {divide_safe(dir.x), divide_safe(dir.y), divide_safe(dir.z)};
  const V3f mins = {(box[0] - org.x) * inv.x, (box[1] - org.y) * inv.y, (box[2] - org.z) * inv.z};
  const V3f maxs = {(box[3] - org.x) * inv.x, (box[4] - org.y) * inv.y, (box[5] - org.z) * inv.z};
  R1f r;
  r.lower = fmax_(fmax_(fmin_(mins.x, maxs.x), fmin_(mins.y, maxs.y)),
                  fmax_(fmin_(mins.z, maxs.z), limit.lower));
  r.upper = fmin_(fmin_(fmax_(mins.x, maxs.x), fmax_(mins.y, maxs.y)),
                  fmin_(fmax_(mins.z, maxs.z), limit.upper));
  return r;
}

// VdbIterator_Initialize, VdbIterator.ispc:36-94.
void intervalInit(IntervalIt &it, const OracleGrid &g, uint32_t attr, int depth, V3f origin, V3f direction,
                  R1f tRange)
{
  it.g    = &g;
  it.attr = attr;
  const V3f root{(float)g.rootOrigin.x, (float)g.rootOrigin.y, (float)g.rootOrigin.z};
  const V3f o   = xfmPoint(g.o2i, origin);
  const V3f org = {o.x - root.x, o.y - root.y, o.z - root.z};
  const V3f dir = xfmVector(g.o2i, direction);
  it.nominalDeltaT = fmin_(fmin_(1.f * rcp_safe(std::fabs(dir.x)), 1.f * rcp_safe(std::fabs(dir.y))),
                           1.f * rcp_safe(std::fabs(dir.z)));
  const R1f ti = intersectBox(org, dir, g.domainBox, tRange);
  if (ti.upper <= ti.lower) {
    it.tMax  = -kInf;
    it.dda.t = 0;
  } else {
    it.tMax = ti.upper;
    ddaInit(it.dda, org, dir, ti.lower, (uint32_t)depth);
  }
}

// VdbIterator_iterateIntervalInternal, VdbIterator.ispc:96-159.
bool intervalNext(IntervalIt &it, const Ranges &q, R1f &tOut, R1f &vOut, float &ndt)
{
  if (!(it.dda.t <= it.tMax)) return false;
  const OracleGrid &g = *it.g;
  // Dense node indices assume a single level-1 node; larger volumes use the
  // general query (the HIP device does the same).
  const bool denseQuery = g.dense && g.activeSize.x <= kRes[1] && g.activeSize.y <= kRes[1] &&
                          g.activeSize.z <= kRes[1];
  VoxelDesc v{0, true, emptyRange()};
  uint32_t steps = 0;
  while (it.dda.t <= it.tMax) {
    // A ray that makes no progress (infinite direction components) is ended
    // after 2^22 steps, matching the HIP device's bound.
    if (++steps > (1u << 22)) {
      it.tMax = -kInf;
      return false;
    }
    const uint32_t l = it.dda.level;
    if (denseQuery)
      queryVoxelDense(g, it.dda.idx[OX(l)], it.dda.idx[OY(l)], it.dda.idx[OZ(l)], l, it.attr, q, v);
    else
      queryVoxel(g, it.dda.idx[OX(l)], it.dda.idx[OY(l)], it.dda.idx[OZ(l)], l, it.attr, q, v);
    if (!v.isEmpty) break;
    hddaStepLevel(it.dda, v.level, 0);
  }
  if (it.dda.t < it.tMax) {
    vOut       = v.valueRange;
    tOut.lower = it.dda.t;
    hddaStepLevel(it.dda, v.level, 0);
    tOut.upper = fmin_(it.dda.t, it.tMax);
    ndt        = it.nominalDeltaT;
    return true;
  }
  return false;
}

Ranges makeRanges(int n, const float *r)
{
  Ranges q;
  for (int i = 0; i < n; i++) q.r.push_back({r[2 * i], r[2 * i + 1]});
  if (n > 0) {  // ValueRanges.ih:54-66
    q.minmax = {kInf, -kInf};
    for (const R1f &x : q.r) {
      q.minmax.lower = fmin_(q.minmax.lower, x.lower);
      q.minmax.upper = fmax_(q.minmax.upper, x.upper);
    }
  }
  return q;
}

// --- hit iteration (DefaultIterator.ispc:141-258, Iterator.ih:17-200) -------

struct HitState
{
  IntervalIt iv;
  V3f origin, direction;
  float time;
  R1f cur;      // currentInterval.tRange
  float ndt;    // currentInterval.nominalDeltaT
  float lastHitOffsetT;
};

inline float sampleAt(const OracleGrid &g, V3f o, V3f d, float t, uint32_t attr)
{
  const V3f p{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
  return sampleObject(g, p, attr);
}

// bisect, Iterator.ih:23-85.
float bisect(const OracleGrid &g, V3f o, V3f d, uint32_t attr, float t0, float s0, float t, float s,
             float iso, float tol)
{
  float tMid = 0.f;
  for (int iter = 0; iter < 10; ++iter) {
    tMid = 0.5f * (t0 + t);
    const float err = 0.5f * (t - t0);
    if (err < tol) break;
    const float sm = sampleAt(g, o, d, tMid, attr);
    if (std::isnan(sm)) return (std::fabs(iso - s0) < std::fabs(iso - s)) ? t0 : t;
    if (sm == iso || (iso - s0) * (iso - sm) < 0.f) {
      t = tMid;
      s = sm;
    } else if (s == iso || (iso - s) * (iso - sm) < 0.f) {
      t0 = tMid;
      s0 = sm;
    } else {
      return kInf;
    }
  }
  return tMid;
}

// intersectSurfacesBisection, Iterator.ih:87-197.
bool intersectBisection(const OracleGrid &g, V3f o, V3f d, R1f tr, uint32_t attr, float step,
                        const std::vector<float> &values, float &hitT, float &hitV, float &hitE)
{
  float t0 = tr.lower;
  float s0 = sampleAt(g, o, d, t0, attr);
  for (int it = 0; it < 10 && std::isnan(s0) && t0 < tr.upper; ++it) {
    t0 += fmax_(1e-5f, 1e-5f * step);
    s0 = sampleAt(g, o, d, t0, attr);
  }
  while (t0 < tr.upper) {
    const float h = fmin_(step, tr.upper - t0);
    const float t = t0 + h;
    if (!(t > t0)) break;  // no float progress
    float s       = sampleAt(g, o, d, t, attr);
    float ts      = t;
    for (int it = 0; it < 10 && std::isnan(s) && ts > t0; ++it) {
      ts -= fmax_(1e-5f, 1e-5f * step);
      s = sampleAt(g, o, d, ts, attr);
    }
    float tHit = kInf, eps = kInf, value = kInf;
    for (float v : values) {
      if (std::fabs(s0 - v) < 1e-6f) {
        if (t0 < tHit && t0 <= tr.upper) {
          tHit  = t0;
          value = v;
          eps   = 0.125f * step;
        }
      } else if (!std::isnan(s0 + s) && (v - s0) * (v - s) < 0.f) {
        const float tIso = bisect(g, o, d, attr, t0, s0, ts, s, v, fmin_(1e-6f, 0.01f * h));
        if (tIso < tHit && tIso <= tr.upper) {
          tHit  = tIso;
          value = v;
          eps   = 0.125f * step;
        }
      }
      if (t == tr.upper && std::fabs(s - v) < 1e-6f) {
        if (ts < tHit && ts <= tr.upper) {
          tHit  = ts;
          value = v;
          eps   = 0.125f * step;
        }
      }
    }
    if (tHit < kInf) {
      hitT = tHit;
      hitV = value;
      hitE = eps * std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
      return true;
    }
    t0 = t;
    s0 = s;
  }
  return false;
}

}  // namespace

// --- C API --------------------------------------------------------------------

extern "C" {

const char *oracle_last_error(void) { return g_lastError.c_str(); }

float oracle_half_to_float(uint16_t h) { return halfToFloat(h); }

void oracle_tricubic_matrix(int32_t *out)
{
  const Tricubic &T = tricubic();
  for (int r = 0; r < 64; r++)
    for (int c = 0; c < 64; c++) out[r * 64 + c] = T.A[r][c];
}

int oracle_map_depth(float hint)
{
  hint = std::max(std::min(1.f, hint), 0.f);
  static const std::pair<float, int> table[4] = {{0.f, 0}, {0.2f, 1}, {0.4f, 2}, {0.8f, 3}};
  int depth = -1;
  for (const auto &m : table) {
    if (hint >= m.first)
      depth = m.second;
    else
      break;
  }
  return depth;
}

OracleGrid *oracle_dense_new(const int32_t dims[3], const int32_t indexOrigin[3], const float i2o[12],
                             int cellCentered, uint32_t numAttributes, const uint32_t *dtypes,
                             const void *const *data, const float *background)
{
  return oracle_dense_new_temporal(dims, indexOrigin, i2o, cellCentered, numAttributes, dtypes, data, background,
                                   0, 0, nullptr, nullptr, 0);
}

OracleGrid *oracle_dense_new_temporal(const int32_t dims[3], const int32_t indexOrigin[3], const float i2o[12],
                                      int cellCentered, uint32_t numAttributes, const uint32_t *dtypes,
                                      const void *const *data, const float *background, uint32_t temporalFormat,
                                      uint32_t numTimesteps, const uint64_t *tuIndices, const float *tuTimes,
                                      uint64_t numTimes)
{
  try {
    auto *g = new OracleGrid();
    g->dense            = true;
    g->constantCellData = cellCentered != 0;
    g->numAttributes    = numAttributes;
    g->dims             = {dims[0], dims[1], dims[2]};
    std::memcpy(g->i2o, i2o, sizeof(g->i2o));
    invertAffine(g->i2o, g->o2i);
    const uint64_t nvox = (uint64_t)dims[0] * dims[1] * dims[2];
    g->temporalFormat   = temporalFormat;
    g->numTimesteps     = numTimesteps;
    uint64_t items      = nvox;
    if (temporalFormat == 1) items = nvox * numTimesteps;
    if (temporalFormat == 2) {
      g->tuIndices.assign(tuIndices, tuIndices + nvox + 1);
      g->tuTimes.assign(tuTimes, tuTimes + numTimes);
      items = g->tuIndices[nvox];
    }
    for (uint32_t a = 0; a < numAttributes; a++) {
      g->attrTypes.push_back(dtypes[a]);
      const size_t bytes = items * typeSize(dtypes[a]);
      g->denseData.emplace_back((const uint8_t *)data[a], (const uint8_t *)data[a] + bytes);
      g->background.push_back(background ? background[a] : vkl_bg());
    }
    g->rootOrigin = {indexOrigin[0], indexOrigin[1], indexOrigin[2]};
    g->activeSize = {(uint32_t)dims[0], (uint32_t)dims[1], (uint32_t)dims[2]};

    // One leaf per 8^3 block, x fastest (DenseVdbVolume.cpp:66-119).
    LeafInput in;
    const int32_t lx = (dims[0] + 7) / 8, ly = (dims[1] + 7) / 8, lz = (dims[2] + 7) / 8;
    for (int32_t k = 0; k < lz; k++)
      for (int32_t j = 0; j < ly; j++)
        for (int32_t i = 0; i < lx; i++) {
          in.origin.push_back({i * 8 + indexOrigin[0], j * 8 + indexOrigin[1], k * 8 + indexOrigin[2]});
          in.level.push_back(3);
          in.format.push_back(kFormatDense);
        }
    g->numLeaves = in.origin.size();
    std::vector<V3u> offsets(g->numLeaves);
    for (uint64_t i = 0; i < g->numLeaves; i++) offsets[i] = toDomainOffset(in.origin[i], g->rootOrigin);
    buildLevels(*g, in, offsets);
    buildValueRanges(*g, in, offsets);

    // VdbVolume.cpp:752-765.
    float box[6];
    for (int i = 0; i < 3; i++) {
      box[i]     = (float)indexOrigin[i];
      box[3 + i] = (float)(indexOrigin[i] + dims[i] - (cellCentered ? 0 : 1));
    }
    finishBounds(*g, box);
    return g;
  } catch (const std::exception &e) {
    g_lastError = e.what();
    return nullptr;
  }
}

OracleGrid *oracle_vdb_new(uint64_t numLeaves, const uint32_t *level, const int32_t *origin,
                           const uint32_t *format, uint32_t numAttributes, const uint32_t *dtypes, int packed,
                           const void *const *packedDense, const void *const *packedTile,
                           const void *const *leafData, const float i2o[12], const float *background,
                           const int32_t *clip)
{
  return oracle_vdb_new_temporal(numLeaves, level, origin, format, numAttributes, dtypes, packed, packedDense,
                                 packedTile, leafData, i2o, background, clip, nullptr, nullptr, nullptr, nullptr,
                                 nullptr);
}

OracleGrid *oracle_vdb_new_temporal(uint64_t numLeaves, const uint32_t *level, const int32_t *origin,
                                    const uint32_t *format, uint32_t numAttributes, const uint32_t *dtypes,
                                    int packed, const void *const *packedDense, const void *const *packedTile,
                                    const void *const *leafData, const float i2o[12], const float *background,
                                    const int32_t *clip, const uint32_t *leafTF, const uint32_t *leafSteps,
                                    const uint64_t *const *leafIdx, const float *const *leafTimes,
                                    const uint64_t *leafTimeCounts)
{
  try {
    if (numLeaves == 0) throw std::runtime_error("Vdb volumes must have at least one leaf node.");
    auto *g = new OracleGrid();
    g->dense            = false;
    g->constantCellData = true;
    g->numAttributes    = numAttributes;
    g->numLeaves        = numLeaves;
    g->packed           = packed != 0;
    std::memcpy(g->i2o, i2o, sizeof(g->i2o));
    invertAffine(g->i2o, g->o2i);
    LeafInput in;
    uint64_t nd = 0, nt = 0;
    for (uint64_t i = 0; i < numLeaves; i++) {
      in.level.push_back(level[i]);
      in.format.push_back(format[i]);
      in.origin.push_back({origin[3 * i], origin[3 * i + 1], origin[3 * i + 2]});
      if (format[i] == kFormatDense) {
        if (level[i] + 1 < (uint32_t)kNumLevels)
          throw std::runtime_error("leaf nodes are only supported on the lowest level.");
        nd++;
      } else if (format[i] == kFormatTile) {
        nt++;
      } else {
        throw std::runtime_error("invalid format specified");
      }
    }
    for (uint32_t a = 0; a < numAttributes; a++) {
      g->attrTypes.push_back(dtypes[a]);
      g->background.push_back(background ? background[a] : vkl_bg());
      const size_t es = typeSize(dtypes[a]);
      if (g->packed) {
        g->packedDense.emplace_back((const uint8_t *)packedDense[a],
                                    (const uint8_t *)packedDense[a] + nd * 512 * es);
        if (packedTile && packedTile[a])
          g->packedTile.emplace_back((const uint8_t *)packedTile[a], (const uint8_t *)packedTile[a] + nt * es);
        else if (nt > 0)
          throw std::runtime_error("tile nodes need packed tile data");
        else
          g->packedTile.emplace_back();
      }
    }
    if (leafTF) {
      g->leafTF.assign(leafTF, leafTF + numLeaves);
      g->leafSteps.assign(numLeaves, 0);
      g->leafIdx.resize(numLeaves);
      g->leafTimes.resize(numLeaves);
      for (uint64_t i = 0; i < numLeaves; i++) {
        const uint64_t voxels = format[i] == kFormatDense ? 512 : 1;
        if (leafTF[i] == 1) g->leafSteps[i] = leafSteps[i];
        if (leafTF[i] == 2) {
          g->leafIdx[i].assign(leafIdx[i], leafIdx[i] + voxels + 1);
          g->leafTimes[i].assign(leafTimes[i], leafTimes[i] + leafTimeCounts[i]);
        }
      }
    }
    if (!g->packed) {
      for (uint64_t i = 0; i < numLeaves; i++)
        for (uint32_t a = 0; a < numAttributes; a++) {
          const uint64_t voxels = format[i] == kFormatDense ? 512 : 1;
          uint64_t items        = voxels;
          if (leafTF && leafTF[i] == 1) items = voxels * leafSteps[i];
          if (leafTF && leafTF[i] == 2) items = g->leafIdx[i][voxels];
          const size_t n = items * typeSize(dtypes[a]);
          const uint8_t *p = (const uint8_t *)leafData[i * numAttributes + a];
          g->leafData.emplace_back(p, p + n);
        }
    }
    // computeBbox / computeRootOrigin, VdbVolume.cpp:71-80, 567-581.
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (uint64_t i = 0; i < numLeaves; i++)
      for (int c = 0; c < 3; c++) {
        lo[c] = std::min(lo[c], origin[3 * i + c]);
        hi[c] = std::max(hi[c], origin[3 * i + c] + (int32_t)kRes[level[i]]);
      }
    for (int c = 0; c < 3; c++)
      if ((uint32_t)(hi[c] - lo[c]) > kRes[0])
        throw std::runtime_error("input leaves do not fit into a single root level node");
    g->rootOrigin = {(int32_t)kRes[1] * (int32_t)std::floor(lo[0] / (float)kRes[1]),
                     (int32_t)kRes[1] * (int32_t)std::floor(lo[1] / (float)kRes[1]),
                     (int32_t)kRes[1] * (int32_t)std::floor(lo[2] / (float)kRes[1])};
    g->activeSize = {(uint32_t)(hi[0] - g->rootOrigin.x), (uint32_t)(hi[1] - g->rootOrigin.y),
                     (uint32_t)(hi[2] - g->rootOrigin.z)};
    std::vector<V3u> offsets(numLeaves);
    for (uint64_t i = 0; i < numLeaves; i++) offsets[i] = toDomainOffset(in.origin[i], g->rootOrigin);
    buildLevels(*g, in, offsets);
    buildValueRanges(*g, in, offsets);
    float box[6] = {(float)lo[0], (float)lo[1], (float)lo[2], (float)hi[0], (float)hi[1], (float)hi[2]};
    if (clip && clip[0] <= clip[3] && clip[1] <= clip[4] && clip[2] <= clip[5])
      for (int i = 0; i < 6; i++) box[i] = (float)clip[i];
    finishBounds(*g, box);
    return g;
  } catch (const std::exception &e) {
    g_lastError = e.what();
    return nullptr;
  }
}

void oracle_free(OracleGrid *g) { delete g; }

void oracle_set_cubic_double(OracleGrid *g, int on) { g->cubicDouble = on != 0; }

void oracle_set_sampler(OracleGrid *g, int filter, int gradientFilter, int maxSamplingDepth)
{
  g->filter           = filter;
  g->gradientFilter   = gradientFilter;
  g->maxSamplingDepth = (uint32_t)std::min(maxSamplingDepth, 3);
}

void oracle_sample_times(const OracleGrid *g, uint64_t n, const float *xyz, const float *times, float *out,
                         uint32_t attr)
{
  for (uint64_t i = 0; i < n; i++) {
    g->curTime = times ? times[i] : 0.f;
    out[i]     = sampleObject(*g, V3f{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, attr);
  }
  g->curTime = 0.f;
}

void oracle_gradient_times(const OracleGrid *g, uint64_t n, const float *xyz, const float *times, float *out3,
                           uint32_t attr)
{
  for (uint64_t i = 0; i < n; i++) {
    g->curTime  = times ? times[i] : 0.f;
    const V3f r = gradientObject(*g, V3f{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, attr);
    out3[3 * i]     = r.x;
    out3[3 * i + 1] = r.y;
    out3[3 * i + 2] = r.z;
  }
  g->curTime = 0.f;
}

void oracle_sample(const OracleGrid *g, uint64_t n, const float *xyz, float *out, uint32_t attr)
{
  g->curTime = 0.f;
  for (uint64_t i = 0; i < n; i++)
    out[i] = sampleObject(*g, V3f{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, attr);
}

void oracle_sample_mt(const OracleGrid *g, uint64_t n, const float *xyz, float *out, uint32_t attr, int threads)
{
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  const uint64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; t++) {
    const uint64_t b = std::min<uint64_t>(n, t * chunk), e = std::min<uint64_t>(n, b + chunk);
    pool.emplace_back([=] { oracle_sample(g, e - b, xyz + 3 * b, out + b, attr); });
  }
  for (auto &th : pool) th.join();
}

void oracle_sample_m(const OracleGrid *g, uint64_t n, const float *xyz, float *out, uint32_t m,
                     const uint32_t *attrs)
{
  g->curTime = 0.f;
  for (uint64_t i = 0; i < n; i++)
    for (uint32_t a = 0; a < m; a++)
      out[i * m + a] = sampleObject(*g, V3f{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, attrs[a]);
}

void oracle_gradient(const OracleGrid *g, uint64_t n, const float *xyz, float *out3, uint32_t attr)
{
  g->curTime = 0.f;
  for (uint64_t i = 0; i < n; i++) {
    const V3f r = gradientObject(*g, V3f{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}, attr);
    out3[3 * i]     = r.x;
    out3[3 * i + 1] = r.y;
    out3[3 * i + 2] = r.z;
  }
}

void oracle_bounding_box(const OracleGrid *g, float out[6]) { std::memcpy(out, g->bbox, sizeof(g->bbox)); }

void oracle_set_leaf_access(OracleGrid *g, uint32_t *buffer) { g->leafAccess = buffer; }

// VdbInnerNodeObserver.cpp:19-42 (visitVoxels) and :96-186 (commit), in the
// sequential order level, node, x, y, z. maxDepth is clamped to 2, the
// deepest level with voxels.
uint64_t oracle_inner_nodes(const OracleGrid *g, int maxDepth, float *out, uint64_t capacity)
{
  const uint32_t md   = std::min<uint32_t>((uint32_t)maxDepth, 2u);
  const uint32_t nf   = 6 + 2 * g->numAttributes;
  uint64_t count      = 0;
  for (uint32_t l = 0; l <= md && !g->dense; ++l) {
    const OracleLevel &L   = g->levels[l];
    const uint32_t logRes  = kLogRes[l];
    const uint32_t res     = 1u << logRes;
    const uint32_t ext     = kRes[l + 1];
    for (uint64_t n = 0; n < L.numNodes; ++n)
      for (uint32_t vx = 0; vx < res; ++vx)
        for (uint32_t vy = 0; vy < res; ++vy)
          for (uint32_t vz = 0; vz < res; ++vz) {
            const uint64_t vidx = (n << (3 * logRes)) + ((uint64_t)vx << (2 * logRes)) + ((uint64_t)vy << logRes) + vz;
            const uint64_t voxel = L.voxels[vidx];
            if (!(voxelIsLeaf(voxel) || (l == md && !voxelIsEmpty(voxel)))) continue;
            if (out && count < capacity) {
              float *node = out + count * nf;
              const int32_t ox = g->rootOrigin.x + (int32_t)(L.origin[n].x + ext * vx);
              const int32_t oy = g->rootOrigin.y + (int32_t)(L.origin[n].y + ext * vy);
              const int32_t oz = g->rootOrigin.z + (int32_t)(L.origin[n].z + ext * vz);
              const V3f lo{(float)ox, (float)oy, (float)oz};
              const V3f hi{lo.x + (float)ext, lo.y + (float)ext, lo.z + (float)ext};
              float b[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
              for (int i = 0; i < 8; ++i) {
                const V3f c = xfmPoint(g->i2o, V3f{(i & 1) ? hi.x : lo.x, (i & 2) ? hi.y : lo.y, (i & 4) ? hi.z : lo.z});
                b[0] = fmin_(b[0], c.x);
                b[1] = fmin_(b[1], c.y);
                b[2] = fmin_(b[2], c.z);
                b[3] = fmax_(b[3], c.x);
                b[4] = fmax_(b[4], c.y);
                b[5] = fmax_(b[5], c.z);
              }
              std::memcpy(node, b, sizeof(b));
              for (uint32_t a = 0; a < g->numAttributes; ++a) {
                node[6 + 2 * a] = L.valueRange[vidx * g->numAttributes + a].lower;
                node[7 + 2 * a] = L.valueRange[vidx * g->numAttributes + a].upper;
              }
            }
            ++count;
          }
  }
  return count;
}

void oracle_value_range(const OracleGrid *g, uint32_t attr, float out[2])
{
  out[0] = g->volumeRange[attr].lower;
  out[1] = g->volumeRange[attr].upper;
}

uint32_t oracle_num_attributes(const OracleGrid *g) { return g->numAttributes; }

void oracle_grid_info(const OracleGrid *g, int32_t rootOrigin[3], uint32_t activeSize[3], float domainBox[6],
                      float o2i[12])
{
  std::memcpy(rootOrigin, &g->rootOrigin, 12);
  std::memcpy(activeSize, &g->activeSize, 12);
  std::memcpy(domainBox, g->domainBox, 24);
  std::memcpy(o2i, g->o2i, 48);
}

uint64_t oracle_level_num_nodes(const OracleGrid *g, int l) { return g->levels[l].numNodes; }

void oracle_level_copy(const OracleGrid *g, int l, uint32_t *origin, uint64_t *voxels, float *ranges)
{
  const OracleLevel &L = g->levels[l];
  if (origin) std::memcpy(origin, L.origin.data(), L.origin.size() * sizeof(V3u));
  if (voxels) std::memcpy(voxels, L.voxels.data(), L.voxels.size() * 8);
  if (ranges) std::memcpy(ranges, L.valueRange.data(), L.valueRange.size() * sizeof(R1f));
}

uint32_t oracle_iterate_intervals(const OracleGrid *g, uint32_t attr, int depth, int numRanges,
                                  const float *ranges, const float org[3], const float dir[3],
                                  const float tRange[2], float *out, uint32_t maxOut)
{
  const Ranges q = makeRanges(numRanges, ranges);
  IntervalIt it;
  intervalInit(it, *g, attr, depth, V3f{org[0], org[1], org[2]}, V3f{dir[0], dir[1], dir[2]},
               R1f{tRange[0], tRange[1]});
  uint32_t n = 0;
  R1f t, v;
  float ndt;
  while (n < maxOut && intervalNext(it, q, t, v, ndt)) {
    out[5 * n]     = t.lower;
    out[5 * n + 1] = t.upper;
    out[5 * n + 2] = v.lower;
    out[5 * n + 3] = v.upper;
    out[5 * n + 4] = ndt;
    n++;
  }
  return n;
}

// vklInitHitIterator: the interval iterator plus resetInterval
// (Interval.ih:15-22).
void hitStart(HitState &h, const OracleGrid &g, uint32_t attr, int depth, V3f org, V3f dir, R1f tRange)
{
  h.origin    = org;
  h.direction = dir;
  intervalInit(h.iv, g, attr, depth, h.origin, h.direction, tRange);
  h.cur            = R1f{1.f, -1.f};
  h.ndt            = 0.f;
  h.lastHitOffsetT = -kInf;
}

// One DefaultHitIterator_iterateHit call (DefaultIterator.ispc:168-258).
bool hitStep(const OracleGrid &g, const Ranges &q, uint32_t attr, const std::vector<float> &vals, HitState &h,
             float &ht, float &hv, float &he)
{
  if (vals.empty() || h.cur.lower == kInf) return false;
  for (;;) {
    const bool need = h.cur.upper <= h.cur.lower;
    bool have       = !need;
    if (need) {
      R1f t, v;
      float ndt;
      have = intervalNext(h.iv, q, t, v, ndt);
      if (have) {
        h.cur = t;
        h.ndt = ndt;
      }
    }
    if (need && have) {
      h.cur.lower = fmax_(h.cur.lower, h.lastHitOffsetT);
      if (h.cur.upper <= h.cur.lower) continue;
    }
    if (!have) {
      h.cur.lower = kInf;
      return false;
    }
    if (intersectBisection(g, h.origin, h.direction, h.cur, attr, h.ndt, vals, ht, hv, he)) {
      h.cur.lower      = ht + he;
      h.lastHitOffsetT = ht + he;
      return true;
    }
    h.cur.lower = kInf;
  }
}

Ranges isoRanges(const std::vector<float> &vals)
{
  std::vector<float> pr;
  for (float v : vals) {
    pr.push_back(v);
    pr.push_back(v);
  }
  return makeRanges((int)vals.size(), pr.data());
}

// Hit iteration of one ray at the grid's current sample time.
static uint32_t iterateHits(const OracleGrid *g, uint32_t attr, int depth, int numValues, const float *values,
                            const float org[3], const float dir[3], const float tRange[2], float *out,
                            uint32_t maxOut)
{
  if (depth < 0) depth = g->dense ? oracle_map_depth(0.5f) : oracle_map_depth(1.f);
  const std::vector<float> vals(values, values + numValues);
  const Ranges q = isoRanges(vals);
  HitState h;
  hitStart(h, *g, attr, depth, V3f{org[0], org[1], org[2]}, V3f{dir[0], dir[1], dir[2]},
           R1f{tRange[0], tRange[1]});
  uint32_t n = 0;
  float ht, hv, he;
  while (n < maxOut && hitStep(*g, q, attr, vals, h, ht, hv, he)) {
    out[3 * n]     = ht;
    out[3 * n + 1] = hv;
    out[3 * n + 2] = he;
    n++;
  }
  return n;
}

uint32_t oracle_iterate_hits(const OracleGrid *g, uint32_t attr, int depth, int numValues, const float *values,
                             const float org[3], const float dir[3], const float tRange[2], float time,
                             float *out, uint32_t maxOut)
{
  g->curTime = time;  // temporal volumes sample at the ray time
  return iterateHits(g, attr, depth, numValues, values, org, dir, tRange, out, maxOut);
}

// Many rays on `threads` host threads (0 = all): interval or hit iteration
// per ray as oracle_iterate_intervals / oracle_iterate_hits (no value
// selector, t range [0, inf)), at most maxPerRay results per ray, written to
// out[r * maxPerRay * width] (width 5 for intervals, 3 for hits) with the
// count in counts[r]. For timing the CPU baseline of the batched iterators
// and checking their output.
void oracle_iterate_many(const OracleGrid *g, uint32_t attr, int depth, int hits, int numValues,
                         const float *values, uint64_t numRays, const float *org, const float *dir,
                         uint32_t maxPerRay, uint32_t threads, uint32_t *counts, float *out)
{
  const uint64_t nt = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
  const float tRange[2] = {0.f, kInf};
  g->curTime = 0.f;
  const int hitDepth = depth >= 0 ? depth : (g->dense ? oracle_map_depth(0.5f) : oracle_map_depth(1.f));
  const uint64_t width = hits ? 3 : 5;
  std::vector<std::thread> pool;
  for (uint64_t t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      for (uint64_t r = t; r < numRays; r += nt) {
        float *o = out + r * maxPerRay * width;
        counts[r] = hits ? iterateHits(g, attr, hitDepth, numValues, values, org + 3 * r, dir + 3 * r, tRange, o,
                                       maxPerRay)
                         : oracle_iterate_intervals(g, attr, depth, 0, nullptr, org + 3 * r, dir + 3 * r,
                                                    tRange, o, maxPerRay);
      }
    });
  for (auto &th : pool) th.join();
}

// ---- Woodcock (delta) tracking --------------------------------------------
// Free-flight sampling of the density path tracer
// (examples/interactive/renderer/density_path_tracer/DensityPathTracer.ispc:
// 61-110), with the renderer's TEA generator (Random.ih:13-46) and transfer
// function (TransferFunction.ih:17-47).

// TEA with `rounds` rounds and key k (Wheeler & Needham 1994); the renderer
// uses 8 rounds and its own key words.
void oracle_tea(uint32_t *v, const uint32_t *k, uint32_t rounds)
{
  uint32_t a = v[0], b = v[1], sum = 0;
  for (uint32_t i = 0; i < rounds; ++i) {
    sum += 0x9e3779b9u;
    a += ((b << 4) + k[0]) ^ (b + sum) ^ ((b >> 5) + k[1]);
    b += ((a << 4) + k[2]) ^ (a + sum) ^ ((a >> 5) + k[3]);
  }
  v[0] = a;
  v[1] = b;
}

// The device library's logarithm: x = (1 + f) 2^e, 1 + f in [sqrt(1/2),
// sqrt(2)); s = f / (2 + f); ln(1 + f) = 2s + s R with R the atanh series
// 2s^2/3 + ... + 2s^8/9 (Abramowitz & Stegun 4.1.27), rearranged with
// 2s = f - s f to f - (f^2/2 - s (f^2/2 + R)); ln 2 = 0x3f317180 + 9.0580013515e-6 (ln 2 minus it).
// Binary32 in this operation order.
float oracle_flight_log(float x)
{
  if (!(x > 0.f)) return x == 0.f ? -kInf : std::numeric_limits<float>::quiet_NaN();
  if (x == kInf) return x;
  int e = 0;
  if (x < kFltMin) {
    x = x * 8388608.f;
    e = -23;
  }
  uint32_t b;
  std::memcpy(&b, &x, 4);
  e += (int)(b >> 23) - 127;
  const uint32_t mb = (b & 0x7fffffu) | 0x3f800000u;
  float m;
  std::memcpy(&m, &mb, 4);
  if (m > 1.41421354f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.f;
  const float s = f / (2.f + f);
  const float z = s * s;
  float series  = 0.222222222222f;
  series        = series * z + 0.285714285714f;
  series        = series * z + 0.4f;
  series        = series * z + 0.666666666667f;
  series        = series * z;
  const float halfF2 = 0.5f * f * f;
  const float ke     = (float)e;
  return ke * 0.693138123f - ((halfF2 - (s * (halfF2 + series) + ke * 9.0580013515e-6f)) - f);
}

static float tfOpacity(const float *tf, uint32_t n, float lo, float hi, float v)
{
  if (std::isnan(v) || n == 0) return 0.f;
  if (v <= lo) return tf[3];
  if (v >= hi) return tf[4 * (n - 1) + 3];
  v               = (v - lo) / (hi - lo) * ((float)n - 1.f);
  // floor(v) kept inside [0, n - 1], so a non-finite range cannot index out
  // of the table (for finite ranges floor(v) is already inside).
  float fi = std::floor(v);
  if (!(fi >= 0.f)) fi = 0.f;
  if (fi > (float)(n - 1)) fi = (float)(n - 1);
  const int i     = (int)fi;
  const float rem = v - fi;
  const int j     = std::min(i + 1, (int)n - 1);
  return (1.f - rem) * tf[4 * i + 3] + rem * tf[4 * j + 3];
}

void oracle_woodcock(const OracleGrid *g, uint64_t numRays, const float *org, const float *dir, const float *tRange,
                     uint32_t *rng, float sigmaMax, const float *tf, uint32_t tfSize, float tfLo, float tfHi,
                     float time, float shutter, uint32_t maxCollisions, uint32_t attr, float *eventT,
                     float *samples, uint32_t *count)
{
  static const uint32_t key[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};
  const float toFloat = 2.3283064365386962890625e-10f;
  for (uint64_t r = 0; r < numRays; ++r) {
    const V3f o{org[3 * r], org[3 * r + 1], org[3 * r + 2]};
    const V3f d{dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]};
    float t = tRange[2 * r];
    const float tfar = tRange[2 * r + 1];
    uint32_t *st = rng + 2 * r;
    float sample = 0.f, hitT = kInf;
    uint32_t n = 0;
    bool hit = false;
    if (sigmaMax > 0.f) {
      hitT = std::numeric_limits<float>::quiet_NaN();
      while (n < maxCollisions) {
        oracle_tea(st, key, 8);
        const float u0 = (float)st[0] * toFloat, u1 = (float)st[1] * toFloat;
        oracle_tea(st, key, 8);
        const float u2 = (float)st[0] * toFloat;
        t = t + -oracle_flight_log(1.f - u0) / sigmaMax;
        if (t > tfar) {
          hitT = kInf;
          break;
        }
        ++n;
        float tm = time;
        if (shutter != 0.f) tm = tm + (u2 - 0.5f) * shutter;
        tm = tm < 0.f ? 0.f : tm;
        tm = tm > 1.f ? 1.f : tm;
        g->curTime = tm;
        sample = sampleObject(*g, V3f{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z}, attr);
        const float sigmaT = sigmaMax * tfOpacity(tf, tfSize, tfLo, tfHi, sample);
        if (u1 < sigmaT / sigmaMax) {
          hitT = t;
          hit  = true;
          break;
        }
      }
    }
    eventT[r]  = hitT;
    samples[r] = hit ? sample : 0.f;
    if (count) count[r] = n;
  }
  g->curTime = 0.f;
}

// ---- interval ray marching with compositing --------------------------------
// The ray-march iterator renderer (examples/interactive/renderer/
// ray_march_iterator/RayMarchIterator.ispc:39-137) with the renderer's
// transfer function (TransferFunction.ih:17-47).

static void tfRgba(const float *tf, uint32_t n, float lo, float hi, float v, float c[4])
{
  if (std::isnan(v) || n == 0) {
    c[0] = c[1] = c[2] = c[3] = 0.f;
    return;
  }
  if (v <= lo || v >= hi) {
    const float *e = v <= lo ? tf : tf + 4 * (n - 1);
    for (int k = 0; k < 4; ++k) c[k] = e[k];
    return;
  }
  v        = (v - lo) / (hi - lo) * ((float)n - 1.f);
  float fi = std::floor(v);
  if (!(fi >= 0.f)) fi = 0.f;
  if (fi > (float)(n - 1)) fi = (float)(n - 1);
  const int i     = (int)fi;
  const int j     = std::min(i + 1, (int)n - 1);
  const float rem = v - fi;
  for (int k = 0; k < 4; ++k) c[k] = (1.f - rem) * tf[4 * i + k] + rem * tf[4 * j + k];
}

void oracle_ray_march_intervals(const OracleGrid *g, uint32_t attr, int depth, int numRanges, const float *ranges,
                                uint64_t numRays, const float *org, const float *dir, const float *tRanges,
                                float time, float samplingRate, const float *tf, uint32_t tfSize, float tfLo,
                                float tfHi, uint32_t maxSamples, float *rgba, uint32_t *count)
{
  const Ranges q = makeRanges(numRanges, ranges);
  g->curTime     = time;
  for (uint64_t r = 0; r < numRays; ++r) {
    const V3f o{org[3 * r], org[3 * r + 1], org[3 * r + 2]};
    const V3f d{dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]};
    const R1f tr = tRanges ? R1f{tRanges[2 * r], tRanges[2 * r + 1]} : R1f{0.f, kInf};
    IntervalIt it;
    intervalInit(it, *g, attr, depth, o, d, tr);
    float color[3] = {0.f, 0.f, 0.f}, alpha = 0.f;
    uint32_t taken = 0;
    R1f tOut, vOut;
    float ndt;
    while (alpha < 0.99f && taken < maxSamples && intervalNext(it, q, tOut, vOut, ndt)) {
      const float step = ndt / samplingRate;
      float lo = tOut.lower;
      float hi = (tOut.lower + step < tOut.upper) ? tOut.lower + step : tOut.upper;
      while (hi - lo > 0.f && alpha < 0.99f && taken < maxSamples) {
        const float t  = 0.5f * (lo + hi);
        const float dt = hi - lo;
        const float s  = sampleObject(*g, V3f{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z}, attr);
        ++taken;
        float c[4];
        tfRgba(tf, tfSize, tfLo, tfHi, s, c);
        float op = c[3] * dt;
        op       = op > 0.f ? op : 0.f;
        op       = op < 1.f ? op : 1.f;
        for (int k = 0; k < 3; ++k) color[k] = color[k] + (1.f - alpha) * (c[k] * op);
        alpha = alpha + (1.f - alpha) * op;
        lo    = hi;
        hi    = (lo + step < tOut.upper) ? lo + step : tOut.upper;
      }
    }
    rgba[4 * r]     = color[0];
    rgba[4 * r + 1] = color[1];
    rgba[4 * r + 2] = color[2];
    rgba[4 * r + 3] = alpha;
    if (count) count[r] = taken;
  }
  g->curTime = 0.f;
}

// ---- hit-iterator renderer ---------------------------------------------------
// examples/interactive/renderer/hit_iterator_renderer/HitIteratorRenderer.ispc:
// 51-150: per hit a gradient normal, ambient 0.1 plus two Lambertian lights
// with one shadow hit query each, transfer-function albedo, surface opacity
// 0.6, until opacity 0.99. normalize(v) = v (1 / sqrt(v . v)).

static V3f normalize3(V3f v)
{
  const float inv = 1.f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  return V3f{v.x * inv, v.y * inv, v.z * inv};
}
static float dot3(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

void oracle_hit_render(const OracleGrid *g, uint32_t attr, int numValues, const float *values, uint64_t numRays,
                       const float *org, const float *dir, const float *tRanges, float time, const float *tf,
                       uint32_t tfSize, float tfLo, float tfHi, uint32_t maxHits, float *rgba, uint32_t *count)
{
  g->curTime      = time;
  const int depth = g->dense ? oracle_map_depth(0.5f) : oracle_map_depth(1.f);
  const std::vector<float> vals(values, values + numValues);
  const Ranges q         = isoRanges(vals);
  const V3f light[2]     = {normalize3(V3f{1.f, 1.f, 1.f}), normalize3(V3f{1.f, 1.f, -1.f})};
  const float emission[2] = {1.f, 0.8f};
  for (uint64_t r = 0; r < numRays; ++r) {
    const V3f o{org[3 * r], org[3 * r + 1], org[3 * r + 2]};
    const V3f d{dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]};
    const R1f tr = tRanges ? R1f{tRanges[2 * r], tRanges[2 * r + 1]} : R1f{0.f, kInf};
    HitState h;
    hitStart(h, *g, attr, depth, o, d, tr);
    const V3f wi   = normalize3(V3f{-1.f * d.x, -1.f * d.y, -1.f * d.z});
    float color[3] = {0.f, 0.f, 0.f}, alpha = 0.f;
    uint32_t n     = 0;
    float ht, hv, he;
    while (alpha < 0.99f && n < maxHits && hitStep(*g, q, attr, vals, h, ht, hv, he)) {
      ++n;
      const V3f c = {o.x + ht * d.x, o.y + ht * d.y, o.z + ht * d.z};
      V3f N       = normalize3(gradientObject(*g, c, attr));
      if (std::isnan(N.x) || std::isnan(N.y) || std::isnan(N.z)) N = V3f{0.f, 0.f, 0.f};
      const float ci = dot3(N, wi);
      float illum    = 0.f;
      if (std::sqrt(dot3(N, N)) > 0.f) {
        illum = 0.1f;
        for (int i = 0; i < 2; ++i) {
          const float co = dot3(N, light[i]);
          if ((co > 0.f) == (ci > 0.f)) {
            HitState sh;
            hitStart(sh, *g, attr, depth, c, light[i], R1f{he, kInf});
            float st, sv, se;
            if (!hitStep(*g, q, attr, vals, sh, st, sv, se)) illum += std::fabs(co) * emission[i];
          }
        }
      } else {
        illum = 1.f;
      }
      float albedo[4];
      tfRgba(tf, tfSize, tfLo, tfHi, hv, albedo);
      for (int k = 0; k < 3; ++k) color[k] = color[k] + (1.f - alpha) * illum * albedo[k];
      alpha = alpha + (1.f - alpha) * 0.6f;
    }
    rgba[4 * r]     = color[0];
    rgba[4 * r + 1] = color[1];
    rgba[4 * r + 2] = color[2];
    rgba[4 * r + 3] = alpha;
    if (count) count[r] = n;
  }
  g->curTime = 0.f;
}

}  // extern "C"
