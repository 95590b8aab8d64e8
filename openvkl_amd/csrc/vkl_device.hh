// Device functions of the HIP sampling / traversal path (gfx950).
//
// Numerics: compiled with -ffp-contract=off so every binary32 operation is
// rounded on its own, in the operation order of the reference ISPC sources
// cited below (files under openvkl/devices/cpu/). Nearest / linear samples,
// gradients and interval iteration are therefore bit-identical to the
// reference algorithm evaluated without FMA contraction; tricubic uses the
// separable Catmull-Rom form (mathematically equal to the Lekien-Marsden
// matrix on central-difference constraints; agrees to ~1e-6 relative).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "vkl_views.h"

namespace vklhip {

#define VKL_DEV __device__ __forceinline__

// ---- scalar helpers -----------------------------------------------------------

VKL_DEV float vmin(float a, float b) { return a < b ? a : b; }  // ISPC min
VKL_DEV float vmax(float a, float b) { return a > b ? a : b; }  // ISPC max
VKL_DEV float vinf() { return __int_as_float(0x7f800000); }

// rkcommon lerp: (1 - t) * a + t * b.
VKL_DEV float lerp(float t, float a, float b) { return (1.f - t) * a + t * b; }

// floor then float -> int32 with NaN / out-of-range -> INT32_MIN (the x86
// cvttss2si result the reference CPU device produces).
VKL_DEV int32_t floorToInt(float f)
{
  const float fl = floorf(f);
  return (fl >= -2147483648.f && fl < 2147483648.f) ? (int32_t)fl : INT32_MIN;
}

// Workgroup index remapped so each of the 8 XCDs (which take workgroups
// round-robin, b -> XCD b mod 8, each with its own L2) walks one contiguous
// band of the launch: neighbouring rays, which share voxel cache lines, then
// meet in one L2 instead of eight. A bijection on [0, 8 floor(n / 8)); the
// tail keeps its index.
VKL_DEV uint32_t xcdBandBlock(uint32_t b, uint32_t n)
{
  const uint32_t per = n / 8u;
  return b < per * 8u ? (b % 8u) * per + b / 8u : b;
}

// rkcommon divide_safe: 1 / x with |x| < FLT_MIN clamped to +-FLT_MIN.
VKL_DEV float divideSafe(float x)
{
  const float fltMin = 1.17549435e-38f;
  const float d      = (fabsf(x) < fltMin) ? (x >= 0.f ? fltMin : -fltMin) : x;
  return 1.f / d;
}

template <typename T>
VKL_DEV float toFloat(T v)
{
  return (float)v;
}
template <>
VKL_DEV float toFloat<__half>(__half v)
{
  return __half2float(v);
}

VKL_DEV float fetchAny(const void *p, uint32_t type, uint64_t i)
{
  switch (type) {
  case kTypeUchar: return (float)((const uint8_t *)p)[i];
  case kTypeShort: return (float)((const int16_t *)p)[i];
  case kTypeUshort: return (float)((const uint16_t *)p)[i];
  case kTypeHalf: return __half2float(((const __half *)p)[i]);
  case kTypeDouble: return (float)((const double *)p)[i];
  default: return ((const float *)p)[i];
  }
}

struct F3
{
  float x, y, z;
};
struct I3
{
  int32_t x, y, z;
};

// VdbGrid.h:56-89.
VKL_DEV F3 xfmVector(const float *M, F3 p)
{
  return {M[0] * p.x + M[1] * p.y + M[2] * p.z, M[3] * p.x + M[4] * p.y + M[5] * p.z,
          M[6] * p.x + M[7] * p.y + M[8] * p.z};
}
VKL_DEV F3 xfmPoint(const float *M, F3 p)
{
  const F3 r = xfmVector(M, p);
  return {r.x + M[9], r.y + M[10], r.z + M[11]};
}
VKL_DEV F3 xfmNormal(const float *MI, F3 p)
{
  return {MI[0] * p.x + MI[3] * p.y + MI[6] * p.z, MI[1] * p.x + MI[4] * p.y + MI[7] * p.z,
          MI[2] * p.x + MI[5] * p.y + MI[8] * p.z};
}

// vdb_topology.h.in:96-170: voxel index inside a level-l node, z fastest.
VKL_DEV uint32_t offsetToLinear(uint32_t l, uint32_t x, uint32_t y, uint32_t z)
{
  const uint32_t m = kRes[l] - 1, s = kTotalLogRes[l + 1], L = kLogRes[l];
  return (((x & m) >> s) << (2 * L)) + (((y & m) >> s) << L) + ((z & m) >> s);
}

// ---- voxel sources --------------------------------------------------------------

// Dense voxels (structuredRegular), VdbSampler_dense.ih:290-330 and
// VdbSampler_traverseAndSample.ih:314-349: out-of-domain corners read the
// background. Addressing per DenseLayout (x-fastest or brick-swizzled): each
// axis contributes an independent term, so the 8 trilinear corners cost 6
// axis terms and 8 adds.
// Element offset of root-relative voxel (x, y, z) under layout L.
VKL_DEV uint64_t denseOffset(const DenseLayout &L, uint32_t x, uint32_t y, uint32_t z)
{
  const uint32_t S = L.sx + L.sy + L.sz;
  return (((((uint64_t)(x >> L.sx)) << S) + (x & ((1u << L.sx) - 1))) << (L.cornerPacked ? 3 : 0)) +
         (uint64_t)(y >> L.sy) * L.strideY + ((y & ((1u << L.sy) - 1)) << L.sx) +
         (uint64_t)(z >> L.sz) * L.strideZ + ((z & ((1u << L.sz) - 1)) << (L.sx + L.sy));
}

// The 8 elements of a corner-packed cell record, one vector load.
template <typename T>
struct alignas(8 * sizeof(T)) CellRecord
{
  T v[8];
};

template <typename T, bool kIdx64>
struct DenseSource
{
  using Off = typename std::conditional<kIdx64, uint64_t, uint32_t>::type;
  const T *data;
  uint32_t nx, ny, nz;
  int32_t rx, ry, rz;
  uint32_t sx, sy, sz, S, pk;
  Off strideY, strideZ;
  float bg;

  VKL_DEV DenseSource(const GridView &g, const AttrView &a, uint32_t = 0)
      : data((const T *)a.data),
        nx(g.activeSize[0]),
        ny(g.activeSize[1]),
        nz(g.activeSize[2]),
        rx(g.rootOrigin[0]),
        ry(g.rootOrigin[1]),
        rz(g.rootOrigin[2]),
        sx(g.layout.sx),
        sy(g.layout.sy),
        sz(g.layout.sz),
        S(g.layout.sx + g.layout.sy + g.layout.sz),
        pk(g.layout.cornerPacked ? 3u : 0u),
        strideY((Off)g.layout.strideY),
        strideZ((Off)g.layout.strideZ),
        bg(a.background)
  {
  }

  VKL_DEV void setTime(float) {}  // temporally constant

  VKL_DEV Off xTerm(uint32_t x) const { return (((Off)(x >> sx) << S) | (Off)(x & ((1u << sx) - 1))) << pk; }
  VKL_DEV Off yTerm(uint32_t y) const
  {
    return (Off)(y >> sy) * strideY + (Off)((y & ((1u << sy) - 1)) << sx);
  }
  VKL_DEV Off zTerm(uint32_t z) const
  {
    return (Off)(z >> sz) * strideZ + (Off)((z & ((1u << sz) - 1)) << (sx + sy));
  }

  VKL_DEV float at(int32_t x, int32_t y, int32_t z) const
  {
    const uint32_t ox = (uint32_t)(x - rx), oy = (uint32_t)(y - ry), oz = (uint32_t)(z - rz);
    if (!(ox < nx && oy < ny && oz < nz)) return bg;
    return toFloat(data[xTerm(ox) + yTerm(oy) + zTerm(oz)]);
  }

  // Corners s[4x + 2y + z] of the cell at ic (VdbSampler_trilinear.ih:48-108):
  // one bounds test when the whole stencil is inside, else per corner.
  VKL_DEV void corners8(I3 ic, float *s) const
  {
    const uint32_t ox = (uint32_t)(ic.x - rx), oy = (uint32_t)(ic.y - ry), oz = (uint32_t)(ic.z - rz);
    if (ox + 1 < nx && oy + 1 < ny && oz + 1 < nz && ox < nx && oy < ny && oz < nz) {
      if (pk) {  // corner-packed: the cell record holds s[0..7]
        const CellRecord<T> c = *reinterpret_cast<const CellRecord<T> *>(data + xTerm(ox) + yTerm(oy) + zTerm(oz));
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] = toFloat(c.v[i]);
        return;
      }
      const Off x0 = xTerm(ox), x1 = xTerm(ox + 1);
      const Off y0 = yTerm(oy), y1 = yTerm(oy + 1);
      const Off z0 = zTerm(oz), z1 = zTerm(oz + 1);
      const Off a00 = y0 + z0, a01 = y0 + z1, a10 = y1 + z0, a11 = y1 + z1;
      s[0] = toFloat(data[x0 + a00]);
      s[4] = toFloat(data[x1 + a00]);
      s[2] = toFloat(data[x0 + a10]);
      s[6] = toFloat(data[x1 + a10]);
      s[1] = toFloat(data[x0 + a01]);
      s[5] = toFloat(data[x1 + a01]);
      s[3] = toFloat(data[x0 + a11]);
      s[7] = toFloat(data[x1 + a11]);
      return;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] = at(ic.x + ((i >> 2) & 1), ic.y + ((i >> 1) & 1), ic.z + (i & 1));
  }
};

// Sparse VDB node table (VdbSampler_traverseAndSample.ih:85-202,
// VdbSampleInner.ih.in:33-77). Leaf data is packed: dense leaves of 512
// z-fastest voxels, tiles of one value. kObserve: the sampler has a
// LeafNodeAccess observer; traversals record their leaf and the corner-packed
// cell records (which bypass traversal) are not used. A separate
// instantiation keeps the unobserved kernels free of the marking code, which
// costs the hit iterator about 27 VGPRs.
template <typename T, bool kObserve = false>
struct SparseSource
{
  LevelView lv[3];
  const uint64_t *leafSlot;
  const T *dense;
  const T *tile;
  const float *cells;         // corner-packed cell records (or null)
  const uint32_t *leafIndex;  // leaf cell -> dense-leaf slot (corner-packed)
  // The same leaf-index grid used as a flattened node table: with full-depth
  // sampling (no error voxels) a voxel of a dense leaf is found with one
  // lookup instead of a three-level traversal (null: traverse).
  const uint32_t *flatIndex;
  uint32_t *leafAccess;       // observed samplers: per leaf index, last access stamp
  int32_t root[3];
  uint32_t size[3];
  uint32_t ldim[3];
  uint32_t maxDepth, numAttr, attr, accessStamp;
  float bg;

  VKL_DEV SparseSource(const GridView &g, const AttrView &a, uint32_t attributeIndex)
      : leafSlot(g.leafSlot),
        dense((const T *)a.data),
        tile((const T *)a.tile),
        cells(!kObserve && g.maxSamplingDepth == g.packedDepth ? a.cells : nullptr),
        leafIndex(g.leafIndex),
        flatIndex(!kObserve && g.maxSamplingDepth >= 3u ? g.leafIndex : nullptr),
        leafAccess(g.leafAccess),
        maxDepth(g.maxSamplingDepth),
        numAttr(g.numAttributes),
        attr(attributeIndex),
        accessStamp(g.accessStamp),
        bg(a.background)
  {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      lv[i]   = g.levels[i];
      root[i] = g.rootOrigin[i];
      size[i] = g.activeSize[i];
      ldim[i] = g.leafDims[i];
    }
  }

  // Voxel code for an in-domain root-relative offset; error voxels report
  // their level's value ranges and voxel offset (VdbSampleInner.ih.in:46-54).
  // The three levels are written out: a loop over lv[l] that the compiler
  // does not unroll indexes lv at run time, which places the whole source in
  // scratch memory.
  VKL_DEV uint64_t traverse(uint32_t ox, uint32_t oy, uint32_t oz, const Range1f *&errRanges, uint32_t &errOffset) const
  {
    const uint32_t vo0 = (uint32_t)offsetToLinear(0, ox, oy, oz);
    const uint64_t v0  = lv[0].voxels[vo0];
    if (((uint32_t)v0 & 3u) >= 2u && maxDepth < 1u) return stop(lv[0].valueRange, vo0, errRanges, errOffset);
    if (((uint32_t)v0 & 3u) != 2u) return found(v0);
    const uint32_t vo1 = (uint32_t)((v0 >> 2) * kNumVoxels[1]) + offsetToLinear(1, ox, oy, oz);
    const uint64_t v1  = lv[1].voxels[vo1];
    if (((uint32_t)v1 & 3u) >= 2u && maxDepth < 2u) return stop(lv[1].valueRange, vo1, errRanges, errOffset);
    if (((uint32_t)v1 & 3u) != 2u) return found(v1);
    const uint32_t vo2 = (uint32_t)((v1 >> 2) * kNumVoxels[2]) + offsetToLinear(2, ox, oy, oz);
    const uint64_t v2  = lv[2].voxels[vo2];
    if (((uint32_t)v2 & 3u) >= 2u && maxDepth < 3u) return stop(lv[2].valueRange, vo2, errRanges, errOffset);
    return found(v2);
  }

  // Error voxel: the stopping level's value ranges and the voxel offset.
  VKL_DEV static uint64_t stop(const Range1f *ranges, uint32_t vo, const Range1f *&errRanges, uint32_t &errOffset)
  {
    errRanges = ranges;
    errOffset = vo;
    return 1u;
  }

  // Leaf access observation (VdbSampleInner.ih.in:69-74): every traversal
  // that ends on a leaf pointer records the leaf index. Concurrent lanes
  // store the same stamp, so the races are benign.
  VKL_DEV uint64_t found(uint64_t v) const
  {
    if (kObserve && ((uint32_t)v & 3u) == 3u && leafAccess) leafAccess[v >> 6] = accessStamp;
    return v;
  }

  VKL_DEV void setTime(float) {}  // temporally constant

  VKL_DEV uint64_t slotOf(uint64_t v) const { return leafSlot ? leafSlot[v >> 6] : (v >> 6); }

  VKL_DEV float errorValue(const Range1f *er, uint32_t eo) const
  {
    const Range1f r = er[(uint64_t)eo * numAttr + attr];
    return 0.5f * (r.lower + r.upper);
  }

  VKL_DEV float leafValue(uint64_t v, uint32_t ox, uint32_t oy, uint32_t oz) const
  {
    const uint64_t leaf = slotOf(v);
    if (((v >> 2) & 3u) == 1u) return toFloat(dense[leaf * 512u + (((ox & 7u) << 6) | ((oy & 7u) << 3) | (oz & 7u))]);
    return toFloat(tile[leaf]);
  }

  VKL_DEV bool inDomain(uint32_t ox, uint32_t oy, uint32_t oz) const
  {
    return ox < size[0] && oy < size[1] && oz < size[2];
  }

  // Leaf index grid entry of the 8^3 leaf cell holding root-relative voxel
  // (ox, oy, oz): a dense-leaf slot, kLeafEmpty or kLeafTraverse.
  VKL_DEV uint32_t flatSlot(uint32_t ox, uint32_t oy, uint32_t oz) const
  {
    return flatIndex[((uint64_t)(oz >> 3) * ldim[1] + (oy >> 3)) * ldim[0] + (ox >> 3)];
  }

  // VdbSampler_sample_inner, VdbSampler_traverseAndSample.ih:144-197.
  VKL_DEV float at(int32_t x, int32_t y, int32_t z) const
  {
    const uint32_t ox = (uint32_t)(x - root[0]), oy = (uint32_t)(y - root[1]), oz = (uint32_t)(z - root[2]);
    if (!inDomain(ox, oy, oz)) return bg;
    if (flatIndex) {
      const uint32_t slot = flatSlot(ox, oy, oz);
      if (slot < kLeafTraverse)
        return toFloat(dense[(uint64_t)slot * 512u + (((ox & 7u) << 6) | ((oy & 7u) << 3) | (oz & 7u))]);
      if (slot == kLeafEmpty) return 0.f;  // in the domain, covered by no node
    }
    const Range1f *er  = nullptr;
    uint32_t eo        = 0;
    const uint64_t v   = traverse(ox, oy, oz, er, eo);
    const uint32_t tag = (uint32_t)v & 3u;
    if (tag == 3u) return leafValue(v, ox, oy, oz);
    if (tag == 1u) return errorValue(er, eo);
    return 0.f;
  }

  // Corner-packed volumes: a cell inside a dense leaf reads its record (the
  // 8 values at() returns for its corners, baked at commit). Otherwise, when
  // the 2x2x2 stencil lies in one leaf (7/8 per axis), one traversal serves
  // all corners; the values equal eight separate traversals.
  VKL_DEV void corners8(I3 ic, float *s) const
  {
    const uint32_t ox = (uint32_t)(ic.x - root[0]), oy = (uint32_t)(ic.y - root[1]), oz = (uint32_t)(ic.z - root[2]);
    if (cells && inDomain(ox, oy, oz)) {
      const uint32_t slot = leafIndex[((uint64_t)(oz >> 3) * ldim[1] + (oy >> 3)) * ldim[0] + (ox >> 3)];
      if (slot < kLeafTraverse) {
        const CellRecord<float> c = *reinterpret_cast<const CellRecord<float> *>(
            cells + ((uint64_t)slot * 512u + (((ox & 7u) << 6) | ((oy & 7u) << 3) | (oz & 7u))) * 8u);
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] = c.v[i];
        return;
      }
    }
    if (inDomain(ox, oy, oz) && inDomain(ox + 1, oy + 1, oz + 1) && (ox & 7u) != 7u && (oy & 7u) != 7u &&
        (oz & 7u) != 7u) {
      if (flatIndex) {
        const uint32_t slot = flatSlot(ox, oy, oz);
        if (slot < kLeafTraverse) {
          leafCorners(((uint64_t)slot << 6) | 7u, nullptr, 0, ox, oy, oz, s, true);
          return;
        }
        if (slot == kLeafEmpty) {  // the stencil lies in one empty cell
#pragma unroll
          for (int i = 0; i < 8; i++) s[i] = 0.f;
          return;
        }
      }
      const Range1f *er = nullptr;
      uint32_t eo       = 0;
      const uint64_t v  = traverse(ox, oy, oz, er, eo);
      leafCorners(v, er, eo, ox, oy, oz, s);
      return;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] = at(ic.x + ((i >> 2) & 1), ic.y + ((i >> 1) & 1), ic.z + (i & 1));
  }

  // The 8 corners of a stencil inside one 8^3 leaf cell, given that cell's
  // traversal result.
  // slotCode: v carries the dense-leaf slot itself (from the flattened
  // index), not a leaf index to map through leafSlot.
  VKL_DEV void leafCorners(uint64_t v, const Range1f *er, uint32_t eo, uint32_t ox, uint32_t oy, uint32_t oz,
                           float *s, bool slotCode = false) const
  {
    const uint32_t tag = (uint32_t)v & 3u;
    if (tag == 3u && ((v >> 2) & 3u) == 1u) {
      const uint64_t slot = slotCode ? (v >> 6) : slotOf(v);
      const T *p = dense + slot * 512u + (((ox & 7u) << 6) | ((oy & 7u) << 3) | (oz & 7u));
      s[0] = toFloat(p[0]);
      s[1] = toFloat(p[1]);
      s[2] = toFloat(p[8]);
      s[3] = toFloat(p[9]);
      s[4] = toFloat(p[64]);
      s[5] = toFloat(p[65]);
      s[6] = toFloat(p[72]);
      s[7] = toFloat(p[73]);
      return;
    }
    const float c = (tag == 3u) ? toFloat(tile[slotOf(v)]) : (tag == 1u) ? errorValue(er, eo) : 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] = c;
  }
};

// interpolateTemporallyStructured, temporal_data_interpolation.ih:13-45:
// voxel v owns samples [steps * v, steps * v + steps) of data. Time is
// clamped to [0, 1] (NaN to 0), which keeps both reads among the voxel's own
// samples; inside [0, 1] this is the reference's arithmetic.
template <typename T>
VKL_DEV float temporallyStructured(const T *data, uint32_t steps, uint64_t v, float time)
{
  const float tc      = vmin(vmax(time, 0.f), 1.f);
  const float tf      = tc * (float)(steps - 1);
  const uint32_t ts   = (uint32_t)tf;
  const float w       = tf - (float)ts;
  const uint32_t next = (w == 0.f) ? ts : ts + 1;
  const uint64_t base = (uint64_t)steps * v;
  return (1.f - w) * toFloat(data[base + ts]) + w * toFloat(data[base + next]);
}

// interpolateTemporallyUnstructured, temporal_data_interpolation.ih:82-152:
// voxel v owns positions [idx[v], idx[v + 1]) of data and times.
template <typename T, typename I>
VKL_DEV float temporallyUnstructured(const T *data, const I *idx, const float *times, uint64_t v, float time)
{
  const uint64_t i0 = idx[v], i1 = idx[v + 1];
  const float t0    = times[i0];
  if (time <= t0) return toFloat(data[i0]);
  const float t1 = times[i1 - 1];
  if (time >= t1) return toFloat(data[i1 - 1]);
  if (isnan(time)) return time;
  const uint64_t n = i1 - i0;
  if (n == 2) {
    const float w = (time - t0) * (1.f / (t1 - t0));
    return (1.f - w) * toFloat(data[i0]) + w * toFloat(data[i1 - 1]);
  }
  // Binary search for [times[lower], times[upper]] around time. Commit
  // checked that the voxel's times increase strictly, so with
  // t0 < time < t1 every read stays inside the voxel's range and the search
  // ends within log2(n) + 1 rounds; the bound and the clamps only guard it.
  uint64_t lower = 0, upper = n;
  bool found     = false;
  for (int it = 0; it < 66 && lower <= upper; ++it) {
    const uint64_t mid = (lower + upper) / 2;
    const float tm     = times[i0 + mid];
    const uint64_t nx1 = i0 + mid + 1 < i1 ? i0 + mid + 1 : i1 - 1;
    if (time >= tm && time <= times[nx1]) {
      lower = i0 + mid;
      upper = lower + 1 < i1 ? lower + 1 : i1 - 1;
      found = true;
      break;
    } else if (time < tm) {
      upper = mid;
    } else {
      lower = mid;
    }
  }
  if (!found) return __int_as_float(0x7fc00000);
  const float tl = times[lower], th = times[upper];
  const float w  = (time - tl) * (1.f / (th - tl));
  return (1.f - w) * toFloat(data[lower]) + w * toFloat(data[upper]);
}

// structuredRegular with temporal data (temporal_data_interpolation.ih:13-152,
// VdbSampler_dense.ih:360-560): every voxel read interpolates that voxel's
// time samples at the source's current time (setTime; 0 by default, as for
// calls without times). Voxels are x-fastest (temporal volumes have no bricks
// or corner packing).
template <typename T>
struct TemporalDenseSource
{
  const T *data;
  const uint32_t *idx32;
  const uint64_t *idx64;
  const float *times;
  uint32_t nx, ny, nz;
  int32_t rx, ry, rz;
  uint32_t format, numTimesteps;
  float bg, time;

  VKL_DEV TemporalDenseSource(const GridView &g, const AttrView &a, uint32_t = 0)
      : data((const T *)a.data),
        idx32(g.tuIndex64 ? nullptr : (const uint32_t *)g.tuIndices),
        idx64(g.tuIndex64 ? (const uint64_t *)g.tuIndices : nullptr),
        times(g.tuTimes),
        nx(g.activeSize[0]),
        ny(g.activeSize[1]),
        nz(g.activeSize[2]),
        rx(g.rootOrigin[0]),
        ry(g.rootOrigin[1]),
        rz(g.rootOrigin[2]),
        format(g.temporalFormat),
        numTimesteps(g.numTimesteps),
        bg(a.background),
        time(0.f)
  {
  }

  VKL_DEV void setTime(float t) { time = t; }

  // The voxel with linear index v at the current time.
  VKL_DEV float voxel(uint64_t v) const
  {
    if (format == 1u) return temporallyStructured(data, numTimesteps, v, time);
    if (idx64) return temporallyUnstructured(data, idx64, times, v, time);
    return temporallyUnstructured(data, idx32, times, v, time);
  }

  VKL_DEV float at(int32_t x, int32_t y, int32_t z) const
  {
    const uint32_t ox = (uint32_t)(x - rx), oy = (uint32_t)(y - ry), oz = (uint32_t)(z - rz);
    if (!(ox < nx && oy < ny && oz < nz)) return bg;
    return voxel(((uint64_t)oz * ny + oy) * nx + ox);
  }

  // A rolled loop: one inlined copy of the time interpolation per stencil
  // instead of eight.
  VKL_DEV void corners8(I3 ic, float *s) const
  {
#pragma unroll 1
    for (int i = 0; i < 8; i++) s[i] = at(ic.x + ((i >> 2) & 1), ic.y + ((i >> 1) & 1), ic.z + (i & 1));
  }
};

// vdb with temporal leaves (VdbSampler_denseZYX.ih / VdbSampler_tile.ih,
// structured and unstructured sampling): the node table is traversed as for
// constant leaves; a leaf pointer's temporal format selects how its voxel is
// read at the source's current time. The traversal records leaf accesses
// when the sampler is observed.
template <typename T>
struct TemporalSparseSource
{
  SparseSource<T, true> nodes;
  const T *data;
  const uint64_t *dataOff, *idxOff, *timesOff, *idx;
  const uint32_t *steps;
  const float *times;
  float time;

  VKL_DEV TemporalSparseSource(const GridView &g, const AttrView &a, uint32_t attributeIndex)
      : nodes(g, a, attributeIndex),
        data((const T *)a.data),
        dataOff(g.leafDataOffset),
        idxOff(g.leafIdxOffset),
        timesOff(g.leafTimesOffset),
        idx((const uint64_t *)g.tuIndices),
        steps(g.leafSteps),
        times(g.tuTimes),
        time(0.f)
  {
    nodes.cells = nullptr;
  }

  VKL_DEV void setTime(float t) { time = t; }

  // Leaf voxel (ox, oy, oz) of the leaf pointer v (tiles: their one voxel).
  VKL_DEV float leafValue(uint64_t v, uint32_t ox, uint32_t oy, uint32_t oz) const
  {
    const uint64_t leaf = v >> 6;
    const uint32_t tf   = (uint32_t)(v >> 4) & 3u;
    const uint64_t vi   = (((v >> 2) & 3u) == 1u) ? (((ox & 7u) << 6) | ((oy & 7u) << 3) | (oz & 7u)) : 0u;
    const T *d          = data + dataOff[leaf];
    if (tf == 1u) return temporallyStructured(d, steps[leaf], vi, time);
    if (tf == 2u) return temporallyUnstructured(d, idx + idxOff[leaf], times + timesOff[leaf], vi, time);
    return toFloat(d[vi]);
  }

  VKL_DEV float at(int32_t x, int32_t y, int32_t z) const
  {
    const uint32_t ox = (uint32_t)(x - nodes.root[0]), oy = (uint32_t)(y - nodes.root[1]),
                   oz = (uint32_t)(z - nodes.root[2]);
    if (!nodes.inDomain(ox, oy, oz)) return nodes.bg;
    const Range1f *er = nullptr;
    uint32_t eo       = 0;
    const uint64_t v  = nodes.traverse(ox, oy, oz, er, eo);
    const uint32_t tag = (uint32_t)v & 3u;
    if (tag == 3u) return leafValue(v, ox, oy, oz);
    if (tag == 1u) return nodes.errorValue(er, eo);
    return 0.f;
  }

  VKL_DEV void corners8(I3 ic, float *s) const
  {
#pragma unroll 1
    for (int i = 0; i < 8; i++) s[i] = at(ic.x + ((i >> 2) & 1), ic.y + ((i >> 1) & 1), ic.z + (i & 1));
  }
};

template <typename T>
VKL_DEV void cubicTaps(const TemporalSparseSource<T> &src, I3 ic, float *S)
{
#pragma unroll 1
  for (int i = 0; i < 64; i++) S[i] = src.at(ic.x + (i >> 4) - 1, ic.y + ((i >> 2) & 3) - 1, ic.z + (i & 3) - 1);
}

// Whether dense offsets (including brick padding) can exceed 32 bits.
VKL_DEV bool denseNeeds64(const GridView &g)
{
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  return nbz * g.layout.strideZ >= (1ull << 32);
}

// ---- filters ------------------------------------------------------------------

// VdbSampler_nearest.ih:25-57: +0.5 for vertex-centered data.
template <typename Src>
VKL_DEV float sampleNearest(const Src &src, F3 p, float off)
{
  return src.at(floorToInt(p.x + off), floorToInt(p.y + off), floorToInt(p.z + off));
}

// VdbSampler_trilinear.ih:378-399.
template <typename Src>
VKL_DEV float sampleLinear(const Src &src, F3 p)
{
  const I3 ic = {floorToInt(p.x), floorToInt(p.y), floorToInt(p.z)};
  const F3 d  = {p.x - (float)ic.x, p.y - (float)ic.y, p.z - (float)ic.z};
  float s[8];
  src.corners8(ic, s);
  return lerp(d.x, lerp(d.y, lerp(d.z, s[0], s[1]), lerp(d.z, s[2], s[3])),
              lerp(d.y, lerp(d.z, s[4], s[5]), lerp(d.z, s[6], s[7])));
}

// VdbSampler_trilinear.ih:751-780 (index-space gradient).
template <typename Src>
VKL_DEV F3 gradientLinear(const Src &src, F3 p)
{
  const I3 ic = {floorToInt(p.x), floorToInt(p.y), floorToInt(p.z)};
  const F3 d  = {p.x - (float)ic.x, p.y - (float)ic.y, p.z - (float)ic.z};
  float s[8];
  src.corners8(ic, s);
  F3 r;
  r.x = lerp(d.y, lerp(d.z, s[4] - s[0], s[5] - s[1]), lerp(d.z, s[6] - s[2], s[7] - s[3]));
  r.y = lerp(d.x, lerp(d.z, s[2] - s[0], s[3] - s[1]), lerp(d.z, s[6] - s[4], s[7] - s[5]));
  r.z = lerp(d.x, lerp(d.y, s[1] - s[0], s[3] - s[2]), lerp(d.y, s[5] - s[4], s[7] - s[6]));
  return r;
}

// Catmull-Rom weights for taps -1, 0, 1, 2 at offset t in [0, 1), and their
// derivatives. The tensor product of these equals the tricubic interpolant
// of VdbSampler_tricubic.ih:34-578 (Hermite on central differences).
VKL_DEV void crWeights(float t, float *w)
{
  const float t2 = t * t, t3 = t2 * t;
  w[0] = 0.5f * (-t3 + 2.f * t2 - t);
  w[1] = 0.5f * (3.f * t3 - 5.f * t2 + 2.f);
  w[2] = 0.5f * (-3.f * t3 + 4.f * t2 + t);
  w[3] = 0.5f * (t3 - t2);
}
VKL_DEV void crDerivWeights(float t, float *w)
{
  const float t2 = t * t;
  w[0] = 0.5f * (-3.f * t2 + 4.f * t - 1.f);
  w[1] = 0.5f * (9.f * t2 - 10.f * t);
  w[2] = 0.5f * (-9.f * t2 + 8.f * t + 1.f);
  w[3] = 0.5f * (3.f * t2 - 2.f * t);
}

template <typename Src>
VKL_DEV void cubicTaps(const Src &src, I3 ic, float *S)
{
#pragma unroll
  for (int i = 0; i < 64; i++) S[i] = src.at(ic.x + (i >> 4) - 1, ic.y + ((i >> 2) & 3) - 1, ic.z + (i & 3) - 1);
}

// Corner-packed dense volumes: the 4x4x4 stencil is 8 cell records (cells at
// ic - 1 + 2 (a, b, c)) when it lies inside the domain; same values as 64
// single reads.
template <typename T, bool kIdx64>
VKL_DEV void cubicTaps(const DenseSource<T, kIdx64> &src, I3 ic, float *S)
{
  const uint32_t ox = (uint32_t)(ic.x - src.rx), oy = (uint32_t)(ic.y - src.ry), oz = (uint32_t)(ic.z - src.rz);
  if (src.pk && ox - 1 < src.nx && ox + 2 < src.nx && oy - 1 < src.ny && oy + 2 < src.ny && oz - 1 < src.nz &&
      oz + 2 < src.nz) {
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const uint32_t a = (c >> 2) & 1, b = (c >> 1) & 1, d = c & 1;
      const CellRecord<T> r = *reinterpret_cast<const CellRecord<T> *>(
          src.data + src.xTerm(ox - 1 + 2 * a) + src.yTerm(oy - 1 + 2 * b) + src.zTerm(oz - 1 + 2 * d));
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint32_t i = 2 * a + ((k >> 2) & 1), j = 2 * b + ((k >> 1) & 1), l = 2 * d + (k & 1);
        S[16 * i + 4 * j + l] = toFloat(r.v[k]);
      }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 64; i++) S[i] = src.at(ic.x + (i >> 4) - 1, ic.y + ((i >> 2) & 3) - 1, ic.z + (i & 3) - 1);
}

// vdb volumes with the flattened leaf index (unobserved, full sampling
// depth): a stencil inside the domain touches at most 2 leaf cells per axis,
// so their 8 index entries are read once and each tap reads its leaf
// directly, instead of one index lookup per tap; the taps equal 64 at()
// calls (dense leaf: its voxel; a cell no node covers: 0). A stencil that
// leaves the domain or meets a tile / error cell takes the 64 at() calls.
template <typename T, bool kObserve>
VKL_DEV void cubicTapsByLeafCells(const SparseSource<T, kObserve> &src, I3 ic, float *S)
{
  const uint32_t ox = (uint32_t)(ic.x - 1 - src.root[0]), oy = (uint32_t)(ic.y - 1 - src.root[1]),
                 oz = (uint32_t)(ic.z - 1 - src.root[2]);
  if (src.flatIndex && src.inDomain(ox, oy, oz) && src.inDomain(ox + 3, oy + 3, oz + 3)) {
    const uint32_t cx0 = ox >> 3, cy0 = oy >> 3, cz0 = oz >> 3;
    const uint32_t cx1 = (ox + 3) >> 3, cy1 = (oy + 3) >> 3, cz1 = (oz + 3) >> 3;
    uint32_t slot[8];
    bool traverse = false;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const uint32_t cx = (c & 4) ? cx1 : cx0, cy = (c & 2) ? cy1 : cy0, cz = (c & 1) ? cz1 : cz0;
      slot[c]  = src.flatIndex[((uint64_t)cz * src.ldim[1] + cy) * src.ldim[0] + cx];
      traverse = traverse || slot[c] == kLeafTraverse;
    }
    if (!traverse) {
      // The cell of each tap by selects (constant register indices only).
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t x  = ox + i;
        const bool hx     = (x >> 3) != cx0;
        const uint32_t t0 = hx ? slot[4] : slot[0], t1 = hx ? slot[5] : slot[1];
        const uint32_t t2 = hx ? slot[6] : slot[2], t3 = hx ? slot[7] : slot[3];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint32_t y  = oy + j;
          const bool hy     = (y >> 3) != cy0;
          const uint32_t u0 = hy ? t2 : t0, u1 = hy ? t3 : t1;
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const uint32_t z  = oz + k;
            const uint32_t sl = (z >> 3) != cz0 ? u1 : u0;
            S[16 * i + 4 * j + k] =
                sl == kLeafEmpty
                    ? 0.f
                    : toFloat(src.dense[(uint64_t)sl * 512u + (((x & 7u) << 6) | ((y & 7u) << 3) | (z & 7u))]);
          }
        }
      }
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < 64; i++) S[i] = src.at(ic.x + (i >> 4) - 1, ic.y + ((i >> 2) & 3) - 1, ic.z + (i & 3) - 1);
}

// Temporal volumes: the 64 reads are a rolled loop. Each read interpolates in
// time (a binary search for unstructured data), and 64 inlined copies of that
// make the kernels that sample them several times larger.
template <typename T>
VKL_DEV void cubicTaps(const TemporalDenseSource<T> &src, I3 ic, float *S)
{
#pragma unroll 1
  for (int i = 0; i < 64; i++) S[i] = src.at(ic.x + (i >> 4) - 1, ic.y + ((i >> 2) & 3) - 1, ic.z + (i & 3) - 1);
}

// The tricubic interpolant at index-space point p from its 64 taps
// S[16 i + 4 j + l] = voxel ic + (i - 1, j - 1, l - 1), ic = floor(p).
VKL_DEV float cubicFromTaps(const float *S, F3 p, I3 ic)
{
  float wx[4], wy[4], wz[4];
  crWeights(p.x - (float)ic.x, wx);
  crWeights(p.y - (float)ic.y, wy);
  crWeights(p.z - (float)ic.z, wz);
  float r = 0.f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float ry = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float *q = S + 16 * i + 4 * j;
      ry += wy[j] * (wz[0] * q[0] + wz[1] * q[1] + wz[2] * q[2] + wz[3] * q[3]);
    }
    r += wx[i] * ry;
  }
  return r;
}

// Its index-space gradient, from the same taps.
VKL_DEV F3 cubicGradientFromTaps(const float *S, F3 p, I3 ic)
{
  float wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
  const F3 d = {p.x - (float)ic.x, p.y - (float)ic.y, p.z - (float)ic.z};
  crWeights(d.x, wx);
  crWeights(d.y, wy);
  crWeights(d.z, wz);
  crDerivWeights(d.x, dx);
  crDerivWeights(d.y, dy);
  crDerivWeights(d.z, dz);
  F3 g = {0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float gy_v = 0.f, gy_dy = 0.f, gy_dz = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float *q = S + 16 * i + 4 * j;
      const float vz = wz[0] * q[0] + wz[1] * q[1] + wz[2] * q[2] + wz[3] * q[3];
      const float dzv = dz[0] * q[0] + dz[1] * q[1] + dz[2] * q[2] + dz[3] * q[3];
      gy_v += wy[j] * vz;
      gy_dy += dy[j] * vz;
      gy_dz += wy[j] * dzv;
    }
    g.x += dx[i] * gy_v;
    g.y += wx[i] * gy_dy;
    g.z += wx[i] * gy_dz;
  }
  return g;
}

template <typename Src>
VKL_DEV float sampleCubic(const Src &src, F3 p)
{
  const I3 ic = {floorToInt(p.x), floorToInt(p.y), floorToInt(p.z)};
  float S[64];
  cubicTaps(src, ic, S);
  return cubicFromTaps(S, p, ic);
}

template <typename Src>
VKL_DEV F3 gradientCubic(const Src &src, F3 p)
{
  const I3 ic = {floorToInt(p.x), floorToInt(p.y), floorToInt(p.z)};
  float S[64];
  cubicTaps(src, ic, S);
  return cubicGradientFromTaps(S, p, ic);
}

// Tricubic in the stream kernels (kSample, kGradient): vdb sources gather
// their taps by leaf cells (cubicTapsByLeafCells). The fused ray and
// iterator kernels keep the per-tap reads of sampleCubic: inlined at their
// several sampler call sites, the unrolled gather makes those kernels'
// compilation take tens of minutes.
template <typename Src>
struct GathersByLeafCells : std::false_type
{
};
template <typename T, bool kObserve>
struct GathersByLeafCells<SparseSource<T, kObserve>> : std::true_type
{
};

template <typename Src>
VKL_DEV float sampleCubicStream(const Src &src, F3 p)
{
  if constexpr (GathersByLeafCells<Src>::value) {
    const I3 ic = {floorToInt(p.x), floorToInt(p.y), floorToInt(p.z)};
    float S[64];
    cubicTapsByLeafCells(src, ic, S);
    return cubicFromTaps(S, p, ic);
  } else {
    return sampleCubic(src, p);
  }
}

template <typename Src>
VKL_DEV F3 gradientCubicStream(const Src &src, F3 p)
{
  if constexpr (GathersByLeafCells<Src>::value) {
    const I3 ic = {floorToInt(p.x), floorToInt(p.y), floorToInt(p.z)};
    float S[64];
    cubicTapsByLeafCells(src, ic, S);
    return cubicGradientFromTaps(S, p, ic);
  } else {
    return gradientCubic(src, p);
  }
}

template <typename Src>
VKL_DEV float sampleFiltered(const Src &src, int filter, F3 p, float nearestOffset)
{
  if (filter == kFilterNearest) return sampleNearest(src, p, nearestOffset);
  if (filter == kFilterCubic) return sampleCubic(src, p);
  return sampleLinear(src, p);
}

// Hit iteration descends to depth 2 on structuredRegular volumes and 3 on vdb
// volumes (HitIteratorContext::commit, IteratorContext.cpp:285-336).
template <typename S>
struct IsDenseSource : std::false_type
{
};
template <typename T, bool B>
struct IsDenseSource<DenseSource<T, B>> : std::true_type
{
};
template <typename T>
struct IsDenseSource<TemporalDenseSource<T>> : std::true_type
{
};
template <typename Src>
constexpr int hitDepthOf()
{
  return IsDenseSource<Src>::value ? 2 : 3;
}

// Object-space sample with one source type and filter per instantiation (the
// sampler entry of VdbSampler_iterator_computeSample_varying,
// VdbSampler.ispc:341-374). Every caller gets its own inlined copy: the
// library has no non-inlined device functions, because branch relaxation in
// a large callee on gfx950 rewrote the return-address register pair s[30:31]
// for its long jumps, which sent the callee's return back into its own body.
template <typename Src, int kFilter>
struct PointSampler
{
  using Source = Src;
  Src src;
  const float *o2i;
  float noff;

  VKL_DEV PointSampler(const GridView &g, const AttrView &a, uint32_t attr)
      : src(g, a, attr), o2i(g.o2i), noff(g.constantCellData ? 0.f : 0.5f)
  {
  }

  VKL_DEV void setTime(float t) { src.setTime(t); }

  VKL_DEV float operator()(F3 oc) const
  {
    const F3 p = xfmPoint(o2i, oc);
    if (kFilter == kFilterNearest) return sampleNearest(src, p, noff);
    if (kFilter == kFilterCubic) return sampleCubic(src, p);
    return sampleLinear(src, p);
  }
};

// ---- interval iteration (VdbIterator.ispc:36-180, Dda.ih, HDDA.ih.in,
//      VdbQueryVoxel.ih.in, VdbQueryVoxelDense.ih) -----------------------------

VKL_DEV int dirSign(float v) { return ((int)(0 < v)) - ((int)(v < 0)); }
VKL_DEV float dirRcp(float v) { return (v == 0) ? vinf() : divideSafe(v); }

// Dda.ih:66-84.
VKL_DEV float planeT(float plane, float org, float inv)
{
  const float t = (plane - org) * inv;
  return (plane == org || isnan(t)) ? vinf() : t;
}

// Dda.ih:127-140.
VKL_DEV int32_t clampCell(float f, int32_t res) { return floorToInt(vmax(f, 0.f)) & ~(res - 1); }

constexpr uint32_t kMaxDdaSteps = 1u << 22;

#define DX(l) (l)
#define DY(l) (4 + (l))
#define DZ(l) (8 + (l))

// Every DDA routine below indexes the state arrays with compile-time
// constants (levels are template arguments or fully unrolled loop indices):
// an array indexed at run time is placed in scratch memory, which costs the
// iterators several times their voxel traffic.

// Per-axis step of the level-L cells: tDelta = res * |1 / dir| (Dda.ih:184-186).
template <int L>
VKL_DEV float ddaTDelta(const DdaState &d, int axis)
{
  return (float)kRes[L + 1] * d.absRcp[axis];
}

// The DDA's deepest level: kDepth when the caller knows it at compile time
// (hit iteration: 2 on structuredRegular, 3 on vdb), else d.level.
template <int kDepth>
VKL_DEV uint32_t ddaLevel(const DdaState &d)
{
  return kDepth >= 0 ? (uint32_t)kDepth : d.level;
}

// Dda.ih:145-196.
template <int kDepth = -1>
VKL_DEV void ddaInit(DdaState &d, F3 org, F3 dir, float tMin, uint32_t level)
{
  const F3 oo = {org.x + tMin * dir.x, org.y + tMin * dir.y, org.z + tMin * dir.z};
  const int sx = dirSign(dir.x), sy = dirSign(dir.y), sz = dirSign(dir.z);
  const F3 id = {dirRcp(dir.x), dirRcp(dir.y), dirRcp(dir.z)};
  d.level     = level < 3u ? level : 3u;
  d.t         = tMin;
  d.rayOrg[0] = org.x;
  d.rayOrg[1] = org.y;
  d.rayOrg[2] = org.z;
  d.rayDir[0] = dir.x;
  d.rayDir[1] = dir.y;
  d.rayDir[2] = dir.z;
  d.absRcp[0] = fabsf(id.x);
  d.absRcp[1] = fabsf(id.y);
  d.absRcp[2] = fabsf(id.z);
  // Levels below d.level are never read; they are written as 0 so that no
  // earlier state stays live through a re-initialisation.
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const bool used   = (uint32_t)l <= ddaLevel<kDepth>(d);
    const int32_t res = (int32_t)kRes[l + 1];
    const int32_t ix = clampCell(oo.x, res), iy = clampCell(oo.y, res), iz = clampCell(oo.z, res);
    d.idx[DX(l)]   = used ? ix : 0;
    d.idx[DY(l)]   = used ? iy : 0;
    d.idx[DZ(l)]   = used ? iz : 0;
    d.tNext[DX(l)] = used ? planeT((float)(ix + res * (sx >= 0)), org.x, id.x) : 0.f;
    d.tNext[DY(l)] = used ? planeT((float)(iy + res * (sy >= 0)), org.y, id.y) : 0.f;
    d.tNext[DZ(l)] = used ? planeT((float)(iz + res * (sz >= 0)), org.z, id.z) : 0.f;
  }
}

// Dda.ih:204-266: levels L + 1 .. d.level restart from the current t, nudged
// 1e-2 across the plane the level-L step crossed.
template <int kDepth, int L>
VKL_DEV void ddaReinitBelow(DdaState &d, int32_t dx, int32_t dy)
{
  const F3 dir = {d.rayDir[0], d.rayDir[1], d.rayDir[2]};
  const F3 id  = {dirRcp(dir.x), dirRcp(dir.y), dirRcp(dir.z)};
  const int sx = dirSign(dir.x), sy = dirSign(dir.y), sz = dirSign(dir.z);
  F3 oo = {d.rayOrg[0] + d.t * dir.x, d.rayOrg[1] + d.t * dir.y, d.rayOrg[2] + d.t * dir.z};
  if (dx != 0)
    oo.x = oo.x + ((float)sx) * 1e-2f;
  else if (dy != 0)
    oo.y = oo.y + ((float)sy) * 1e-2f;
  else
    oo.z = oo.z + ((float)sz) * 1e-2f;
#pragma unroll
  for (int l = L + 1; l < 4; ++l) {
    if ((uint32_t)l > ddaLevel<kDepth>(d)) continue;
    const int32_t res = (int32_t)kRes[l + 1];
    d.idx[DX(l)]      = clampCell(oo.x, res);
    d.idx[DY(l)]      = clampCell(oo.y, res);
    d.idx[DZ(l)]      = clampCell(oo.z, res);
    d.tNext[DX(l)]    = planeT((float)(d.idx[DX(l)] + res * (sx >= 0)), d.rayOrg[0], id.x);
    d.tNext[DY(l)]    = planeT((float)(d.idx[DY(l)] + res * (sy >= 0)), d.rayOrg[1], id.y);
    d.tNext[DZ(l)]    = planeT((float)(d.idx[DZ(l)] + res * (sz >= 0)), d.rayOrg[2], id.z);
  }
}

// Dda.ih:271-309 (ties: x, then y, then z).
template <int L>
VKL_DEV void ddaStep(DdaState &d)
{
  // Every field is written on every path (values selected, not addresses),
  // so the state stays in registers.
  const float tx = d.tNext[DX(L)], ty = d.tNext[DY(L)], tz = d.tNext[DZ(L)];
  const float minT  = vmin(vmin(tx, ty), tz);
  const int32_t res = (int32_t)kRes[L + 1];
  const bool sx     = minT == tx;
  const bool sy     = !sx && minT == ty;
  const bool sz     = !sx && !sy;
  d.t               = sx ? tx : (sy ? ty : tz);
  d.tNext[DX(L)]    = sx ? tx + ddaTDelta<L>(d, 0) : tx;
  d.tNext[DY(L)]    = sy ? ty + ddaTDelta<L>(d, 1) : ty;
  d.tNext[DZ(L)]    = sz ? tz + ddaTDelta<L>(d, 2) : tz;
  d.idx[DX(L)]      = d.idx[DX(L)] + (sx ? dirSign(d.rayDir[0]) * res : 0);
  d.idx[DY(L)]      = d.idx[DY(L)] + (sy ? dirSign(d.rayDir[1]) * res : 0);
  d.idx[DZ(L)]      = d.idx[DZ(L)] + (sz ? dirSign(d.rayDir[2]) * res : 0);
}

// HDDA.ih.in:52-101, parents of a stepped level from the nearest one up:
// each one steps only if its child left its cell.
template <int P>
VKL_DEV void hddaParents(DdaState &d)
{
  if constexpr (P >= 0) {
    constexpr int C   = P + 1;
    const int32_t res = (int32_t)kRes[P + 1];
    const bool left   = d.idx[DX(C)] < d.idx[DX(P)] || d.idx[DY(C)] < d.idx[DY(P)] || d.idx[DZ(C)] < d.idx[DZ(P)] ||
                      d.idx[DX(C)] >= d.idx[DX(P)] + res || d.idx[DY(C)] >= d.idx[DY(P)] + res ||
                      d.idx[DZ(C)] >= d.idx[DZ(P)] + res;
    if (!left) return;
    ddaStep<P>(d);
    hddaParents<P - 1>(d);
  }
}

// HDDA.ih.in:24-101: step on level L (the level of the empty voxel),
// re-initialise the levels below, and let parents follow a child that left
// their cell.
template <int kDepth, int L>
VKL_DEV void hddaStepAt(DdaState &d)
{
  if constexpr (L >= 3) {
    ddaStep<3>(d);
  } else {
    const int32_t ox = d.idx[DX(L)], oy = d.idx[DY(L)];
    ddaStep<L>(d);
    if ((uint32_t)L < ddaLevel<kDepth>(d)) ddaReinitBelow<kDepth, L>(d, d.idx[DX(L)] - ox, d.idx[DY(L)] - oy);
  }
  hddaParents<(L >= 3 ? 3 : L) - 1>(d);
}

// descLevel never exceeds the DDA's deepest level (queries descend no
// deeper), so with a compile-time depth the deeper cases are not generated.
template <int kDepth>
VKL_DEV void hddaStep(DdaState &d, uint32_t descLevel)
{
  constexpr int kMax = kDepth >= 0 ? kDepth : 3;
  switch (descLevel) {
  case 0: hddaStepAt<kDepth, 0>(d); break;
  case 1: hddaStepAt<kDepth, (kMax < 1 ? kMax : 1)>(d); break;
  case 2: hddaStepAt<kDepth, (kMax < 2 ? kMax : 2)>(d); break;
  default: hddaStepAt<kDepth, kMax>(d); break;
  }
}

// Pins a value to a register: the optimizer can no longer turn a select
// between array elements into a load from a selected address (a run-time
// array index, which sends the whole array to scratch memory).
VKL_DEV int32_t opaque(int32_t v)
{
  asm volatile("" : "+v"(v));
  return v;
}

// The cell of the DDA's deepest level, selected without a run-time array
// index.
template <int kDepth>
VKL_DEV I3 ddaCell(const DdaState &d)
{
  if constexpr (kDepth >= 0) return I3{d.idx[DX(kDepth)], d.idx[DY(kDepth)], d.idx[DZ(kDepth)]};
  I3 c = {opaque(d.idx[DX(0)]), opaque(d.idx[DY(0)]), opaque(d.idx[DZ(0)])};
#pragma unroll
  for (int l = 1; l < 4; ++l) {
    const I3 cl = {opaque(d.idx[DX(l)]), opaque(d.idx[DY(l)]), opaque(d.idx[DZ(l)])};
    if (d.level == (uint32_t)l) c = cl;
  }
  return c;
}

// ValueRanges.ih:22-52.
VKL_DEV bool rangesOverlap(const IteratorContextView &c, Range1f r)
{
  if (c.numRanges == 0) return true;
  if (!(c.rangesMinMax.upper >= r.lower && c.rangesMinMax.lower <= r.upper)) return false;
  for (int i = 0; i < c.numRanges; i++)
    if (c.ranges[i].upper >= r.lower && c.ranges[i].lower <= r.upper) return true;
  return false;
}

struct VoxelDesc
{
  uint32_t level;
  bool isEmpty;
  Range1f valueRange;
};

// VdbQueryVoxel.ih.in:42-116. With kNodeCache, a query in the 128^3 block
// of the previous query's level-2 node starts at level 2 (it.ck / it.cnode):
// the level-0 and level-1 voxels on the way there, and so their outcome, are
// the same. Hit iteration uses it (+7 % on the 4096^3 vdb); plain interval
// iteration does not (its kernel runs 11 % slower with the extra state).
template <bool kNodeCache>
VKL_DEV void queryVoxel(const IteratorContextView &c, int32_t x, int32_t y, int32_t z, uint32_t depth,
                        VoxelDesc &desc, IntervalState &it)
{
  const GridView &g = c.sampler.grid;
  uint64_t node     = 0;
  uint32_t first    = 0;
  const int32_t kx = x >> 7, ky = y >> 7, kz = z >> 7;
  if (kNodeCache && it.ck[0] == kx && it.ck[1] == ky && it.ck[2] == kz) {
    node  = it.cnode;
    first = 2;
  }
  for (uint32_t L = first; L < 3; ++L) {
    const uint32_t vo = (uint32_t)(offsetToLinear(L, (uint32_t)x, (uint32_t)y, (uint32_t)z) + node * kNumVoxels[L]);
    const uint64_t v  = g.levels[L].voxels[vo];
    desc.valueRange   = g.levels[L].valueRange[(uint64_t)vo * g.numAttributes + c.attributeIndex];
    desc.level        = L;
    desc.isEmpty      = ((v & 3u) == 0u) || !rangesOverlap(c, desc.valueRange);
    if (desc.isEmpty) return;
    const bool atDepth = (L + 1 > depth);
    const bool isLeaf  = (v & 3u) == 3u;
    if (isLeaf && !atDepth) desc.level = L + 1;
    if (L < 2 && !(atDepth || isLeaf)) {
      node = v >> 2;
      if (kNodeCache && L == 1) {
        it.ck[0]  = kx;
        it.ck[1]  = ky;
        it.ck[2]  = kz;
        it.cnode = (uint32_t)node;
      }
      continue;
    }
    return;
  }
}
// VdbQueryVoxelDense.ih:115-164 (nodes numbered x-fastest, one level-1 node).
VKL_DEV void queryVoxelDense(const IteratorContextView &c, int32_t x, int32_t y, int32_t z, uint32_t depth,
                             VoxelDesc &desc)
{
  const GridView &g = c.sampler.grid;
  desc.level        = depth;
  const uint32_t ox = (uint32_t)x, oy = (uint32_t)y, oz = (uint32_t)z;
  if (!(ox < g.activeSize[0] && oy < g.activeSize[1] && oz < g.activeSize[2])) {
    desc.isEmpty = true;
    return;
  }
  const uint64_t n1x = (g.activeSize[0] + kRes[1] - 1) >> kTotalLogRes[1];
  const uint64_t n1y = (g.activeSize[1] + kRes[1] - 1) >> kTotalLogRes[1];
  const uint64_t l1  = (uint64_t)(oz >> kTotalLogRes[1]) * n1y * n1x + (uint64_t)(oy >> kTotalLogRes[1]) * n1x +
                      (ox >> kTotalLogRes[1]);
  uint64_t vi;
  int lvl;
  if (depth >= 2) {
    const uint64_t n2x = (g.activeSize[0] + kRes[2] - 1) >> kTotalLogRes[2];
    const uint64_t n2y = (g.activeSize[1] + kRes[2] - 1) >> kTotalLogRes[2];
    const uint64_t l2  = (uint64_t)(oz >> kTotalLogRes[2]) * n2y * n2x + (uint64_t)(oy >> kTotalLogRes[2]) * n2x +
                        (ox >> kTotalLogRes[2]);
    vi  = l1 * kNumVoxels[1] + l2 * kNumVoxels[2] + offsetToLinear(2, ox, oy, oz);
    lvl = 2;
  } else if (depth == 1) {
    vi  = l1 * kNumVoxels[1] + offsetToLinear(1, ox, oy, oz);
    lvl = 1;
  } else {
    vi  = offsetToLinear(0, ox, oy, oz);
    lvl = 0;
  }
  desc.valueRange = g.levels[lvl].valueRange[vi * g.numAttributes + c.attributeIndex];
  desc.isEmpty    = !rangesOverlap(c, desc.valueRange);
}

// box_utility.ih:62-75.
VKL_DEV Range1f intersectBox(F3 org, F3 dir, const float *box, Range1f lim)
{
  const F3 inv  = {divideSafe(dir.x), divideSafe(dir.y), divideSafe(dir.z)};
  const F3 mins = {(box[0] - org.x) * inv.x, (box[1] - org.y) * inv.y, (box[2] - org.z) * inv.z};
  const F3 maxs = {(box[3] - org.x) * inv.x, (box[4] - org.y) * inv.y, (box[5] - org.z) * inv.z};
  Range1f r;
  r.lower = vmax(vmax(vmin(mins.x, maxs.x), vmin(mins.y, maxs.y)), vmax(vmin(mins.z, maxs.z), lim.lower));
  r.upper = vmin(vmin(vmax(mins.x, maxs.x), vmax(mins.y, maxs.y)), vmin(vmax(mins.z, maxs.z), lim.upper));
  return r;
}

// VdbIterator_Initialize, VdbIterator.ispc:36-94.
template <int kDepth = -1>
VKL_DEV void intervalInit(const IteratorContextView &c, IntervalState &it, F3 origin, F3 direction, Range1f tRange)
{
  const GridView &g = c.sampler.grid;
  const F3 o        = xfmPoint(g.o2i, origin);
  const F3 org      = {o.x - (float)g.rootOrigin[0], o.y - (float)g.rootOrigin[1], o.z - (float)g.rootOrigin[2]};
  const F3 dir      = xfmVector(g.o2i, direction);
  it.nominalDeltaT  = vmin(vmin(1.f * divideSafe(fabsf(dir.x)), 1.f * divideSafe(fabsf(dir.y))),
                          1.f * divideSafe(fabsf(dir.z)));
  it.ck[0] = it.ck[1] = it.ck[2] = INT32_MIN;
  it.cnode = 0;
  const Range1f ti = intersectBox(org, dir, g.domainBox, tRange);
  // A ray that misses the box gets tMax = -inf (and t = 0), which ends its
  // iteration; the DDA is initialised either way so that every field of the
  // state is written on both paths.
  const bool miss = ti.upper <= ti.lower;
  it.tMax         = miss ? -vinf() : ti.upper;
  ddaInit<kDepth>(it.dda, org, dir, miss ? 0.f : ti.lower, c.maxIteratorDepth);
}

// Whether a context's queries use the dense node numbering: dense node
// indices assume one level-1 node (VdbQueryVoxelDense.ih:48-57); larger
// volumes use the general query, which stays inside the node table.
VKL_DEV bool denseQuery(const GridView &g)
{
  return g.dense != 0 && g.activeSize[0] <= kRes[1] && g.activeSize[1] <= kRes[1] && g.activeSize[2] <= kRes[1];
}

// VdbIterator_iterateIntervalInternal, VdbIterator.ispc:96-159. kQuery: 0
// the general query, 1 the dense one, -1 chosen per context at run time.
template <bool kNodeCache = false, int kDepth = -1, int kQuery = -1>
VKL_DEV bool intervalNext(const IteratorContextView &c, IntervalState &it, IntervalOut &out)
{
  if (!(it.dda.t <= it.tMax)) return false;
  const bool dense = kQuery >= 0 ? kQuery == 1 : denseQuery(c.sampler.grid);
  VoxelDesc v{0, true, {0.f, 0.f}};
  // Every step advances t by a positive tDelta; a ray whose step makes no
  // progress (infinite direction components) is ended after kMaxDdaSteps.
  uint32_t steps = 0;
  while (it.dda.t <= it.tMax) {
    if (++steps > kMaxDdaSteps) {
      it.tMax = -vinf();
      return false;
    }
    const uint32_t l = ddaLevel<kDepth>(it.dda);
    const I3 cell    = ddaCell<kDepth>(it.dda);
    if (dense)
      queryVoxelDense(c, cell.x, cell.y, cell.z, l, v);
    else
      queryVoxel<kNodeCache>(c, cell.x, cell.y, cell.z, l, v, it);
    if (!v.isEmpty) break;
    hddaStep<kDepth>(it.dda, v.level);
  }
  if (it.dda.t < it.tMax) {
    out.vLower = v.valueRange.lower;
    out.vUpper = v.valueRange.upper;
    out.tLower = it.dda.t;
    hddaStep<kDepth>(it.dda, v.level);
    out.tUpper        = vmin(it.dda.t, it.tMax);
    out.nominalDeltaT = it.nominalDeltaT;
    return true;
  }
  return false;
}

// ---- hit iteration (DefaultIterator.ispc:141-258, Iterator.ih:17-200) --------

template <typename S>
VKL_DEV float sampleAlong(const S &smp, F3 o, F3 d, float t)
{
  return smp(F3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z});
}

// bisect, Iterator.ih:23-85.
template <typename S>
VKL_DEV float bisect(const S &smp, F3 o, F3 d, float t0, float s0, float t, float s, float iso, float tol)
{
  float tMid = 0.f;
#pragma unroll 1
  for (int iter = 0; iter < 10; ++iter) {
    tMid            = 0.5f * (t0 + t);
    const float err = 0.5f * (t - t0);
    if (err < tol) break;
    const float sm = sampleAlong(smp, o, d, tMid);
    if (isnan(sm)) return (fabsf(iso - s0) < fabsf(iso - s)) ? t0 : t;
    if (sm == iso || (iso - s0) * (iso - sm) < 0.f) {
      t = tMid;
      s = sm;
    } else if (s == iso || (iso - s) * (iso - sm) < 0.f) {
      t0 = tMid;
      s0 = sm;
    } else {
      return vinf();
    }
  }
  return tMid;
}

// intersectSurfacesBisection, Iterator.ih:87-197.
template <typename S>
VKL_DEV bool intersectBisection(const IteratorContextView &c, const S &smp, F3 o, F3 d, float tLo, float tHi,
                                float step, HitOut &hit)
{
  float t0 = tLo;
  float s0 = sampleAlong(smp, o, d, t0);
  for (int it = 0; it < 10 && isnan(s0) && t0 < tHi; ++it) {
    t0 += vmax(1e-5f, 1e-5f * step);
    s0 = sampleAlong(smp, o, d, t0);
  }
  while (t0 < tHi) {
    const float h = vmin(step, tHi - t0);
    const float t = t0 + h;
    if (!(t > t0)) break;  // no float progress: the march cannot continue
    float s       = sampleAlong(smp, o, d, t);
    float ts      = t;
    for (int it = 0; it < 10 && isnan(s) && ts > t0; ++it) {
      ts -= vmax(1e-5f, 1e-5f * step);
      s = sampleAlong(smp, o, d, ts);
    }
    float tHit = vinf(), eps = vinf(), value = vinf();
    for (int i = 0; i < c.numValues; i++) {
      const float v = c.values[i];
      if (fabsf(s0 - v) < 1e-6f) {
        if (t0 < tHit && t0 <= tHi) {
          tHit  = t0;
          value = v;
          eps   = 0.125f * step;
        }
      } else if (!isnan(s0 + s) && (v - s0) * (v - s) < 0.f) {
        const float tIso = bisect(smp, o, d, t0, s0, ts, s, v, vmin(1e-6f, 0.01f * h));
        if (tIso < tHit && tIso <= tHi) {
          tHit  = tIso;
          value = v;
          eps   = 0.125f * step;
        }
      }
      if (t == tHi && fabsf(s - v) < 1e-6f) {
        if (ts < tHit && ts <= tHi) {
          tHit  = ts;
          value = v;
          eps   = 0.125f * step;
        }
      }
    }
    if (tHit < vinf()) {
      hit.t       = tHit;
      hit.sample  = value;
      hit.epsilon = eps * sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
      return true;
    }
    t0 = t;
    s0 = s;
  }
  return false;
}

template <int kDepth = -1>
VKL_DEV void hitInit(const IteratorContextView &c, HitState &h, F3 origin, F3 direction, Range1f tRange, float time)
{
  intervalInit<kDepth>(c, h.iv, origin, direction, tRange);
  h.origin[0] = origin.x;
  h.origin[1] = origin.y;
  h.origin[2] = origin.z;
  h.direction[0] = direction.x;
  h.direction[1] = direction.y;
  h.direction[2] = direction.z;
  h.time             = time;
  h.curLower         = 1.f;  // resetInterval, Interval.ih:15-22
  h.curUpper         = -1.f;
  h.curNominalDeltaT = 0.f;
  h.lastHitOffsetT   = -vinf();
}

// DefaultHitIterator_iterateHit, DefaultIterator.ispc:168-258.
template <typename S, int kQuery = -1>
VKL_DEV bool hitNext(const IteratorContextView &c, const S &smp, HitState &h, HitOut &hit)
{
  if (c.numValues == 0) return false;
  if (h.curLower == vinf()) return false;
  const F3 o = {h.origin[0], h.origin[1], h.origin[2]};
  const F3 d = {h.direction[0], h.direction[1], h.direction[2]};
  for (;;) {
    const bool need = h.curUpper <= h.curLower;
    bool have       = !need;
    if (need) {
      IntervalOut iv;
      have = intervalNext<true, hitDepthOf<typename S::Source>(), kQuery>(c, h.iv, iv);
      if (have) {
        h.curLower         = iv.tLower;
        h.curUpper         = iv.tUpper;
        h.curNominalDeltaT = iv.nominalDeltaT;
      }
    }
    if (need && have) {
      h.curLower = vmax(h.curLower, h.lastHitOffsetT);
      if (h.curUpper <= h.curLower) continue;
    }
    if (!have) {
      h.curLower = vinf();
      return false;
    }
    hit.t = vinf();
    if (intersectBisection(c, smp, o, d, h.curLower, h.curUpper, h.curNominalDeltaT, hit)) {
      h.curLower       = hit.t + hit.epsilon;
      h.lastHitOffsetT = hit.t + hit.epsilon;
      return true;
    }
    h.curLower = vinf();
  }
}

}  // namespace vklhip
