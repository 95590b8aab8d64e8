// Spatially binned stream sampling for large dense volumes
// (vklComputeSampleN with many points on a volume far larger than L2).
//
// A random point stream gathers 8 corners from 4 different 128-byte lines per
// sample; on a 512^3 float volume nearly every one misses L2, so a direct
// gather moves ~20x the compulsory bytes. Two kernels instead:
//
//  1. kBinScatter: each workgroup takes 16 K points, computes their
//     index-space positions (the same xfmPoint as the direct kernel), bins
//     them by the brick region of the volume they fall in (LDS histogram),
//     reserves room in the bin's segment for this XCD with one global atomic
//     per (bin, workgroup), and appends 16-byte records {p.x, p.y, p.z, index}.
//     Segments are per XCD (workgroup index mod 8) so that a segment's open
//     cache line is only ever written through one L2.
//  2. kBinnedSample: one workgroup per bin stages the bin's voxel brick
//     (region plus a one-voxel halo) in LDS once, then samples every record of
//     the bin from LDS and writes out[index].
//
// Records that do not fit their segment go to an overflow list that the
// second kernel samples with the direct path. Corners outside the domain or
// the staged brick are read by the direct path's own source, so every sample
// equals the direct kernel's bit for bit.

#include "binned.h"
#include "kernels.h"
#include "vkl_device.hh"

namespace vklhip {

constexpr int kScatterBlock = 1024;
constexpr int kScatterPts   = 16;
constexpr int kSampleBlock  = 512;
constexpr int kXcds         = 8;

// Root-relative cell clamped into [0, n - 1] (NaN positions land in cell 0).
__device__ __forceinline__ uint32_t clampAxis(float p, int32_t root, uint32_t n)
{
  const int64_t o = (int64_t)floorToInt(p) - root;
  return o < 0 ? 0u : (o >= (int64_t)n ? n - 1 : (uint32_t)o);
}

__device__ __forceinline__ uint32_t binOf(const BinGeometry &b, const int32_t root[3], const uint32_t size[3], F3 p)
{
  const uint32_t x = clampAxis(p.x, root[0], size[0]) >> b.shift[0];
  const uint32_t y = clampAxis(p.y, root[1], size[1]) >> b.shift[1];
  const uint32_t z = clampAxis(p.z, root[2], size[2]) >> b.shift[2];
  return (z * b.count[1] + y) * b.count[0] + x;
}

__global__ __launch_bounds__(kScatterBlock) void kBinScatter(GridView g, BinGeometry b, uint64_t n,
                                                              const float *__restrict__ xyz, BinScratch s)
{
  // Per-bin counts, then (in place) the bin's base in this XCD's segment.
  extern __shared__ uint32_t hist[];
  const uint32_t nb = b.numBins;
  for (uint32_t i = threadIdx.x; i < nb; i += kScatterBlock) hist[i] = 0;

  // Issue every point load of this thread before using any of them.
  const uint64_t first = (uint64_t)blockIdx.x * (kScatterBlock * kScatterPts) + threadIdx.x;
  F3 p[kScatterPts];
#pragma unroll
  for (int k = 0; k < kScatterPts; k++) {
    const uint64_t i = first + (uint64_t)k * kScatterBlock;
    const uint64_t j = i < n ? i : 0;
    p[k]             = F3{xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2]};
  }
  __syncthreads();

  // bin << 14 | rank within the workgroup (both < 2^14: at most 16384 bins
  // and 16384 points per workgroup).
  static_assert(kScatterBlock * kScatterPts <= (1 << 14), "rank field");
  uint32_t binRank[kScatterPts];
#pragma unroll
  for (int k = 0; k < kScatterPts; k++) {
    const uint64_t i = first + (uint64_t)k * kScatterBlock;
    binRank[k]       = 0xffffffffu;
    if (i < n) {
      p[k]             = xfmPoint(g.o2i, p[k]);
      const uint32_t c = binOf(b, g.rootOrigin, g.activeSize, p[k]);
      binRank[k]       = (c << 14) | atomicAdd(&hist[c], 1u);
    }
  }
  __syncthreads();

  // Reserve segment room: all atomics of this thread in flight together.
  constexpr int kPerThread = 16384 / kScatterBlock;  // bins per thread at most
  const uint32_t xcd = blockIdx.x % kXcds;
  uint32_t base[kPerThread];
#pragma unroll
  for (int j = 0; j < kPerThread; j++) {
    const uint32_t i = threadIdx.x + j * kScatterBlock;
    const uint32_t c = i < nb ? hist[i] : 0u;
    base[j]          = c ? atomicAdd(&s.segCount[i * kXcds + xcd], c) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPerThread; j++) {
    const uint32_t i = threadIdx.x + j * kScatterBlock;
    if (i < nb) hist[i] = base[j];
  }
  __syncthreads();

  // Record stores (no atomics in this loop, so nothing waits on them);
  // points whose segment is full are collected for the overflow list.
  uint32_t full = 0;
#pragma unroll
  for (int k = 0; k < kScatterPts; k++) {
    if (binRank[k] == 0xffffffffu) continue;
    const uint64_t i   = first + (uint64_t)k * kScatterBlock;
    const uint32_t c   = binRank[k] >> 14;
    const uint32_t pos = hist[c] + (binRank[k] & 0x3fffu);
    if (pos < s.capacity) {
      float4 r = {p[k].x, p[k].y, p[k].z, __uint_as_float((uint32_t)i)};
      s.records[((uint64_t)c * kXcds + xcd) * s.capacity + pos] = r;
    } else {
      full |= 1u << k;
    }
  }
  if (full) {
    for (int k = 0; k < kScatterPts; k++)
      if (full & (1u << k)) s.overflow[atomicAdd(s.overflowCount, 1u)] = (uint32_t)(first + (uint64_t)k * kScatterBlock);
  }
}

// Dense source whose taps come from an LDS brick when they lie inside the
// domain and the brick; otherwise from the global source (same values, same
// domain test, so results equal the direct kernels').
template <typename T, bool kIdx64>
struct LdsBrickSource
{
  DenseSource<T, kIdx64> global;
  const T *brick;
  int32_t bx, by, bz;   // brick origin (root-relative voxels; may be -1)
  uint32_t ex, ey, ez;  // brick extents (elements)

  VKL_DEV float at(int32_t x, int32_t y, int32_t z) const
  {
    const uint32_t ox = (uint32_t)(x - global.rx), oy = (uint32_t)(y - global.ry), oz = (uint32_t)(z - global.rz);
    const uint32_t lx = ox - (uint32_t)bx, ly = oy - (uint32_t)by, lz = oz - (uint32_t)bz;
    if (ox < global.nx && oy < global.ny && oz < global.nz && lx < ex && ly < ey && lz < ez)
      return toFloat(brick[(lz * ey + ly) * ex + lx]);
    return global.at(x, y, z);
  }

  VKL_DEV void corners8(I3 ic, float *s) const
  {
    const uint32_t ox = (uint32_t)(ic.x - global.rx), oy = (uint32_t)(ic.y - global.ry),
                   oz = (uint32_t)(ic.z - global.rz);
    const uint32_t lx = ox - (uint32_t)bx, ly = oy - (uint32_t)by, lz = oz - (uint32_t)bz;
    // In-domain test identical to DenseSource::corners8, plus brick cover.
    if (ox + 1 < global.nx && oy + 1 < global.ny && oz + 1 < global.nz && ox < global.nx && oy < global.ny &&
        oz < global.nz && lx + 1 < ex && ly + 1 < ey && lz + 1 < ez) {
      const T *q = brick + (lz * ey + ly) * ex + lx;
      const uint32_t sy = ex, sz = ex * ey;
      s[0] = toFloat(q[0]);
      s[4] = toFloat(q[1]);
      s[2] = toFloat(q[sy]);
      s[6] = toFloat(q[sy + 1]);
      s[1] = toFloat(q[sz]);
      s[5] = toFloat(q[sz + 1]);
      s[3] = toFloat(q[sz + sy]);
      s[7] = toFloat(q[sz + sy + 1]);
      return;
    }
    global.corners8(ic, s);
  }
};

// One point of a binned launch (the direct kernels' arithmetic).
template <int kOp, typename Src>
VKL_DEV void binnedPoint(const GridView &g, const Src &src, F3 p, float noff, uint64_t i, float *out,
                         uint32_t outStride)
{
  if (kOp == kBinGradientLinear || kOp == kBinGradientCubic) {
    const F3 gi = (kOp == kBinGradientCubic) ? gradientCubic(src, p) : gradientLinear(src, p);
    const F3 r  = xfmNormal(g.o2i, gi);
    out[3 * i]     = r.x;
    out[3 * i + 1] = r.y;
    out[3 * i + 2] = r.z;
    return;
  }
  float r;
  if (kOp == kBinSampleNearest)
    r = sampleNearest(src, p, noff);
  else if (kOp == kBinSampleCubic)
    r = sampleCubic(src, p);
  else
    r = sampleLinear(src, p);
  out[i * outStride] = r;
}

template <typename T, bool kIdx64, int kOp>
__global__ __launch_bounds__(kSampleBlock) void kBinnedSample(GridView g, AttrView a, BinGeometry b,
                                                               const float *__restrict__ xyz,
                                                               float *__restrict__ out, uint32_t outStride,
                                                               float noff, BinScratch s)
{
  extern __shared__ __attribute__((aligned(16))) unsigned char ldsRaw[];
  T *brick = reinterpret_cast<T *>(ldsRaw);
  const uint32_t bin = blockIdx.x;
  const uint32_t cx = bin % b.count[0], cy = (bin / b.count[0]) % b.count[1], cz = bin / (b.count[0] * b.count[1]);
  const int32_t lo = (int32_t)b.haloLo;
  const uint32_t w = b.haloLo + b.haloHi;
  LdsBrickSource<T, kIdx64> src{DenseSource<T, kIdx64>(g, a),
                                brick,
                                (int32_t)(cx << b.shift[0]) - lo,
                                (int32_t)(cy << b.shift[1]) - lo,
                                (int32_t)(cz << b.shift[2]) - lo,
                                (1u << b.shift[0]) + w,
                                (1u << b.shift[1]) + w,
                                (1u << b.shift[2]) + w};
  // Stage the brick row by row (a row is ex consecutive x voxels, one per
  // thread), 8 rows in flight per thread; voxels outside the domain are never
  // read from LDS, so their slots stay unwritten.
  const uint32_t perPass = kSampleBlock / src.ex;
  if (threadIdx.x < perPass * src.ex) {
    const uint32_t lx = threadIdx.x % src.ex, r0 = threadIdx.x / src.ex;
    const uint32_t x  = (uint32_t)(src.bx + (int32_t)lx);
    const bool xIn    = x < src.global.nx;
    const auto xOff   = src.global.xTerm(x);
    const uint32_t dly = perPass % src.ey, dlz = perPass / src.ey;
    uint32_t ly = r0 % src.ey, lz = r0 / src.ey;
    while (lz < src.ez) {
      T v[8];
      uint32_t slot[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        // Unconditional loads (element 0 stands in for skipped voxels) so
        // that all eight are in flight together.
        const uint32_t y = (uint32_t)(src.by + (int32_t)ly), z = (uint32_t)(src.bz + (int32_t)lz);
        const bool ok    = lz < src.ez && xIn && y < src.global.ny && z < src.global.nz;
        v[u]             = src.global.data[ok ? xOff + src.global.yTerm(y) + src.global.zTerm(z) : 0];
        slot[u]          = ok ? (lz * src.ey + ly) * src.ex + lx : 0xffffffffu;
        ly += dly;
        lz += dlz;
        if (ly >= src.ey) {
          ly -= src.ey;
          lz++;
        }
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (slot[u] != 0xffffffffu) brick[slot[u]] = v[u];
    }
  }
  __syncthreads();
  // The bin's records: 8 XCD segments, read 4 per thread at a time.
  uint32_t segEnd[kXcds];
  uint32_t total = 0;
#pragma unroll
  for (int q = 0; q < kXcds; q++) {
    total += min(s.segCount[bin * kXcds + q], s.capacity);
    segEnd[q] = total;
  }
  const float4 *rec0 = s.records + (uint64_t)bin * kXcds * s.capacity;
  for (uint32_t k0 = threadIdx.x; k0 < total; k0 += 4 * kSampleBlock) {
    float4 r[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      // Past the end, re-read record k0 (unconditional loads stay in flight
      // together); its result is not used.
      const uint32_t k = k0 + u * kSampleBlock < total ? k0 + u * kSampleBlock : k0;
      uint32_t q = 0, before = 0;
#pragma unroll
      for (int t = 0; t < kXcds - 1; t++)
        if (k >= segEnd[t]) {
          q      = t + 1;
          before = segEnd[t];
        }
      r[u] = rec0[(uint64_t)q * s.capacity + (k - before)];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (k0 + u * kSampleBlock < total)
        binnedPoint<kOp>(g, src, F3{r[u].x, r[u].y, r[u].z}, noff, __float_as_uint(r[u].w), out, outStride);
  }
  // Overflow records (segments that were full): direct path.
  const uint32_t over = *s.overflowCount;
  for (uint64_t k = (uint64_t)bin * kSampleBlock + threadIdx.x; k < over; k += (uint64_t)b.numBins * kSampleBlock) {
    const uint32_t i = s.overflow[k];
    const F3 oc      = {xyz[3 * (uint64_t)i], xyz[3 * (uint64_t)i + 1], xyz[3 * (uint64_t)i + 2]};
    binnedPoint<kOp>(g, src.global, xfmPoint(g.o2i, oc), noff, i, out, outStride);
  }
}

namespace {

template <typename T, int kOp>
hipError_t launchBinnedT(const GridView &g, const AttrView &a, const BinGeometry &b, const float *xyz, float *out,
                         uint32_t outStride, float noff, const BinScratch &s, hipStream_t st)
{
  const size_t lds   = binnedBrickBytes(b, sizeof(T));
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  if (nbz * g.layout.strideZ >= (1ull << 32))
    hipLaunchKernelGGL((kBinnedSample<T, true, kOp>), dim3(b.numBins), dim3(kSampleBlock), lds, st, g, a, b, xyz,
                       out, outStride, noff, s);
  else
    hipLaunchKernelGGL((kBinnedSample<T, false, kOp>), dim3(b.numBins), dim3(kSampleBlock), lds, st, g, a, b, xyz,
                       out, outStride, noff, s);
  return hipGetLastError();
}

template <typename T>
hipError_t launchBinnedOp(const GridView &g, const AttrView &a, const BinGeometry &b, BinOp op, const float *xyz,
                          float *out, uint32_t outStride, float noff, const BinScratch &s, hipStream_t st)
{
  switch (op) {
  case kBinSampleNearest: return launchBinnedT<T, kBinSampleNearest>(g, a, b, xyz, out, outStride, noff, s, st);
  case kBinSampleLinear: return launchBinnedT<T, kBinSampleLinear>(g, a, b, xyz, out, outStride, noff, s, st);
  case kBinSampleCubic: return launchBinnedT<T, kBinSampleCubic>(g, a, b, xyz, out, outStride, noff, s, st);
  case kBinGradientLinear: return launchBinnedT<T, kBinGradientLinear>(g, a, b, xyz, out, outStride, noff, s, st);
  case kBinGradientCubic: return launchBinnedT<T, kBinGradientCubic>(g, a, b, xyz, out, outStride, noff, s, st);
  default: return hipErrorInvalidValue;
  }
}

__global__ void kZeroCounters(uint32_t *__restrict__ seg, uint32_t n, uint32_t *__restrict__ overflowCount)
{
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) seg[i] = 0u;
  if (i == 0) *overflowCount = 0u;
}

void haloOf(BinOp op, uint32_t &lo, uint32_t &hi)
{
  const bool cubic = op == kBinSampleCubic || op == kBinGradientCubic;
  lo               = cubic ? 1 : 0;
  hi               = cubic ? 2 : 1;
}

}  // namespace

size_t binnedBrickBytes(const BinGeometry &b, uint32_t elemSize)
{
  const uint32_t w = b.haloLo + b.haloHi;
  return (size_t)((1u << b.shift[0]) + w) * ((1u << b.shift[1]) + w) * ((1u << b.shift[2]) + w) * elemSize;
}

bool chooseBinGeometry(const uint32_t dims[3], uint32_t elemSize, BinOp op, size_t maxBrickBytes, uint32_t maxBins,
                       BinGeometry &out)
{
  // Candidate regions, largest first; x gets the longest edge (rows are the
  // contiguous direction of the staging loads).
  static const uint32_t kShapes[][3] = {{6, 6, 5}, {6, 5, 5}, {5, 5, 5}, {5, 5, 4}, {5, 4, 4}, {4, 4, 4}, {4, 4, 3}};
  for (const auto &sh : kShapes) {
    BinGeometry b{};
    haloOf(op, b.haloLo, b.haloHi);
    uint64_t bins = 1;
    for (int i = 0; i < 3; i++) {
      b.shift[i] = sh[i];
      b.count[i] = (dims[i] + (1u << sh[i]) - 1) >> sh[i];
      bins *= b.count[i];
    }
    // Rows of the staging loop must fit one workgroup.
    if (binnedBrickBytes(b, elemSize) > maxBrickBytes || (1u << sh[0]) + b.haloLo + b.haloHi > kSampleBlock) continue;
    if (bins > maxBins) return false;  // smaller regions only add bins
    b.numBins = (uint32_t)bins;
    out       = b;
    return true;
  }
  return false;
}

hipError_t launchBinned(const SamplerView &sv, const AttrView &a, const BinGeometry &b, BinOp op, uint64_t n,
                        const float *xyz, float *out, uint32_t outStride, const BinScratch &s, hipStream_t st)
{
  if (n == 0) return hipSuccess;
  // The segment counters and the overflow count start at zero. A fill kernel
  // rather than hipMemsetAsync, so that callers may capture the call in a
  // graph.
  const uint32_t nc = b.numBins * kXcds;
  hipLaunchKernelGGL(kZeroCounters, dim3((nc + 255) / 256), dim3(256), 0, st, s.segCount, nc, s.overflowCount);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t perBlock = (uint64_t)kScatterBlock * kScatterPts;
  const unsigned blocks   = (unsigned)((n + perBlock - 1) / perBlock);
  hipLaunchKernelGGL(kBinScatter, dim3(blocks), dim3(kScatterBlock), b.numBins * sizeof(uint32_t), st, sv.grid, b,
                     n, xyz, s);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const float noff = sv.grid.constantCellData ? 0.f : 0.5f;
  switch (a.type) {
  case kTypeUchar: return launchBinnedOp<uint8_t>(sv.grid, a, b, op, xyz, out, outStride, noff, s, st);
  case kTypeShort: return launchBinnedOp<int16_t>(sv.grid, a, b, op, xyz, out, outStride, noff, s, st);
  case kTypeUshort: return launchBinnedOp<uint16_t>(sv.grid, a, b, op, xyz, out, outStride, noff, s, st);
  case kTypeHalf: return launchBinnedOp<__half>(sv.grid, a, b, op, xyz, out, outStride, noff, s, st);
  case kTypeFloat: return launchBinnedOp<float>(sv.grid, a, b, op, xyz, out, outStride, noff, s, st);
  case kTypeDouble: return launchBinnedOp<double>(sv.grid, a, b, op, xyz, out, outStride, noff, s, st);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace vklhip
