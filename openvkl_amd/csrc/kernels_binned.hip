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
constexpr int kSampleBlock  = 256;
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
  __syncthreads();

  const uint64_t first = (uint64_t)blockIdx.x * (kScatterBlock * kScatterPts) + threadIdx.x;
  // bin << 14 | rank within the workgroup (both < 2^14: at most 16384 bins
  // and 16384 points per workgroup).
  static_assert(kScatterBlock * kScatterPts <= (1 << 14), "rank field");
  F3 p[kScatterPts];
  uint32_t binRank[kScatterPts];
#pragma unroll
  for (int k = 0; k < kScatterPts; k++) {
    const uint64_t i = first + (uint64_t)k * kScatterBlock;
    binRank[k]       = 0xffffffffu;
    if (i < n) {
      const F3 oc      = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
      p[k]             = xfmPoint(g.o2i, oc);
      const uint32_t c = binOf(b, g.rootOrigin, g.activeSize, p[k]);
      binRank[k]       = (c << 14) | atomicAdd(&hist[c], 1u);
    }
  }
  __syncthreads();
  const uint32_t xcd = blockIdx.x % kXcds;
  for (uint32_t i = threadIdx.x; i < nb; i += kScatterBlock) {
    const uint32_t c = hist[i];
    hist[i]          = c ? atomicAdd(&s.segCount[i * kXcds + xcd], c) : 0u;
  }
  __syncthreads();
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
      s.overflow[atomicAdd(s.overflowCount, 1u)] = (uint32_t)i;
    }
  }
}

// Dense source whose corners come from an LDS brick when the whole 2x2x2
// stencil is inside the domain and the brick; otherwise from the global
// source (same values, same domain test).
template <typename T, bool kIdx64>
struct LdsBrickSource
{
  DenseSource<T, kIdx64> global;
  const T *brick;
  uint32_t bx, by, bz;  // brick origin (root-relative voxels)
  uint32_t ex, ey, ez;  // brick extents (elements)

  VKL_DEV float at(int32_t x, int32_t y, int32_t z) const { return global.at(x, y, z); }

  VKL_DEV void corners8(I3 ic, float *s) const
  {
    const uint32_t ox = (uint32_t)(ic.x - global.rx), oy = (uint32_t)(ic.y - global.ry),
                   oz = (uint32_t)(ic.z - global.rz);
    const uint32_t lx = ox - bx, ly = oy - by, lz = oz - bz;
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

template <typename T, bool kIdx64>
__global__ __launch_bounds__(kSampleBlock) void kBinnedSample(GridView g, AttrView a, BinGeometry b, uint64_t n,
                                                               const float *__restrict__ xyz,
                                                               float *__restrict__ out, uint32_t outStride,
                                                               BinScratch s)
{
  extern __shared__ __attribute__((aligned(16))) unsigned char ldsRaw[];
  T *brick = reinterpret_cast<T *>(ldsRaw);
  const uint32_t bin = blockIdx.x;
  const uint32_t cx = bin % b.count[0], cy = (bin / b.count[0]) % b.count[1], cz = bin / (b.count[0] * b.count[1]);
  LdsBrickSource<T, kIdx64> src{DenseSource<T, kIdx64>(g, a), brick, cx << b.shift[0], cy << b.shift[1],
                                cz << b.shift[2], (1u << b.shift[0]) + 1, (1u << b.shift[1]) + 1,
                                (1u << b.shift[2]) + 1};
  const uint32_t total = src.ex * src.ey * src.ez;
  // Stage the brick; voxels outside the domain are never read from LDS.
  for (uint32_t e = threadIdx.x; e < total; e += kSampleBlock) {
    const uint32_t lx = e % src.ex, ly = (e / src.ex) % src.ey, lz = e / (src.ex * src.ey);
    const uint32_t x = src.bx + lx, y = src.by + ly, z = src.bz + lz;
    T v{};
    if (x < src.global.nx && y < src.global.ny && z < src.global.nz)
      v = src.global.data[src.global.xTerm(x) + src.global.yTerm(y) + src.global.zTerm(z)];
    brick[e] = v;
  }
  __syncthreads();
  for (uint32_t seg = 0; seg < kXcds; seg++) {
    const uint32_t cnt = min(s.segCount[bin * kXcds + seg], s.capacity);
    const float4 *rec  = s.records + ((uint64_t)bin * kXcds + seg) * s.capacity;
    for (uint32_t k = threadIdx.x; k < cnt; k += kSampleBlock) {
      const float4 r = rec[k];
      const uint32_t i = __float_as_uint(r.w);
      out[(uint64_t)i * outStride] = sampleLinear(src, F3{r.x, r.y, r.z});
    }
  }
  // Overflow records (segments that were full): direct path.
  const uint32_t over = *s.overflowCount;
  for (uint64_t k = (uint64_t)bin * kSampleBlock + threadIdx.x; k < over; k += (uint64_t)b.numBins * kSampleBlock) {
    const uint32_t i = s.overflow[k];
    const F3 oc      = {xyz[3 * (uint64_t)i], xyz[3 * (uint64_t)i + 1], xyz[3 * (uint64_t)i + 2]};
    out[(uint64_t)i * outStride] = sampleLinear(src.global, xfmPoint(g.o2i, oc));
  }
}

namespace {

template <typename T>
hipError_t launchBinnedSampleT(const GridView &g, const AttrView &a, const BinGeometry &b, uint64_t n,
                               const float *xyz, float *out, uint32_t outStride, const BinScratch &s,
                               hipStream_t st)
{
  const size_t lds   = binnedBrickBytes(b, sizeof(T));
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  if (nbz * g.layout.strideZ >= (1ull << 32))
    hipLaunchKernelGGL((kBinnedSample<T, true>), dim3(b.numBins), dim3(kSampleBlock), lds, st, g, a, b, n, xyz, out,
                       outStride, s);
  else
    hipLaunchKernelGGL((kBinnedSample<T, false>), dim3(b.numBins), dim3(kSampleBlock), lds, st, g, a, b, n, xyz, out,
                       outStride, s);
  return hipGetLastError();
}

}  // namespace

size_t binnedBrickBytes(const BinGeometry &b, uint32_t elemSize)
{
  return (size_t)((1u << b.shift[0]) + 1) * ((1u << b.shift[1]) + 1) * ((1u << b.shift[2]) + 1) * elemSize;
}

bool chooseBinGeometry(const uint32_t dims[3], uint32_t elemSize, size_t maxBrickBytes, uint32_t maxBins,
                       BinGeometry &out)
{
  // Candidate bricks, largest first; x gets the longest edge (rows are the
  // contiguous direction of the staging loads).
  static const uint32_t kShapes[][3] = {{6, 6, 5}, {6, 5, 5}, {5, 5, 5}, {5, 5, 4}, {5, 4, 4}, {4, 4, 4}, {4, 4, 3}};
  for (const auto &sh : kShapes) {
    BinGeometry b{};
    uint64_t bins = 1;
    for (int i = 0; i < 3; i++) {
      b.shift[i] = sh[i];
      b.count[i] = (dims[i] + (1u << sh[i]) - 1) >> sh[i];
      bins *= b.count[i];
    }
    if (binnedBrickBytes(b, elemSize) > maxBrickBytes) continue;
    if (bins > maxBins) return false;  // smaller bricks only add bins
    b.numBins = (uint32_t)bins;
    out       = b;
    return true;
  }
  return false;
}

hipError_t launchSampleBinned(const SamplerView &sv, const AttrView &a, const BinGeometry &b, uint64_t n,
                              const float *xyz, float *out, uint32_t outStride, const BinScratch &s,
                              hipStream_t st)
{
  if (n == 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(s.segCount, 0, (size_t)b.numBins * kXcds * sizeof(uint32_t), st);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(s.overflowCount, 0, sizeof(uint32_t), st);
  if (e != hipSuccess) return e;
  const uint64_t perBlock = (uint64_t)kScatterBlock * kScatterPts;
  const unsigned blocks   = (unsigned)((n + perBlock - 1) / perBlock);
  hipLaunchKernelGGL(kBinScatter, dim3(blocks), dim3(kScatterBlock), b.numBins * sizeof(uint32_t), st, sv.grid, b,
                     n, xyz, s);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  switch (a.type) {
  case kTypeUchar: return launchBinnedSampleT<uint8_t>(sv.grid, a, b, n, xyz, out, outStride, s, st);
  case kTypeShort: return launchBinnedSampleT<int16_t>(sv.grid, a, b, n, xyz, out, outStride, s, st);
  case kTypeUshort: return launchBinnedSampleT<uint16_t>(sv.grid, a, b, n, xyz, out, outStride, s, st);
  case kTypeHalf: return launchBinnedSampleT<__half>(sv.grid, a, b, n, xyz, out, outStride, s, st);
  case kTypeFloat: return launchBinnedSampleT<float>(sv.grid, a, b, n, xyz, out, outStride, s, st);
  case kTypeDouble: return launchBinnedSampleT<double>(sv.grid, a, b, n, xyz, out, outStride, s, st);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace vklhip
