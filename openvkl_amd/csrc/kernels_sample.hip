// Stream sampling kernels: vklComputeSampleN / GradientN / SampleMN on the
// structuredRegular (dense) and vdb (sparse) volumes.
//
// One work-item per point, 256-wide workgroups (4 wavefronts). Points are
// read as packed vec3f (12 B, coalesced per wavefront), samples written as
// float (or vec3f for gradients). Each kernel instance is specialised on the
// voxel element type, the filter and 32- vs 64-bit voxel addressing.

#include "kernels.h"
#include "vkl_device.hh"

namespace vklhip {

constexpr int kBlock = 256;

template <typename Src, int kFilter>
__global__ __launch_bounds__(kBlock) void kSample(GridView g, AttrView a, uint32_t attr, uint64_t n,
                                                   const float *__restrict__ xyz, float *__restrict__ out, float noff,
                                                   uint32_t outStride, const float *__restrict__ times)
{
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  // Points and samples stream through once: non-temporal accesses keep them
  // from displacing voxel lines in the caches.
  const F3 oc = {__builtin_nontemporal_load(xyz + 3 * i), __builtin_nontemporal_load(xyz + 3 * i + 1),
                 __builtin_nontemporal_load(xyz + 3 * i + 2)};
  const F3 p  = xfmPoint(g.o2i, oc);
  Src src(g, a, attr);
  if (times) src.setTime(times[i]);
  float r;
  if (kFilter == kFilterNearest)
    r = sampleNearest(src, p, noff);
  else if (kFilter == kFilterCubic)
    r = sampleCubicStream(src, p);
  else
    r = sampleLinear(src, p);
  __builtin_nontemporal_store(r, out + i * outStride);
}

// VdbSampler.ispc:258-335: index-space gradient, then xfmNormal with the
// object->index matrix (the transposed inverse of index->object).
template <typename Src, int kFilter>
__global__ __launch_bounds__(kBlock) void kGradient(GridView g, AttrView a, uint32_t attr, uint64_t n,
                                                     const float *__restrict__ xyz, float *__restrict__ out,
                                                     const float *__restrict__ times)
{
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  F3 r = {0.f, 0.f, 0.f};
  if (kFilter != kFilterNearest) {
    const F3 oc = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    const F3 p  = xfmPoint(g.o2i, oc);
    Src src(g, a, attr);
    if (times) src.setTime(times[i]);
    const F3 gi = (kFilter == kFilterCubic) ? gradientCubicStream(src, p) : gradientLinear(src, p);
    r           = xfmNormal(g.o2i, gi);
  }
  out[3 * i]     = r.x;
  out[3 * i + 1] = r.y;
  out[3 * i + 2] = r.z;
}

// Fixed-step ray marching fused with sampling (one ray per work-item):
// out[r] = sum over s of sample(o + t_s d), t_s = t0 + (s + 0.5) dt,
// dt = (t1 - t0) / steps, summed in step order. Adjacent rays of a
// wavefront read adjacent voxels, so the gathers hit L1 / L2.
template <typename Src, int kFilter>
__global__ __launch_bounds__(kBlock) void kRayMarch(GridView g, AttrView a, uint32_t attr, uint32_t numRays,
                                                     const float *__restrict__ org, const float *__restrict__ dir,
                                                     const float *__restrict__ tRange, uint32_t steps, float noff,
                                                     float *__restrict__ out)
{
  const uint32_t r = xcdBandBlock(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (r >= numRays) return;
  const F3 o = {org[3 * r], org[3 * r + 1], org[3 * r + 2]};
  const F3 d = {dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]};
  const float t0 = tRange[2 * r], dt = (tRange[2 * r + 1] - t0) / (float)steps;
  const Src src(g, a, attr);
  float acc = 0.f;
  for (uint32_t k = 0; k < steps; k++) {
    const float t = t0 + ((float)k + 0.5f) * dt;
    const F3 p    = xfmPoint(g.o2i, F3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z});
    float v;
    if (kFilter == kFilterNearest)
      v = sampleNearest(src, p, noff);
    else if (kFilter == kFilterCubic)
      v = sampleCubic(src, p);
    else
      v = sampleLinear(src, p);
    acc += v;
  }
  out[r] = acc;
}

namespace {

inline unsigned blocksFor(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct MarchArgs
{
  uint32_t numRays, steps;
  const float *org, *dir, *tRange;
  float *out;
};

template <typename Src>
hipError_t marchFilter(const SamplerView &s, const AttrView &a, uint32_t attr, const MarchArgs &m, hipStream_t st)
{
  const GridView &g = s.grid;
  const float noff  = g.constantCellData ? 0.f : 0.5f;
  const dim3 grid(blocksFor(m.numRays)), block(kBlock);
  if (s.filter == kFilterNearest)
    hipLaunchKernelGGL((kRayMarch<Src, kFilterNearest>), grid, block, 0, st, g, a, attr, m.numRays, m.org, m.dir,
                       m.tRange, m.steps, noff, m.out);
  else if (s.filter == kFilterCubic)
    hipLaunchKernelGGL((kRayMarch<Src, kFilterCubic>), grid, block, 0, st, g, a, attr, m.numRays, m.org, m.dir,
                       m.tRange, m.steps, noff, m.out);
  else
    hipLaunchKernelGGL((kRayMarch<Src, kFilterLinear>), grid, block, 0, st, g, a, attr, m.numRays, m.org, m.dir,
                       m.tRange, m.steps, noff, m.out);
  return hipGetLastError();
}

template <typename T>
hipError_t marchDense(const SamplerView &s, const AttrView &a, uint32_t attr, const MarchArgs &m, hipStream_t st)
{
  const GridView &g = s.grid;
  if (g.temporalFormat) return marchFilter<TemporalDenseSource<T>>(s, a, attr, m, st);
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  if (nbz * g.layout.strideZ >= (1ull << 32)) return marchFilter<DenseSource<T, true>>(s, a, attr, m, st);
  return marchFilter<DenseSource<T, false>>(s, a, attr, m, st);
}

template <typename Src>
hipError_t dispatchFilter(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                          float *out, uint32_t stride, const float *times, hipStream_t st, bool gradient)
{
  const GridView &g = s.grid;
  const float noff  = g.constantCellData ? 0.f : 0.5f;
  const int f       = gradient ? s.gradientFilter : s.filter;
  const dim3 grid(blocksFor(n)), block(kBlock);
  if (gradient) {
    if (f == kFilterNearest)
      hipLaunchKernelGGL((kGradient<Src, kFilterNearest>), grid, block, 0, st, g, a, attr, n, xyz, out, times);
    else if (f == kFilterCubic)
      hipLaunchKernelGGL((kGradient<Src, kFilterCubic>), grid, block, 0, st, g, a, attr, n, xyz, out, times);
    else
      hipLaunchKernelGGL((kGradient<Src, kFilterLinear>), grid, block, 0, st, g, a, attr, n, xyz, out, times);
  } else {
    if (f == kFilterNearest)
      hipLaunchKernelGGL((kSample<Src, kFilterNearest>), grid, block, 0, st, g, a, attr, n, xyz, out, noff,
                         stride, times);
    else if (f == kFilterCubic)
      hipLaunchKernelGGL((kSample<Src, kFilterCubic>), grid, block, 0, st, g, a, attr, n, xyz, out, noff,
                         stride, times);
    else
      hipLaunchKernelGGL((kSample<Src, kFilterLinear>), grid, block, 0, st, g, a, attr, n, xyz, out, noff,
                         stride, times);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t dispatchDense(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                         float *out, uint32_t stride, const float *times, hipStream_t st, bool gradient)
{
  const GridView &g = s.grid;
  if (g.temporalFormat)
    return dispatchFilter<TemporalDenseSource<T>>(s, a, attr, n, xyz, out, stride, times, st, gradient);
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  if (nbz * g.layout.strideZ >= (1ull << 32))
    return dispatchFilter<DenseSource<T, true>>(s, a, attr, n, xyz, out, stride, times, st, gradient);
  return dispatchFilter<DenseSource<T, false>>(s, a, attr, n, xyz, out, stride, times, st, gradient);
}

hipError_t dispatch(const SamplerView &s, uint32_t attr, const AttrView &a, uint64_t n, const float *xyz,
                    float *out, uint32_t stride, const float *times, hipStream_t st, bool gradient)
{
  // Tricubic on volumes with cubic records: wavefront-cooperative stencil
  // loads (kernels_cubic.hip), the same samples.
  if (s.grid.dense && !s.grid.temporalFormat && a.cubic && (a.cubicStencils || s.grid.activeSize[0] >= 4) &&
      (gradient ? s.gradientFilter : s.filter) == kFilterCubic)
    return launchCubicRecords(s, a, attr, n, xyz, out, stride, gradient, st);
  if (s.grid.dense) {
    switch (a.type) {
    case kTypeUchar: return dispatchDense<uint8_t>(s, a, attr, n, xyz, out, stride, times, st, gradient);
    case kTypeShort: return dispatchDense<int16_t>(s, a, attr, n, xyz, out, stride, times, st, gradient);
    case kTypeUshort: return dispatchDense<uint16_t>(s, a, attr, n, xyz, out, stride, times, st, gradient);
    case kTypeHalf: return dispatchDense<__half>(s, a, attr, n, xyz, out, stride, times, st, gradient);
    case kTypeFloat: return dispatchDense<float>(s, a, attr, n, xyz, out, stride, times, st, gradient);
    case kTypeDouble: return dispatchDense<double>(s, a, attr, n, xyz, out, stride, times, st, gradient);
    default: return hipErrorInvalidValue;
    }
  }
  if (s.grid.temporalFormat) {
    if (a.type == kTypeHalf)
      return dispatchFilter<TemporalSparseSource<__half>>(s, a, attr, n, xyz, out, stride, times, st, gradient);
    if (a.type == kTypeFloat)
      return dispatchFilter<TemporalSparseSource<float>>(s, a, attr, n, xyz, out, stride, times, st, gradient);
    return hipErrorInvalidValue;
  }
  const bool obs = s.grid.leafAccess != nullptr;
  if (a.type == kTypeHalf)
    return obs ? dispatchFilter<SparseSource<__half, true>>(s, a, attr, n, xyz, out, stride, times, st, gradient)
               : dispatchFilter<SparseSource<__half>>(s, a, attr, n, xyz, out, stride, times, st, gradient);
  if (a.type == kTypeFloat)
    return obs ? dispatchFilter<SparseSource<float, true>>(s, a, attr, n, xyz, out, stride, times, st, gradient)
               : dispatchFilter<SparseSource<float>>(s, a, attr, n, xyz, out, stride, times, st, gradient);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launchSample(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                        float *out, hipStream_t stream, const float *times)
{
  if (n == 0) return hipSuccess;
  return dispatch(s, attr, a, n, xyz, out, 1, times, stream, false);
}

hipError_t launchGradient(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                          float *out, hipStream_t stream, const float *times)
{
  if (n == 0) return hipSuccess;
  return dispatch(s, attr, a, n, xyz, out, 1, times, stream, true);
}

hipError_t launchRayMarch(const SamplerView &s, const AttrView &a, uint32_t attr, uint32_t numRays, const float *org,
                          const float *dir, const float *tRange, uint32_t steps, float *out, hipStream_t stream)
{
  if (numRays == 0) return hipSuccess;
  const MarchArgs m{numRays, steps, org, dir, tRange, out};
  if (s.grid.dense) {
    switch (a.type) {
    case kTypeUchar: return marchDense<uint8_t>(s, a, attr, m, stream);
    case kTypeShort: return marchDense<int16_t>(s, a, attr, m, stream);
    case kTypeUshort: return marchDense<uint16_t>(s, a, attr, m, stream);
    case kTypeHalf: return marchDense<__half>(s, a, attr, m, stream);
    case kTypeFloat: return marchDense<float>(s, a, attr, m, stream);
    case kTypeDouble: return marchDense<double>(s, a, attr, m, stream);
    default: return hipErrorInvalidValue;
    }
  }
  if (s.grid.temporalFormat) {
    if (a.type == kTypeHalf) return marchFilter<TemporalSparseSource<__half>>(s, a, attr, m, stream);
    if (a.type == kTypeFloat) return marchFilter<TemporalSparseSource<float>>(s, a, attr, m, stream);
    return hipErrorInvalidValue;
  }
  const bool obs = s.grid.leafAccess != nullptr;
  if (a.type == kTypeHalf)
    return obs ? marchFilter<SparseSource<__half, true>>(s, a, attr, m, stream)
               : marchFilter<SparseSource<__half>>(s, a, attr, m, stream);
  if (a.type == kTypeFloat)
    return obs ? marchFilter<SparseSource<float, true>>(s, a, attr, m, stream)
               : marchFilter<SparseSource<float>>(s, a, attr, m, stream);
  return hipErrorInvalidValue;
}

hipError_t launchSampleStrided(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                               float *out, uint32_t outStride, hipStream_t stream, const float *times)
{
  if (n == 0) return hipSuccess;
  return dispatch(s, attr, a, n, xyz, out, outStride, times, stream, false);
}

// Re-lay an x-fastest array into the brick layout L (one work-item per
// destination element; padding elements are zero and never read).
template <typename E>
__global__ __launch_bounds__(kBlock) void kSwizzle(const E *__restrict__ src, E *__restrict__ dst, uint32_t nx,
                                                    uint32_t ny, uint32_t nz, DenseLayout L, uint64_t nbx,
                                                    uint64_t nby, uint64_t total)
{
  const uint64_t d = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (d >= total) return;
  const uint32_t S     = L.sx + L.sy + L.sz;
  const uint64_t b     = d >> S;
  const uint32_t inner = (uint32_t)(d & ((1ull << S) - 1));
  const uint64_t bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
  const uint64_t x = (bx << L.sx) | (inner & ((1u << L.sx) - 1));
  const uint64_t y = (by << L.sy) | ((inner >> L.sx) & ((1u << L.sy) - 1));
  const uint64_t z = (bz << L.sz) | (inner >> (L.sx + L.sy));
  E v{};
  if (x < nx && y < ny && z < nz) v = src[(z * ny + y) * (uint64_t)nx + x];
  dst[d] = v;
}

// Corner-packed layout: cell (x, y, z) <- the 8 voxels v(x + dx, y + dy,
// z + dz), record order 4 dx + 2 dy + dz, coordinates past the last voxel
// clamped to it (those corners are never read through the record: the
// stencil is then out of the domain). One work-item per cell.
template <typename E>
__global__ __launch_bounds__(kBlock) void kPackCorners(const E *__restrict__ src, E *__restrict__ dst, uint32_t nx,
                                                        uint32_t ny, uint32_t nz)
{
  const uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= (uint64_t)nx * ny * nz) return;
  const uint32_t x = (uint32_t)(c % nx), y = (uint32_t)((c / nx) % ny), z = (uint32_t)(c / ((uint64_t)nx * ny));
  const uint32_t x1 = x + 1 < nx ? x + 1 : x, y1 = y + 1 < ny ? y + 1 : y, z1 = z + 1 < nz ? z + 1 : z;
  CellRecord<E> r;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t xx = (k & 4) ? x1 : x, yy = (k & 2) ? y1 : y, zz = (k & 1) ? z1 : z;
    r.v[k] = src[((uint64_t)zz * ny + yy) * nx + xx];
  }
  reinterpret_cast<CellRecord<E> *>(dst)[c] = r;
}

hipError_t launchPackCorners(const void *src, void *dst, uint32_t elemSize, const uint32_t dims[3],
                             hipStream_t stream)
{
  const uint64_t cells = (uint64_t)dims[0] * dims[1] * dims[2];
  const dim3 grid(blocksFor(cells)), block(kBlock);
  switch (elemSize) {
  case 1:
    hipLaunchKernelGGL(kPackCorners<uint8_t>, grid, block, 0, stream, (const uint8_t *)src, (uint8_t *)dst, dims[0],
                       dims[1], dims[2]);
    break;
  case 2:
    hipLaunchKernelGGL(kPackCorners<uint16_t>, grid, block, 0, stream, (const uint16_t *)src, (uint16_t *)dst,
                       dims[0], dims[1], dims[2]);
    break;
  case 4:
    hipLaunchKernelGGL(kPackCorners<uint32_t>, grid, block, 0, stream, (const uint32_t *)src, (uint32_t *)dst,
                       dims[0], dims[1], dims[2]);
    break;
  case 8:
    hipLaunchKernelGGL(kPackCorners<uint64_t>, grid, block, 0, stream, (const uint64_t *)src, (uint64_t *)dst,
                       dims[0], dims[1], dims[2]);
    break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launchSwizzle(const void *src, void *dst, uint32_t elemSize, const uint32_t dims[3],
                         const DenseLayout &L, uint64_t total, hipStream_t stream)
{
  const uint64_t nbx = ((uint64_t)dims[0] + (1u << L.sx) - 1) >> L.sx;
  const uint64_t nby = ((uint64_t)dims[1] + (1u << L.sy) - 1) >> L.sy;
  const dim3 grid(blocksFor(total)), block(kBlock);
  switch (elemSize) {
  case 1:
    hipLaunchKernelGGL(kSwizzle<uint8_t>, grid, block, 0, stream, (const uint8_t *)src, (uint8_t *)dst, dims[0],
                       dims[1], dims[2], L, nbx, nby, total);
    break;
  case 2:
    hipLaunchKernelGGL(kSwizzle<uint16_t>, grid, block, 0, stream, (const uint16_t *)src, (uint16_t *)dst, dims[0],
                       dims[1], dims[2], L, nbx, nby, total);
    break;
  case 4:
    hipLaunchKernelGGL(kSwizzle<uint32_t>, grid, block, 0, stream, (const uint32_t *)src, (uint32_t *)dst, dims[0],
                       dims[1], dims[2], L, nbx, nby, total);
    break;
  case 8:
    hipLaunchKernelGGL(kSwizzle<uint64_t>, grid, block, 0, stream, (const uint64_t *)src, (uint64_t *)dst, dims[0],
                       dims[1], dims[2], L, nbx, nby, total);
    break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace vklhip
