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
                                                   const float *__restrict__ xyz, float *__restrict__ out, float noff)
{
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const F3 oc = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  const F3 p  = xfmPoint(g.o2i, oc);
  const Src src(g, a, attr);
  float r;
  if (kFilter == kFilterNearest)
    r = sampleNearest(src, p, noff);
  else if (kFilter == kFilterCubic)
    r = sampleCubic(src, p);
  else
    r = sampleLinear(src, p);
  out[i] = r;
}

// VdbSampler.ispc:258-335: index-space gradient, then xfmNormal with the
// object->index matrix (the transposed inverse of index->object).
template <typename Src, int kFilter>
__global__ __launch_bounds__(kBlock) void kGradient(GridView g, AttrView a, uint32_t attr, uint64_t n,
                                                     const float *__restrict__ xyz, float *__restrict__ out)
{
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  F3 r = {0.f, 0.f, 0.f};
  if (kFilter != kFilterNearest) {
    const F3 oc = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    const F3 p  = xfmPoint(g.o2i, oc);
    const Src src(g, a, attr);
    const F3 gi = (kFilter == kFilterCubic) ? gradientCubic(src, p) : gradientLinear(src, p);
    r           = xfmNormal(g.o2i, gi);
  }
  out[3 * i]     = r.x;
  out[3 * i + 1] = r.y;
  out[3 * i + 2] = r.z;
}

// Sampler.h:226-244 layout samples[i * M + a]; attributes may differ in type,
// so this path dispatches per attribute at run time.
__global__ __launch_bounds__(kBlock) void kSampleM(SamplerView s, uint32_t m, const uint32_t *__restrict__ attrs,
                                                    uint64_t n, const float *__restrict__ xyz,
                                                    float *__restrict__ out)
{
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const F3 oc = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  for (uint32_t k = 0; k < m; k++) out[i * m + k] = sampleObject(s, attrs[k], oc);
}

namespace {

inline unsigned blocksFor(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <typename Src>
hipError_t dispatchFilter(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                          float *out, hipStream_t st, bool gradient)
{
  const GridView &g = s.grid;
  const float noff  = g.constantCellData ? 0.f : 0.5f;
  const int f       = gradient ? s.gradientFilter : s.filter;
  const dim3 grid(blocksFor(n)), block(kBlock);
  if (gradient) {
    if (f == kFilterNearest)
      hipLaunchKernelGGL((kGradient<Src, kFilterNearest>), grid, block, 0, st, g, a, attr, n, xyz, out);
    else if (f == kFilterCubic)
      hipLaunchKernelGGL((kGradient<Src, kFilterCubic>), grid, block, 0, st, g, a, attr, n, xyz, out);
    else
      hipLaunchKernelGGL((kGradient<Src, kFilterLinear>), grid, block, 0, st, g, a, attr, n, xyz, out);
  } else {
    if (f == kFilterNearest)
      hipLaunchKernelGGL((kSample<Src, kFilterNearest>), grid, block, 0, st, g, a, attr, n, xyz, out, noff);
    else if (f == kFilterCubic)
      hipLaunchKernelGGL((kSample<Src, kFilterCubic>), grid, block, 0, st, g, a, attr, n, xyz, out, noff);
    else
      hipLaunchKernelGGL((kSample<Src, kFilterLinear>), grid, block, 0, st, g, a, attr, n, xyz, out, noff);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t dispatchDense(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                         float *out, hipStream_t st, bool gradient)
{
  const GridView &g = s.grid;
  const uint64_t nvox = (uint64_t)g.activeSize[0] * g.activeSize[1] * g.activeSize[2];
  if (nvox >= (1ull << 32)) return dispatchFilter<DenseSource<T, true>>(s, a, attr, n, xyz, out, st, gradient);
  return dispatchFilter<DenseSource<T, false>>(s, a, attr, n, xyz, out, st, gradient);
}

hipError_t dispatch(const SamplerView &s, uint32_t attr, const AttrView &a, uint64_t n, const float *xyz,
                    float *out, hipStream_t st, bool gradient)
{
  if (s.grid.dense) {
    switch (a.type) {
    case kTypeUchar: return dispatchDense<uint8_t>(s, a, attr, n, xyz, out, st, gradient);
    case kTypeShort: return dispatchDense<int16_t>(s, a, attr, n, xyz, out, st, gradient);
    case kTypeUshort: return dispatchDense<uint16_t>(s, a, attr, n, xyz, out, st, gradient);
    case kTypeHalf: return dispatchDense<__half>(s, a, attr, n, xyz, out, st, gradient);
    case kTypeFloat: return dispatchDense<float>(s, a, attr, n, xyz, out, st, gradient);
    case kTypeDouble: return dispatchDense<double>(s, a, attr, n, xyz, out, st, gradient);
    default: return hipErrorInvalidValue;
    }
  }
  if (a.type == kTypeHalf) return dispatchFilter<SparseSource<__half>>(s, a, attr, n, xyz, out, st, gradient);
  if (a.type == kTypeFloat) return dispatchFilter<SparseSource<float>>(s, a, attr, n, xyz, out, st, gradient);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launchSample(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                        float *out, hipStream_t stream)
{
  if (n == 0) return hipSuccess;
  return dispatch(s, attr, a, n, xyz, out, stream, false);
}

hipError_t launchGradient(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                          float *out, hipStream_t stream)
{
  if (n == 0) return hipSuccess;
  return dispatch(s, attr, a, n, xyz, out, stream, true);
}

hipError_t launchSampleM(const SamplerView &s, uint32_t m, const uint32_t *attrsDev, uint64_t n, const float *xyz,
                         float *out, hipStream_t stream)
{
  if (n == 0 || m == 0) return hipSuccess;
  hipLaunchKernelGGL(kSampleM, dim3(blocksFor(n)), dim3(kBlock), 0, stream, s, m, attrsDev, n, xyz, out);
  return hipGetLastError();
}

}  // namespace vklhip
