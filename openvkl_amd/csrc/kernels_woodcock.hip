// Woodcock (delta) tracking fused with sampling: the free-flight sampler of
// the reference density path tracer (examples/interactive/renderer/
// density_path_tracer/DensityPathTracer.ispc:61-110, sampleWoodcock), one ray
// per work-item, the whole collision loop in registers.
//
// Per tentative collision a ray draws two TEA pairs, advances by an
// exponential step of rate sigmaMax, samples the volume there and accepts the
// collision with probability opacity(sample) (transfer-function w channel).
// The cost per collision is ~16 TEA rounds of integer ALU work plus one
// filtered sample, so the kernel is VALU-bound on a resident volume.

#include "kernels.h"
#include "vkl_device.hh"

namespace vklhip {

constexpr int kWoodBlock = 256;

// 8-round TEA (Wheeler & Needham's Tiny Encryption Algorithm with the
// reference's key words, examples/interactive/renderer/Random.ih:13-27).
VKL_DEV void tea8(uint32_t &v0, uint32_t &v1)
{
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    sum += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
  }
}

// Natural logarithm, binary32 without contraction: x = (1 + f) 2^e with
// 1 + f in [sqrt(1/2), sqrt(2)); with s = f / (2 + f), ln(1 + f) = 2 atanh(s)
// = 2s + s R, R = 2s^2/3 + 2s^4/5 + 2s^6/7 + 2s^8/9 (Abramowitz & Stegun
// 4.1.27; |s| <= 0.172, truncation below 2^-28), evaluated as
// f - (f^2/2 - s (f^2/2 + R)) (using 2s = f - s f) so the rounding of the
// small correction does not reach f; ln 2 is split as 0x3f317180 (low bits
// clear, e times it is exact) + 9.0580013515e-6 (ln 2 minus it). ln 0 = -inf; negative or NaN
// input gives NaN.
VKL_DEV float flightLog(float x)
{
  if (!(x > 0.f)) return x == 0.f ? -__builtin_inff() : __builtin_nanf("");
  if (x == __builtin_inff()) return x;
  int e = 0;
  if (x < 1.17549435e-38f) {
    x = x * 8388608.f;  // 2^23: subnormal -> normal
    e = -23;
  }
  const uint32_t b = __float_as_uint(x);
  e += (int)(b >> 23) - 127;
  float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421354f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f    = m - 1.f;
  const float s    = f / (2.f + f);
  const float z    = s * s;
  float R          = 0.222222222222f;
  R                = R * z + 0.285714285714f;
  R                = R * z + 0.4f;
  R                = R * z + 0.666666666667f;
  R                = R * z;
  const float hfsq = 0.5f * f * f;
  const float k    = (float)e;
  return k * 0.693138123f - ((hfsq - (s * (hfsq + R) + k * 9.0580013515e-6f)) - f);
}

// TransferFunction_sample (examples/interactive/renderer/
// TransferFunction.ih:17-47), w channel only: tf holds 4 floats per entry.
VKL_DEV float tfOpacity(const float *__restrict__ tf, uint32_t n, float lo, float hi, float v)
{
  if (__builtin_isnan(v) || n == 0) return 0.f;
  if (v <= lo) return tf[3];
  if (v >= hi) return tf[4 * (n - 1) + 3];
  v               = (v - lo) / (hi - lo) * ((float)n - 1.f);
  // floor(v) kept inside [0, n - 1], so a non-finite range cannot index out
  // of the table (for finite ranges floor(v) is already inside).
  float fi = floorf(v);
  if (!(fi >= 0.f)) fi = 0.f;
  if (fi > (float)(n - 1)) fi = (float)(n - 1);
  const int i     = (int)fi;
  const float rem = v - fi;
  const int j     = i + 1 < (int)n - 1 ? i + 1 : (int)n - 1;
  return (1.f - rem) * tf[4 * i + 3] + rem * tf[4 * j + 3];
}

// ray r: t starts at tRange[2r]; eventT[r] = t of the accepted collision,
// +inf when t passes tRange[2r + 1] first (transmittance 1), NaN when
// maxCollisions samples were all rejected. rng: (v0, v1) per ray, updated.
template <typename Src, int kFilter>
__global__ __launch_bounds__(kWoodBlock) void kWoodcock(GridView g, AttrView a, uint32_t attr, WoodcockArgs w,
                                                       float noff)
{
  const uint32_t r = xcdBandBlock(blockIdx.x, gridDim.x) * kWoodBlock + threadIdx.x;
  if (r >= w.numRays) return;
  const F3 o = {w.org[3 * r], w.org[3 * r + 1], w.org[3 * r + 2]};
  const F3 d = {w.dir[3 * r], w.dir[3 * r + 1], w.dir[3 * r + 2]};
  float t          = w.tRange[2 * r];
  const float tfar = w.tRange[2 * r + 1];
  uint32_t v0 = w.rng[2 * r], v1 = w.rng[2 * r + 1];
  Src src(g, a, attr);
  float sample = 0.f, tOut = __builtin_inff();
  uint32_t n = 0;
  bool hit   = false;
  if (w.sigmaMax > 0.f) {
    tOut = __builtin_nanf("");
    const float toFloat = 2.3283064365386962890625e-10f;  // 2^-32
    for (; n < w.maxCollisions; n++) {
      tea8(v0, v1);
      const float u0 = (float)v0 * toFloat, u1 = (float)v1 * toFloat;
      tea8(v0, v1);
      const float u2 = (float)v0 * toFloat;
      t = t + -flightLog(1.f - u0) / w.sigmaMax;
      if (t > tfar) {
        tOut = __builtin_inff();
        break;
      }
      float time = w.time;
      if (w.shutter != 0.f) time = time + (u2 - 0.5f) * w.shutter;
      time = time < 0.f ? 0.f : time;
      time = time > 1.f ? 1.f : time;
      src.setTime(time);
      const F3 p = xfmPoint(g.o2i, F3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z});
      if (kFilter == kFilterNearest)
        sample = sampleNearest(src, p, noff);
      else if (kFilter == kFilterCubic)
        sample = sampleCubic(src, p);
      else
        sample = sampleLinear(src, p);
      const float sigmaT = w.sigmaMax * tfOpacity(w.tf, w.tfSize, w.tfLo, w.tfHi, sample);
      if (u1 < sigmaT / w.sigmaMax) {
        tOut = t;
        hit  = true;
        n++;
        break;
      }
    }
  }
  w.eventT[r]  = tOut;
  w.samples[r] = hit ? sample : 0.f;
  w.rng[2 * r]     = v0;
  w.rng[2 * r + 1] = v1;
  if (w.count) w.count[r] = n;
}

namespace {

template <typename Src>
hipError_t woodFilter(const SamplerView &s, const AttrView &a, uint32_t attr, const WoodcockArgs &w, hipStream_t st)
{
  const GridView &g = s.grid;
  const float noff  = g.constantCellData ? 0.f : 0.5f;
  const dim3 grid((w.numRays + kWoodBlock - 1) / kWoodBlock), block(kWoodBlock);
  if (s.filter == kFilterNearest)
    hipLaunchKernelGGL((kWoodcock<Src, kFilterNearest>), grid, block, 0, st, g, a, attr, w, noff);
  else if (s.filter == kFilterCubic)
    hipLaunchKernelGGL((kWoodcock<Src, kFilterCubic>), grid, block, 0, st, g, a, attr, w, noff);
  else
    hipLaunchKernelGGL((kWoodcock<Src, kFilterLinear>), grid, block, 0, st, g, a, attr, w, noff);
  return hipGetLastError();
}

template <typename T>
hipError_t woodDense(const SamplerView &s, const AttrView &a, uint32_t attr, const WoodcockArgs &w, hipStream_t st)
{
  const GridView &g = s.grid;
  if (g.temporalFormat) return woodFilter<TemporalDenseSource<T>>(s, a, attr, w, st);
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  if (nbz * g.layout.strideZ >= (1ull << 32)) return woodFilter<DenseSource<T, true>>(s, a, attr, w, st);
  return woodFilter<DenseSource<T, false>>(s, a, attr, w, st);
}

}  // namespace

hipError_t launchWoodcock(const SamplerView &s, const AttrView &a, uint32_t attr, const WoodcockArgs &w,
                          hipStream_t stream)
{
  if (w.numRays == 0) return hipSuccess;
  if (s.grid.dense) {
    switch (a.type) {
    case kTypeUchar: return woodDense<uint8_t>(s, a, attr, w, stream);
    case kTypeShort: return woodDense<int16_t>(s, a, attr, w, stream);
    case kTypeUshort: return woodDense<uint16_t>(s, a, attr, w, stream);
    case kTypeHalf: return woodDense<__half>(s, a, attr, w, stream);
    case kTypeFloat: return woodDense<float>(s, a, attr, w, stream);
    case kTypeDouble: return woodDense<double>(s, a, attr, w, stream);
    default: return hipErrorInvalidValue;
    }
  }
  if (s.grid.temporalFormat) {
    if (a.type == kTypeHalf) return woodFilter<TemporalSparseSource<__half>>(s, a, attr, w, stream);
    if (a.type == kTypeFloat) return woodFilter<TemporalSparseSource<float>>(s, a, attr, w, stream);
    return hipErrorInvalidValue;
  }
  const bool obs = s.grid.leafAccess != nullptr;
  if (a.type == kTypeHalf)
    return obs ? woodFilter<SparseSource<__half, true>>(s, a, attr, w, stream)
               : woodFilter<SparseSource<__half>>(s, a, attr, w, stream);
  if (a.type == kTypeFloat)
    return obs ? woodFilter<SparseSource<float, true>>(s, a, attr, w, stream)
               : woodFilter<SparseSource<float>>(s, a, attr, w, stream);
  return hipErrorInvalidValue;
}

}  // namespace vklhip
