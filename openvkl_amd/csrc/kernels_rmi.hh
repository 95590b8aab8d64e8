// Interval-driven ray marching with transfer-function compositing, fused
// with the interval iterator and the sampler: the reference's ray-march
// iterator renderer (examples/interactive/renderer/ray_march_iterator/
// RayMarchIterator.ispc:39-137), one ray per work-item. Instantiated by
// kernels_rmi_dense.hip (structuredRegular sources) and
// kernels_rmi_sparse.hip (vdb and temporal sources), which compile in
// parallel.
#pragma once

#include "kernels.h"
#include "kernels_hit.hh"

namespace vklhip {

// TransferFunction_sample (examples/interactive/renderer/
// TransferFunction.ih:17-47): tf holds n (r, g, b, opacity) entries over
// [lo, hi]; floor(v) is kept inside [0, n - 1] so a non-finite range cannot
// index out of the table.
VKL_DEV void tfSampleRgba(const float *__restrict__ tf, uint32_t n, float lo, float hi, float v, float c[4])
{
  if (__builtin_isnan(v) || n == 0) {
    c[0] = c[1] = c[2] = c[3] = 0.f;
    return;
  }
  if (v <= lo || v >= hi) {
    const float *e = v <= lo ? tf : tf + 4 * (n - 1);
    c[0] = e[0], c[1] = e[1], c[2] = e[2], c[3] = e[3];
    return;
  }
  v        = (v - lo) / (hi - lo) * ((float)n - 1.f);
  float fi = floorf(v);
  if (!(fi >= 0.f)) fi = 0.f;
  if (fi > (float)(n - 1)) fi = (float)(n - 1);
  const int i     = (int)fi;
  const int j     = i + 1 < (int)n - 1 ? i + 1 : (int)n - 1;
  const float rem = v - fi;
#pragma unroll
  for (int k = 0; k < 4; k++) c[k] = (1.f - rem) * tf[4 * i + k] + rem * tf[4 * j + k];
}

VKL_DEV float minf_(float a, float b) { return a < b ? a : b; }

// Samples of an interval taken together (independent loads in flight): 4
// for vdb sources, whose samples wait on dependent index and leaf loads; 1
// for structuredRegular ones, whose coherent rays hit the caches and which
// lose more to the extra registers than they gain.
template <typename Src>
constexpr int samplesAhead()
{
  return IsDenseSource<Src>::value ? 1 : 4;
}

// rgba[4r..4r+3] = composited colour and accumulated opacity of ray r.
// structuredRegular sources: 5 waves per SIMD (96 VGPRs instead of 126 at 4
// waves; with the dense query alone a 12-byte spill, 128 bytes with both
// queries) measured fastest on a 1080p frame of the 512^3 volume (6 waves:
// 80 VGPRs and a 112-byte spill, 1-7 % slower); vdb sources keep the
// compiler's choice (no spill), which with 4 samples in flight measured 14 %
// faster on the 4096^3 vdb; cubic filters keep the compiler's choice.
template <typename Src, int kFilter, int kDepth, int kQuery>
__global__ __launch_bounds__(kIterBlock) __attribute__((
    amdgpu_waves_per_eu(kFilter == kFilterCubic || !IsDenseSource<Src>::value ? 1 : 5)))
void kRayMarchIntervals(const IteratorContextView *__restrict__ cp,
                                                                  AttrView a, RayMarchIntervalsArgs m)
{
  const uint32_t r = xcdBandBlock(blockIdx.x, gridDim.x) * kIterBlock + threadIdx.x;
  if (r >= m.numRays) return;
  const IteratorContextView &c = *cp;
  PointSampler<Src, kFilter> smp(c.sampler.grid, a, c.attributeIndex);
  smp.setTime(m.time);
  const F3 o = {m.org[3 * r], m.org[3 * r + 1], m.org[3 * r + 2]};
  const F3 d = {m.dir[3 * r], m.dir[3 * r + 1], m.dir[3 * r + 2]};
  IntervalState st;
  intervalInit<kDepth>(c, st, o, d, rayRange(m.tRanges, r));
  float cr = 0.f, cg = 0.f, cb = 0.f, alpha = 0.f;
  uint32_t taken = 0;
  IntervalOut iv;
  constexpr int kAhead = samplesAhead<Src>();
  while (alpha < 0.99f && taken < m.maxSamples && intervalNext<false, kDepth, kQuery>(c, st, iv)) {
    const float dtNominal = iv.nominalDeltaT / m.samplingRate;
    float lo = iv.tLower, hi = minf_(iv.tLower + dtNominal, iv.tUpper);
    if constexpr (kAhead == 1) {
      while (hi - lo > 0.f && alpha < 0.99f && taken < m.maxSamples) {
        const float t  = 0.5f * (lo + hi);
        const float dt = hi - lo;
        const float s  = smp(F3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z});
        taken++;
        float rgba[4];
        tfSampleRgba(m.tf, m.tfSize, m.tfLo, m.tfHi, s, rgba);
        float op = rgba[3] * dt;
        op       = op > 0.f ? op : 0.f;
        op       = op < 1.f ? op : 1.f;
        cr       = cr + (1.f - alpha) * (rgba[0] * op);
        cg       = cg + (1.f - alpha) * (rgba[1] * op);
        cb       = cb + (1.f - alpha) * (rgba[2] * op);
        alpha    = alpha + (1.f - alpha) * op;
        lo       = hi;
        hi       = minf_(lo + dtNominal, iv.tUpper);
      }
    } else {
      while (hi - lo > 0.f && alpha < 0.99f && taken < m.maxSamples) {
        // The next kAhead sub-intervals of this interval (the same running
        // sums as one at a time), their samples issued together so that the
        // loads overlap, then composited in order with the same stopping
        // rules; samples past a stop are not counted.
        float tt[kAhead], dd[kAhead], ss[kAhead];
        uint32_t nk = 0;
#pragma unroll
        for (int k = 0; k < kAhead; k++) {
          const bool live = hi - lo > 0.f;
          tt[k]           = 0.5f * (lo + hi);
          dd[k]           = hi - lo;
          if (live) {
            nk = k + 1;
            lo = hi;
            hi = minf_(lo + dtNominal, iv.tUpper);
          }
        }
#pragma unroll
        for (int k = 0; k < kAhead; k++)
          ss[k] = (uint32_t)k < nk ? smp(F3{o.x + tt[k] * d.x, o.y + tt[k] * d.y, o.z + tt[k] * d.z}) : 0.f;
#pragma unroll
        for (int k = 0; k < kAhead; k++) {
          if ((uint32_t)k >= nk || !(alpha < 0.99f) || taken >= m.maxSamples) continue;
          taken++;
          float rgba[4];
          tfSampleRgba(m.tf, m.tfSize, m.tfLo, m.tfHi, ss[k], rgba);
          float op = rgba[3] * dd[k];
          op       = op > 0.f ? op : 0.f;
          op       = op < 1.f ? op : 1.f;
          cr       = cr + (1.f - alpha) * (rgba[0] * op);
          cg       = cg + (1.f - alpha) * (rgba[1] * op);
          cb       = cb + (1.f - alpha) * (rgba[2] * op);
          alpha    = alpha + (1.f - alpha) * op;
        }
      }
    }
  }
  m.rgba[4 * r]     = cr;
  m.rgba[4 * r + 1] = cg;
  m.rgba[4 * r + 2] = cb;
  m.rgba[4 * r + 3] = alpha;
  if (m.count) m.count[r] = taken;
}

// Launches the instantiation for the context's iterator depth on vdb
// sources (the DDA keeps only the levels it walks) and for its query kind
// (denseQueryOf: structuredRegular volumes within one level-1 node use the
// dense node numbering, and their kernels then carry no node-table query).
struct RayMarchIntervalsLaunch
{
  const IteratorContextView *c;
  AttrView a;
  RayMarchIntervalsArgs m;
  uint32_t depth;
  hipStream_t st;
  bool denseQuery;
  template <typename Src, int kFilter, int kDepth, int kQuery>
  hipError_t run() const
  {
    hipLaunchKernelGGL((kRayMarchIntervals<Src, kFilter, kDepth, kQuery>),
                       dim3((m.numRays + kIterBlock - 1) / kIterBlock), dim3(kIterBlock), 0, st, c, a, m);
    return hipGetLastError();
  }
  template <typename Src, int kFilter>
  hipError_t operator()() const
  {
    // structuredRegular sources keep the depth-generic kernel, which
    // measured faster for them (5.1 against 5.6 ms per 1080p frame).
    if constexpr (IsDenseSource<Src>::value) {
      return denseQuery ? run<Src, kFilter, -1, 1>() : run<Src, kFilter, -1, 0>();
    } else {
      switch (depth) {
      case 0: return run<Src, kFilter, 0, 0>();
      case 1: return run<Src, kFilter, 1, 0>();
      case 2: return run<Src, kFilter, 2, 0>();
      default: return run<Src, kFilter, 3, 0>();
      }
    }
  }
};

}  // namespace vklhip
