// The hit-iterator renderer fused with hit iteration, gradients and shadow
// queries: examples/interactive/renderer/hit_iterator_renderer/
// HitIteratorRenderer.ispc:51-150, one ray per work-item. Instantiated by
// kernels_hitrender_dense.hip (structuredRegular sources) and
// kernels_hitrender_sparse.hip (vdb and temporal sources).
//
// The primary ray and its shadow rays share one hit-iterator state and one
// hitNext call site (the primary state is parked while a hit's shadow rays
// run, and only when the hit casts one),
// so the large hit-iteration code is inlined once per specialisation.
#pragma once

#include "kernels.h"
#include "kernels_hit.hh"
#include "kernels_rmi.hh"

namespace vklhip {

// normalize(v) = v (1 / sqrt(v . v)), correctly rounded square root and
// division.
VKL_DEV F3 normalize3(F3 v)
{
  const float inv = 1.f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return F3{v.x * inv, v.y * inv, v.z * inv};
}
VKL_DEV float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename Src, int kFilter>
__global__ __launch_bounds__(kIterBlock) void kHitRender(const IteratorContextView *__restrict__ cp, AttrView a,
                                                          HitRenderArgs m)
{
  const uint32_t r = xcdBandBlock(blockIdx.x, gridDim.x) * kIterBlock + threadIdx.x;
  if (r >= m.numRays) return;
  const IteratorContextView &c = *cp;
  PointSampler<Src, kFilter> smp(c.sampler.grid, a, c.attributeIndex);
  smp.setTime(m.time);
  const F3 o = {m.org[3 * r], m.org[3 * r + 1], m.org[3 * r + 2]};
  const F3 d = {m.dir[3 * r], m.dir[3 * r + 1], m.dir[3 * r + 2]};
  const F3 light0 = normalize3(F3{1.f, 1.f, 1.f}), light1 = normalize3(F3{1.f, 1.f, -1.f});
  const F3 wi     = normalize3(F3{-1.f * d.x, -1.f * d.y, -1.f * d.z});
  HitState st, parked;
  hitInit(c, st, o, d, rayRange(m.tRanges, r), m.time);
  float cr = 0.f, cg = 0.f, cb = 0.f, alpha = 0.f;
  uint32_t hits = 0;
  bool shadowRay = false, isParked = false;
  int light      = 0;  // next light of the hit being shaded
  F3 p = {0.f, 0.f, 0.f}, N = {0.f, 0.f, 0.f};
  float ci = 0.f, co = 0.f, illum = 0.f, hitSample = 0.f, hitEps = 0.f;
  for (;;) {
    if (!shadowRay && !(alpha < 0.99f && hits < m.maxHits)) break;
    HitOut h;
    const bool got = hitNext(c, smp, st, h);
    if (!shadowRay) {
      if (!got) break;
      hits++;
      p = F3{o.x + h.t * d.x, o.y + h.t * d.y, o.z + h.t * d.z};
      const F3 q = xfmPoint(c.sampler.grid.o2i, p);
      F3 gi      = {0.f, 0.f, 0.f};
      if (c.sampler.gradientFilter == kFilterLinear)
        gi = gradientLinear(smp.src, q);
      else if (c.sampler.gradientFilter == kFilterCubic)
        gi = gradientCubic(smp.src, q);
      N = normalize3(xfmNormal(c.sampler.grid.o2i, gi));
      if (isnan(N.x) || isnan(N.y) || isnan(N.z)) N = F3{0.f, 0.f, 0.f};
      ci        = dot3(N, wi);
      hitSample = h.sample;
      hitEps    = h.epsilon;
      if (sqrtf(dot3(N, N)) > 0.f) {
        illum = 0.1f;  // ambient
        light = 0;
      } else {
        illum = 1.f;  // no normal: unshaded
        light = 2;
      }
    } else {
      if (!got) illum += fabsf(co) * (light == 0 ? 1.f : 0.8f);
      light++;
      shadowRay = false;
    }
    // The next light this hit is shaded by, if it lies on the viewer's side.
    while (light < 2) {
      co = dot3(N, light == 0 ? light0 : light1);
      if ((co > 0.f) == (ci > 0.f)) break;
      light++;
    }
    if (light < 2) {
      if (!isParked) {  // first shadow ray of this hit: park the primary ray
        parked   = st;
        isParked = true;
      }
      hitInit(c, st, p, light == 0 ? light0 : light1, Range1f{hitEps, vinf()}, m.time);
      shadowRay = true;
      continue;
    }
    float albedo[4];
    tfSampleRgba(m.tf, m.tfSize, m.tfLo, m.tfHi, hitSample, albedo);
    cr    = cr + (1.f - alpha) * illum * albedo[0];
    cg    = cg + (1.f - alpha) * illum * albedo[1];
    cb    = cb + (1.f - alpha) * illum * albedo[2];
    alpha = alpha + (1.f - alpha) * 0.6f;
    if (isParked) {
      st       = parked;
      isParked = false;
    }
  }
  m.rgba[4 * r]     = cr;
  m.rgba[4 * r + 1] = cg;
  m.rgba[4 * r + 2] = cb;
  m.rgba[4 * r + 3] = alpha;
  if (m.count) m.count[r] = hits;
}

struct HitRenderLaunch
{
  const IteratorContextView *c;
  AttrView a;
  HitRenderArgs m;
  hipStream_t st;
  template <typename Src, int kFilter>
  hipError_t operator()() const
  {
    hipLaunchKernelGGL((kHitRender<Src, kFilter>), dim3((m.numRays + kIterBlock - 1) / kIterBlock),
                       dim3(kIterBlock), 0, st, c, a, m);
    return hipGetLastError();
  }
};

}  // namespace vklhip
