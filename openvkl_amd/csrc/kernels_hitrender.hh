// The hit-iterator renderer fused with hit iteration, gradients and shadow
// queries: examples/interactive/renderer/hit_iterator_renderer/
// HitIteratorRenderer.ispc:51-150, one ray per work-item. Instantiated by
// kernels_hitrender_dense.hip (structuredRegular sources) and
// kernels_hitrender_sparse.hip (vdb and temporal sources).
//
// The primary ray's hits are collected first and shaded afterwards, so the
// primary and shadow rays never need two hit-iterator states at once.
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

// Opacity 0.6 per shaded hit takes the accumulated opacity past 0.99 after
// 6 hits (binary32: 0.6, 0.84, 0.936, 0.9744, 0.98976, 0.995904), so a ray
// shades at most this many hits.
constexpr int kMaxShadedHits = 8;

// Two phases per ray, in the reference's order of results: first the
// primary ray's hits that will be shaded (the compositing opacity does not
// depend on the shading, so their number is known in advance; they wait in
// LDS), then each hit in order: gradient normal, shadow rays, compositing.
// Primary and shadow queries share one hit-iterator state and one hitNext
// call site (the hit iteration is inlined once), and no second state is
// parked, so the state stays in registers.
// Occupancy: on structuredRegular sources 5 waves per SIMD (128 VGPRs at
// most, 60-104 bytes of spill) measured fastest on a 1080p frame of the
// 512^3 volume: 23.2 ms against 24.8 at 4 waves, 29.5 at 3 and 26.6 at 6.
// Other sources keep 3 waves (168 VGPRs instead of 204 at 2, 35 % faster).
// Cubic filters keep the compiler's choice.
template <typename Src, int kFilter, int kQuery>
__global__ __launch_bounds__(kIterBlock) __attribute__((
    amdgpu_waves_per_eu(kFilter == kFilterCubic ? 1 : (IsDenseSource<Src>::value ? 5 : 3))))
void kHitRender(const IteratorContextView *__restrict__ cp, AttrView a,
                                                          HitRenderArgs m)
{
  __shared__ float hitT[kMaxShadedHits][kIterBlock], hitS[kMaxShadedHits][kIterBlock],
      hitE[kMaxShadedHits][kIterBlock];
  const uint32_t r = xcdBandBlock(blockIdx.x, gridDim.x) * kIterBlock + threadIdx.x;
  if (r >= m.numRays) return;
  const uint32_t lane          = threadIdx.x;
  const IteratorContextView &c = *cp;
  PointSampler<Src, kFilter> smp(c.sampler.grid, a, c.attributeIndex);
  smp.setTime(m.time);
  const F3 o = {m.org[3 * r], m.org[3 * r + 1], m.org[3 * r + 2]};
  const F3 d = {m.dir[3 * r], m.dir[3 * r + 1], m.dir[3 * r + 2]};
  const F3 light0 = normalize3(F3{1.f, 1.f, 1.f}), light1 = normalize3(F3{1.f, 1.f, -1.f});
  const F3 wi     = normalize3(F3{-1.f * d.x, -1.f * d.y, -1.f * d.z});
  const uint32_t maxShade = m.maxHits < (uint32_t)kMaxShadedHits ? m.maxHits : (uint32_t)kMaxShadedHits;

  HitState st;
  hitInit<hitDepthOf<Src>()>(c, st, o, d, rayRange(m.tRanges, r), m.time);
  bool primary = true;
  uint32_t hits = 0, k = 0;
  float alphaP = 0.f;  // opacity after the hits collected so far
  // the hit being shaded
  F3 p = {0.f, 0.f, 0.f}, N = {0.f, 0.f, 0.f};
  float ci = 0.f, co = 0.f, illum = 0.f;
  int light = -1;  // next light to test; -1: hit k not started
  float cr = 0.f, cg = 0.f, cb = 0.f, alpha = 0.f;
  for (;;) {
    if (primary && !(alphaP < 0.99f && hits < maxShade)) {
      primary = false;
    } else {
      HitOut h;
      const bool got = hitNext<decltype(smp), kQuery>(c, smp, st, h);
      if (primary) {
        if (got) {
          hitT[hits][lane] = h.t;
          hitS[hits][lane] = h.sample;
          hitE[hits][lane] = h.epsilon;
          hits++;
          alphaP = alphaP + (1.f - alphaP) * 0.6f;
          continue;
        }
        primary = false;
      } else {
        if (!got) illum += fabsf(co) * (light == 0 ? 1.f : 0.8f);
        light++;
      }
    }
    // Advance the shading to the next shadow query, compositing every hit
    // whose lights are done.
    bool query = false;
    while (k < hits) {
      if (light < 0) {  // start hit k
        const float t = hitT[k][lane];
        p             = F3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
        const F3 q    = xfmPoint(c.sampler.grid.o2i, p);
        F3 gi         = {0.f, 0.f, 0.f};
        if (c.sampler.gradientFilter == kFilterLinear)
          gi = gradientLinear(smp.src, q);
        else if (c.sampler.gradientFilter == kFilterCubic)
          gi = gradientCubic(smp.src, q);
        N = normalize3(xfmNormal(c.sampler.grid.o2i, gi));
        if (isnan(N.x) || isnan(N.y) || isnan(N.z)) N = F3{0.f, 0.f, 0.f};
        ci = dot3(N, wi);
        if (sqrtf(dot3(N, N)) > 0.f) {
          illum = 0.1f;  // ambient
          light = 0;
        } else {
          illum = 1.f;  // no normal: unshaded
          light = 2;
        }
      }
      // the next light on the viewer's side of the surface
      while (light < 2) {
        co = dot3(N, light == 0 ? light0 : light1);
        if ((co > 0.f) == (ci > 0.f)) break;
        light++;
      }
      if (light < 2) {
        query = true;
        break;
      }
      float albedo[4];
      tfSampleRgba(m.tf, m.tfSize, m.tfLo, m.tfHi, hitS[k][lane], albedo);
      cr    = cr + (1.f - alpha) * illum * albedo[0];
      cg    = cg + (1.f - alpha) * illum * albedo[1];
      cb    = cb + (1.f - alpha) * illum * albedo[2];
      alpha = alpha + (1.f - alpha) * 0.6f;
      k++;
      light = -1;
    }
    if (!query) break;
    hitInit<hitDepthOf<Src>()>(c, st, p, light == 0 ? light0 : light1, Range1f{hitE[k][lane], vinf()}, m.time);
  }
  m.rgba[4 * r]     = cr;
  m.rgba[4 * r + 1] = cg;
  m.rgba[4 * r + 2] = cb;
  m.rgba[4 * r + 3] = alpha;
  if (m.count) m.count[r] = hits;
}

// The kernel is specialised on the context's query kind (denseQueryOf):
// structuredRegular volumes within one level-1 node carry no node-table
// query.
struct HitRenderLaunch
{
  const IteratorContextView *c;
  AttrView a;
  HitRenderArgs m;
  hipStream_t st;
  bool denseQuery;
  template <typename Src, int kFilter, int kQuery>
  hipError_t run() const
  {
    hipLaunchKernelGGL((kHitRender<Src, kFilter, kQuery>), dim3((m.numRays + kIterBlock - 1) / kIterBlock),
                       dim3(kIterBlock), 0, st, c, a, m);
    return hipGetLastError();
  }
  template <typename Src, int kFilter>
  hipError_t operator()() const
  {
    if constexpr (IsDenseSource<Src>::value)
      return denseQuery ? run<Src, kFilter, 1>() : run<Src, kFilter, 0>();
    else
      return run<Src, kFilter, 0>();
  }
};

}  // namespace vklhip
