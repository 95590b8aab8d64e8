// Host-side selection of a kernel specialisation: voxel source type (dense
// element type and 32- / 64-bit offsets, or sparse vdb float / half) and
// filter; sparse sources of observed samplers get their own instantiation.
// `f` is a functor with
//   template <typename Src, int kFilter> hipError_t operator()() const;
// that launches the matching instantiation.
#pragma once

#include "vkl_device.hh"

namespace vklhip {

template <typename Src, typename F>
hipError_t withFilter(int filter, const F &f)
{
  if (filter == kFilterNearest) return f.template operator()<Src, kFilterNearest>();
  if (filter == kFilterCubic) return f.template operator()<Src, kFilterCubic>();
  return f.template operator()<Src, kFilterLinear>();
}

// Source kinds a dispatch instantiates kernels for; the hit kernels split
// them over translation units that compile in parallel.
enum : unsigned
{
  kDenseIntKinds       = 1u,  // uint8, int16, uint16 voxels
  kSparseKinds         = 2u,  // vdb, unobserved samplers
  kTemporalKinds       = 4u,
  kDenseFloatKinds     = 8u,  // half, float, double voxels
  kSparseObservedKinds = 16u,
  kDenseKinds          = kDenseIntKinds | kDenseFloatKinds,
  kAllKinds            = 31u
};

template <typename T>
constexpr unsigned denseKindOf()
{
  return std::is_integral<T>::value ? kDenseIntKinds : kDenseFloatKinds;
}

template <typename T, unsigned kKinds, typename F>
hipError_t withDense(const GridView &g, int filter, const F &f)
{
  if (g.temporalFormat) {
    if constexpr ((kKinds & kTemporalKinds) != 0)
      return withFilter<TemporalDenseSource<T>>(filter, f);
    else
      return hipErrorInvalidValue;
  } else {
    if constexpr ((kKinds & denseKindOf<T>()) != 0) {
      const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
      if (nbz * g.layout.strideZ >= (1ull << 32)) return withFilter<DenseSource<T, true>>(filter, f);
      return withFilter<DenseSource<T, false>>(filter, f);
    } else {
      return hipErrorInvalidValue;
    }
  }
}

template <unsigned kKinds = kAllKinds, typename F>
hipError_t dispatchSource(const GridView &g, const AttrView &a, int filter, const F &f)
{
  if (g.dense) {
    if constexpr ((kKinds & (kDenseKinds | kTemporalKinds)) != 0) {
      switch (a.type) {
      case kTypeUchar: return withDense<uint8_t, kKinds>(g, filter, f);
      case kTypeShort: return withDense<int16_t, kKinds>(g, filter, f);
      case kTypeUshort: return withDense<uint16_t, kKinds>(g, filter, f);
      case kTypeHalf: return withDense<__half, kKinds>(g, filter, f);
      case kTypeFloat: return withDense<float, kKinds>(g, filter, f);
      case kTypeDouble: return withDense<double, kKinds>(g, filter, f);
      default: return hipErrorInvalidValue;
      }
    } else {
      return hipErrorInvalidValue;
    }
  }
  if (g.temporalFormat) {
    if constexpr ((kKinds & kTemporalKinds) != 0) {
      if (a.type == kTypeHalf) return withFilter<TemporalSparseSource<__half>>(filter, f);
      if (a.type == kTypeFloat) return withFilter<TemporalSparseSource<float>>(filter, f);
    }
    return hipErrorInvalidValue;
  }
  if (g.leafAccess) {
    if constexpr ((kKinds & kSparseObservedKinds) != 0) {
      if (a.type == kTypeHalf) return withFilter<SparseSource<__half, true>>(filter, f);
      if (a.type == kTypeFloat) return withFilter<SparseSource<float, true>>(filter, f);
    }
    return hipErrorInvalidValue;
  }
  if constexpr ((kKinds & kSparseKinds) != 0) {
    if (a.type == kTypeHalf) return withFilter<SparseSource<__half>>(filter, f);
    if (a.type == kTypeFloat) return withFilter<SparseSource<float>>(filter, f);
  }
  return hipErrorInvalidValue;
}

}  // namespace vklhip
