// Tricubic samples and gradients of structuredRegular volumes with cubic
// records (AttrView::cubic), and the commit-time packing of those records.
//
// The 4x4x4 stencil of a cell is 4 consecutive records, 64 elements in a row
// (256 B for float32): 2 or 3 cache lines, against ~6 for the 8
// corner-packed cell records and ~16 for the rows of the x-fastest array.
// Loaded one lane at a time, a lane's 16-byte loads each touch 64 different
// lines across the wavefront and the lines leave L1 between them; so the
// lanes of a wavefront load their 64 stencils together. Each load
// instruction reads 64 / C whole stencils of C 16-byte chunks (C =
// 4 sizeof(T)), the chunks are staged in LDS (rotated per stencil so that
// neither the writes nor the reads of a wavefront meet in one bank), and each
// lane reads its own stencil back and evaluates it with the arithmetic of the
// per-point path (cubicFromTaps / cubicGradientFromTaps): the taps are the
// same values, so samples are bit-identical to it.

#include "kernels.h"
#include "vkl_device.hh"

namespace vklhip {

template <typename T>
struct StencilChunks
{
  static constexpr uint32_t kChunks  = 4 * sizeof(T);                  // 16-byte chunks per stencil
  static constexpr uint32_t kRecord  = sizeof(T);                      // chunks per record
  static constexpr uint32_t kPerLoad = 64 / kChunks;                   // stencils per load instruction
  static constexpr uint32_t kBatch   = kChunks < 16 ? kChunks : 16;    // loads in flight per lane
  static constexpr uint32_t kElems   = 16 / sizeof(T);                 // elements per chunk
  static constexpr uint32_t kWaves   = sizeof(T) == 8 ? 2 : 4;         // wavefronts per workgroup
  static constexpr uint32_t kThreads = 64 * kWaves;
};

constexpr uint64_t kNoStencil = ~0ull;

// Stencils on the domain's faces come from the records too: the 4 records
// loaded start at x0 = clamp(x - 1, 0, nx - 4) (so they exist), tap column
// x - 1 + i reads record x - 1 + i - x0, and taps outside the domain (in x,
// or the y / z elements a record holds clamped) take the background, as the
// per-point path's reads do. Only points whose own y or z cell lies outside
// the domain (no record to start from) take the per-point path.
// kWhole: the records are whole stencils (AttrView::cubicStencils), one per
// cell; stencils whose cell lies in the domain load their own record.
template <typename T, bool kIdx64, bool kGradient, bool kWhole>
__global__ __launch_bounds__(StencilChunks<T>::kThreads) void kCubicRecords(GridView g, AttrView a, uint32_t attr,
                                                                             uint64_t n, const float *__restrict__ xyz,
                                                                             float *__restrict__ out,
                                                                             uint32_t outStride)
{
  using K              = StencilChunks<T>;
  constexpr uint32_t C = K::kChunks;
  __shared__ uint4 stage[K::kWaves][64 * C];
  __shared__ uint64_t first[K::kWaves][64];
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t i = (uint64_t)blockIdx.x * K::kThreads + threadIdx.x;
  const bool live  = i < n;
  F3 p             = {0.f, 0.f, 0.f};
  if (live) {
    // Points stream through once (non-temporal), as in the per-point kernels.
    const F3 oc = {__builtin_nontemporal_load(xyz + 3 * i), __builtin_nontemporal_load(xyz + 3 * i + 1),
                   __builtin_nontemporal_load(xyz + 3 * i + 2)};
    p = xfmPoint(g.o2i, oc);
  }
  const DenseSource<T, kIdx64> src(g, a, attr);
  const I3 ic       = {floorToInt(p.x), floorToInt(p.y), floorToInt(p.z)};
  const uint32_t oy = (uint32_t)(ic.y - src.ry), oz = (uint32_t)(ic.z - src.rz);
  const bool staged = live && oy < src.ny && oz < src.nz && (!kWhole || (uint32_t)(ic.x - src.rx) < src.nx);
  // First stencil column and first loaded record column (row records:
  // nx >= 4; whole stencils: the cell's own record).
  const int64_t sx = (int64_t)ic.x - src.rx - 1;
  const int64_t x0 = kWhole ? sx + 1 : (sx < 0 ? 0 : (sx > (int64_t)src.nx - 4 ? (int64_t)src.nx - 4 : sx));
  first[w][lane]   = staged ? (((uint64_t)oz * src.ny + oy) * src.nx + (uint64_t)x0) * (kWhole ? C : K::kRecord)
                            : kNoStencil;
  __syncthreads();
  const uint4 *rec = static_cast<const uint4 *>(a.cubic);
  const uint32_t c = lane % C;
#pragma unroll
  for (uint32_t r0 = 0; r0 < C; r0 += K::kBatch) {
    uint4 v[K::kBatch];
#pragma unroll
    for (uint32_t k = 0; k < K::kBatch; k++) {
      const uint32_t s = (r0 + k) * K::kPerLoad + lane / C;
      const uint64_t f = first[w][s];
      v[k]             = f != kNoStencil ? rec[f + c] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (uint32_t k = 0; k < K::kBatch; k++) {
      const uint32_t s                = (r0 + k) * K::kPerLoad + lane / C;
      stage[w][s * C + ((c + s) % C)] = v[k];
    }
  }
  __syncthreads();
  if (!live) return;
  float S[64];
  if (staged) {
    bool vy[4], vz[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      vy[j] = oy - 1 + j < src.ny;
      vz[j] = oz - 1 + j < src.nz;
    }
#pragma unroll
    for (uint32_t t = 0; t < 4; t++) {
      const int64_t x   = sx + t;
      const bool vx     = x >= 0 && x < (int64_t)src.nx;
      // record within the 4 loaded (whole stencils: the stencil's column t)
      const uint32_t at = kWhole ? t : (vx ? (uint32_t)(x - x0) : 0u);
#pragma unroll
      for (uint32_t cc = 0; cc < K::kRecord; cc++) {
        const uint32_t q = at * K::kRecord + cc;
        const uint4 u    = stage[w][lane * C + ((q + lane) % C)];
        T e[K::kElems];
        __builtin_memcpy(e, &u, 16);
#pragma unroll
        for (uint32_t m = 0; m < K::kElems; m++) {
          const uint32_t el = cc * K::kElems + m;  // 4 j + k within the record
          S[16 * t + el]    = vx && vy[el >> 2] && vz[el & 3] ? toFloat(e[m]) : src.bg;
        }
      }
    }
  }
  if (kGradient) {
    const F3 gi    = staged ? cubicGradientFromTaps(S, p, ic) : gradientCubic(src, p);
    const F3 r     = xfmNormal(g.o2i, gi);
    out[3 * i]     = r.x;
    out[3 * i + 1] = r.y;
    out[3 * i + 2] = r.z;
  } else {
    const float r = staged ? cubicFromTaps(S, p, ic) : sampleCubic(src, p);
    __builtin_nontemporal_store(r, out + i * outStride);
  }
}

template <typename E>
struct alignas(16 * sizeof(E)) CubicRecord
{
  E v[16];
};

// Record of voxel (x, y, z): e[4 j + k] = v(x, y - 1 + j, z - 1 + k), y and
// z clamped into the domain (clamped elements are never read: such stencils
// leave the domain and take the per-point path). One work-item per voxel.
template <typename E>
__global__ __launch_bounds__(256) void kPackCubicRecords(const E *__restrict__ src, E *__restrict__ dst, uint32_t nx,
                                                         uint32_t ny, uint32_t nz)
{
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= (uint64_t)nx * ny * nz) return;
  const uint32_t x = (uint32_t)(c % nx), y = (uint32_t)((c / nx) % ny), z = (uint32_t)(c / ((uint64_t)nx * ny));
  CubicRecord<E> r;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int32_t yj   = (int32_t)y - 1 + j;
    const uint32_t yy  = yj < 0 ? 0u : (yj >= (int32_t)ny ? ny - 1 : (uint32_t)yj);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int32_t zk  = (int32_t)z - 1 + k;
      const uint32_t zz = zk < 0 ? 0u : (zk >= (int32_t)nz ? nz - 1 : (uint32_t)zk);
      r.v[4 * j + k]    = src[((uint64_t)zz * ny + yy) * nx + x];
    }
  }
  reinterpret_cast<CubicRecord<E> *>(dst)[c] = r;
}

template <typename E>
struct alignas(64 * sizeof(E) < 256 ? 64 * sizeof(E) : 256) StencilRecord
{
  E v[64];
};

// Whole stencil of cell (x, y, z): e[16 i + 4 j + k] = v(x - 1 + i, y - 1 + j,
// z - 1 + k), coordinates clamped into the domain (clamped elements are
// replaced by the background where the sampler reads them). One work-item
// per cell.
template <typename E>
__global__ __launch_bounds__(256) void kPackCubicStencils(const E *__restrict__ src, E *__restrict__ dst, uint32_t nx,
                                                          uint32_t ny, uint32_t nz)
{
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= (uint64_t)nx * ny * nz) return;
  const uint32_t x = (uint32_t)(c % nx), y = (uint32_t)((c / nx) % ny), z = (uint32_t)(c / ((uint64_t)nx * ny));
  auto clampTo = [](uint32_t v, int d, uint32_t n) {
    const int32_t q = (int32_t)v - 1 + d;
    return q < 0 ? 0u : (q >= (int32_t)n ? n - 1 : (uint32_t)q);
  };
  StencilRecord<E> r;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t xx = clampTo(x, i, nx);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t yy = clampTo(y, j, ny);
#pragma unroll
      for (int k = 0; k < 4; k++) r.v[16 * i + 4 * j + k] = src[((uint64_t)clampTo(z, k, nz) * ny + yy) * nx + xx];
    }
  }
  reinterpret_cast<StencilRecord<E> *>(dst)[c] = r;
}

namespace {

template <typename T, bool kIdx64>
hipError_t launchTyped(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                       float *out, uint32_t stride, bool gradient, hipStream_t st)
{
  using K = StencilChunks<T>;
  const dim3 grid((unsigned)((n + K::kThreads - 1) / K::kThreads)), block(K::kThreads);
  if (a.cubicStencils) {
    if (gradient)
      hipLaunchKernelGGL((kCubicRecords<T, kIdx64, true, true>), grid, block, 0, st, s.grid, a, attr, n, xyz, out,
                         stride);
    else
      hipLaunchKernelGGL((kCubicRecords<T, kIdx64, false, true>), grid, block, 0, st, s.grid, a, attr, n, xyz, out,
                         stride);
  } else {
    if (gradient)
      hipLaunchKernelGGL((kCubicRecords<T, kIdx64, true, false>), grid, block, 0, st, s.grid, a, attr, n, xyz, out,
                         stride);
    else
      hipLaunchKernelGGL((kCubicRecords<T, kIdx64, false, false>), grid, block, 0, st, s.grid, a, attr, n, xyz, out,
                         stride);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launchType(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                      float *out, uint32_t stride, bool gradient, hipStream_t st)
{
  // The per-point fallback addresses the volume's own layout; 64-bit
  // offsets when it needs them (the records are always addressed in 64 bits).
  const GridView &g  = s.grid;
  const uint64_t nbz = ((uint64_t)g.activeSize[2] + (1u << g.layout.sz) - 1) >> g.layout.sz;
  if (nbz * g.layout.strideZ >= (1ull << 32))
    return launchTyped<T, true>(s, a, attr, n, xyz, out, stride, gradient, st);
  return launchTyped<T, false>(s, a, attr, n, xyz, out, stride, gradient, st);
}

}  // namespace

hipError_t launchCubicRecords(const SamplerView &s, const AttrView &a, uint32_t attr, uint64_t n, const float *xyz,
                              float *out, uint32_t stride, bool gradient, hipStream_t stream)
{
  if (n == 0) return hipSuccess;
  if (!s.grid.dense || s.grid.temporalFormat || !a.cubic || (!a.cubicStencils && s.grid.activeSize[0] < 4))
    return hipErrorInvalidValue;
  switch (a.type) {
  case kTypeUchar: return launchType<uint8_t>(s, a, attr, n, xyz, out, stride, gradient, stream);
  case kTypeShort: return launchType<int16_t>(s, a, attr, n, xyz, out, stride, gradient, stream);
  case kTypeUshort: return launchType<uint16_t>(s, a, attr, n, xyz, out, stride, gradient, stream);
  case kTypeHalf: return launchType<__half>(s, a, attr, n, xyz, out, stride, gradient, stream);
  case kTypeFloat: return launchType<float>(s, a, attr, n, xyz, out, stride, gradient, stream);
  case kTypeDouble: return launchType<double>(s, a, attr, n, xyz, out, stride, gradient, stream);
  default: return hipErrorInvalidValue;
  }
}

namespace {

template <typename E>
void packTyped(bool stencils, const void *src, void *dst, const uint32_t dims[3], dim3 grid, hipStream_t stream)
{
  if (stencils)
    hipLaunchKernelGGL(kPackCubicStencils<E>, grid, dim3(256), 0, stream, (const E *)src, (E *)dst, dims[0],
                       dims[1], dims[2]);
  else
    hipLaunchKernelGGL(kPackCubicRecords<E>, grid, dim3(256), 0, stream, (const E *)src, (E *)dst, dims[0], dims[1],
                       dims[2]);
}

}  // namespace

hipError_t launchPackCubicRecords(const void *src, void *dst, uint32_t elemSize, const uint32_t dims[3],
                                  bool stencils, hipStream_t stream)
{
  const uint64_t voxels = (uint64_t)dims[0] * dims[1] * dims[2];
  const dim3 grid((unsigned)((voxels + 255) / 256));
  switch (elemSize) {
  case 1: packTyped<uint8_t>(stencils, src, dst, dims, grid, stream); break;
  case 2: packTyped<uint16_t>(stencils, src, dst, dims, grid, stream); break;
  case 4: packTyped<uint32_t>(stencils, src, dst, dims, grid, stream); break;
  case 8: packTyped<uint64_t>(stencils, src, dst, dims, grid, stream); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace vklhip
