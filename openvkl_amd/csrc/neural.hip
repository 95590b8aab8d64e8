// Neural-field sampler kernels (gfx950).
//
// Encoding: 16-level hash grid (Mueller et al. 2022), 2 binary16 features per
// entry, 2^19 entries per level, trilinear interpolation of the 8 cell
// corners; dense indexing on levels whose (res+1)^3 grid fits in the table.
// MLP: 32 -> 64 -> 64 -> 1, ReLU, no biases. Hidden layers and all backward
// matrix products run on v_mfma_f32_16x16x32_f16 (binary16 operands, binary32
// accumulation); one wavefront owns a tile of 64 samples, staged through LDS.
//
// MFMA 16x16x32 lane layout (lane l, lr = l % 16, lg = l / 16):
//   A (16 x 32): row lr, k = 8 lg .. 8 lg + 7
//   B (32 x 16): col lr, k = 8 lg .. 8 lg + 7
//   C (16 x 16): col lr, rows 4 lg .. 4 lg + 3

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "neural.h"
#include "vkl_device.hh"

namespace vklhip {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWave   = 64;
constexpr int kLdX    = 40;  // sX row stride (halves): 32 + 8 pad, 80 B
constexpr int kLdH    = 72;  // hidden row stride (halves): 64 + 8 pad, 144 B
constexpr uint32_t kP1 = 2654435761u, kP2 = 805459861u;

__device__ __forceinline__ f4 mma(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }

__device__ __forceinline__ h8 ld8(const _Float16 *p) { return *reinterpret_cast<const h8 *>(p); }

__device__ __forceinline__ float h2f(uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); }

// ---- encoding --------------------------------------------------------------------

__device__ __forceinline__ uint32_t nfIndex(uint32_t res, uint32_t x, uint32_t y, uint32_t z)
{
  const uint32_t side = res + 1;
  if ((uint64_t)side * side * side <= kNfT) return x + side * (y + side * z);
  return ((x * 1u) ^ (y * kP1) ^ (z * kP2)) & (kNfT - 1);
}

// Normalised box coordinates in [0, 1] of an object-space point.
__device__ __forceinline__ F3 nfUnit(const NfModel &m, F3 oc)
{
  const F3 p = xfmPoint(m.o2i, oc);
  F3 x = {(p.x - m.lo[0]) * m.invExtent, (p.y - m.lo[1]) * m.invExtent, (p.z - m.lo[2]) * m.invExtent};
  x.x  = fminf(fmaxf(x.x, 0.f), 1.f);
  x.y  = fminf(fmaxf(x.y, 0.f), 1.f);
  x.z  = fminf(fmaxf(x.z, 0.f), 1.f);
  return x;
}

struct Corners
{
  uint32_t idx[8];
  float w[8];
};

__device__ __forceinline__ void nfCorners(const NfModel &m, int l, F3 x, Corners &c)
{
  const uint32_t res = m.res[l];
  const float r      = (float)res;
  const float px = x.x * r, py = x.y * r, pz = x.z * r;
  const int32_t cx = min(max((int32_t)floorf(px), 0), (int32_t)res - 1);
  const int32_t cy = min(max((int32_t)floorf(py), 0), (int32_t)res - 1);
  const int32_t cz = min(max((int32_t)floorf(pz), 0), (int32_t)res - 1);
  const float wx = px - (float)cx, wy = py - (float)cy, wz = pz - (float)cz;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int dx = (k >> 2) & 1, dy = (k >> 1) & 1, dz = k & 1;
    c.w[k]   = (dx ? wx : 1.f - wx) * (dy ? wy : 1.f - wy) * (dz ? wz : 1.f - wz);
    c.idx[k] = nfIndex(res, (uint32_t)(cx + dx), (uint32_t)(cy + dy), (uint32_t)(cz + dz));
  }
}

// One sample's 32 encoding features, rounded to binary16, into row `row`.
__device__ __forceinline__ void nfEncode(const NfModel &m, F3 x, _Float16 *row)
{
#pragma unroll 1
  for (int l = 0; l < kNfLevels; l++) {
    Corners c;
    nfCorners(m, l, x, c);
    float f0 = 0.f, f1 = 0.f;
    const uint32_t *tab = reinterpret_cast<const uint32_t *>(m.w.table) + (size_t)l * kNfT;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t e = tab[c.idx[k]];
      f0 += c.w[k] * h2f((uint16_t)(e & 0xffffu));
      f1 += c.w[k] * h2f((uint16_t)(e >> 16));
    }
    row[2 * l]     = (_Float16)f0;
    row[2 * l + 1] = (_Float16)f1;
  }
}

// ---- MLP forward on a 64-sample tile ------------------------------------------------

// sX (64 x 32) -> sH1 (64 x 64) -> sH2 (64 x 64), ReLU, binary16 activations.
__device__ __forceinline__ void mlpHidden(const NfModel &m, const _Float16 *sX, _Float16 *sH1, _Float16 *sH2,
                                          int lr, int lg)
{
  const _Float16 *W1T = reinterpret_cast<const _Float16 *>(m.w.W1T);
  const _Float16 *W2T = reinterpret_cast<const _Float16 *>(m.w.W2T);
  h8 b1[4];
#pragma unroll
  for (int c = 0; c < 4; c++) b1[c] = ld8(W1T + (16 * c + lr) * kNfIn + 8 * lg);
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const h8 a = ld8(sX + (16 * r + lr) * kLdX + 8 * lg);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      acc    = mma(a, b1[c], acc);
#pragma unroll
      for (int i = 0; i < 4; i++) sH1[(16 * r + 4 * lg + i) * kLdH + 16 * c + lr] = (_Float16)fmaxf(acc[i], 0.f);
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const h8 bA = ld8(W2T + (16 * c + lr) * kNfHidden + 8 * lg);
    const h8 bB = ld8(W2T + (16 * c + lr) * kNfHidden + 32 + 8 * lg);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      acc    = mma(ld8(sH1 + (16 * r + lr) * kLdH + 8 * lg), bA, acc);
      acc    = mma(ld8(sH1 + (16 * r + lr) * kLdH + 32 + 8 * lg), bB, acc);
#pragma unroll
      for (int i = 0; i < 4; i++) sH2[(16 * r + 4 * lg + i) * kLdH + 16 * c + lr] = (_Float16)fmaxf(acc[i], 0.f);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ float mlpOut(const NfModel &m, const _Float16 *sH2, int lane)
{
  float y = 0.f;
#pragma unroll 8
  for (int j = 0; j < kNfHidden; j++) y += (float)sH2[lane * kLdH + j] * m.w.W3[j];
  return y;
}

// Forward pass of inference: the same binary16 H1 and H2 as mlpHidden, but
// H2 never reaches LDS. Each lane folds its MFMA outputs (4 rows x 4 column
// tiles, column 16c + lr) into per-row partial dot products with W3, which
// go through the input tile's space (dead once H1 is written: 64 rows x
// kLdP floats fit in 64 x kLdX halves) and are summed per row by the row's
// lane. Only the fp32 summation order differs from mlpHidden + mlpOut.
// Without the H2 tile a block holds 14 KB of LDS instead of 23 KB.
constexpr int kLdP = 17;  // partial-sum row stride (floats)
static_assert(kWave * kLdP * sizeof(float) <= kWave * kLdX * sizeof(_Float16), "partials must fit the input tile");

__device__ __forceinline__ float mlpForward(const NfModel &m, _Float16 *sX, _Float16 *sH1, int lr, int lg, int lane)
{
  const _Float16 *W1T = reinterpret_cast<const _Float16 *>(m.w.W1T);
  const _Float16 *W2T = reinterpret_cast<const _Float16 *>(m.w.W2T);
  h8 b1[4];
#pragma unroll
  for (int c = 0; c < 4; c++) b1[c] = ld8(W1T + (16 * c + lr) * kNfIn + 8 * lg);
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const h8 a = ld8(sX + (16 * r + lr) * kLdX + 8 * lg);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      acc    = mma(a, b1[c], acc);
#pragma unroll
      for (int i = 0; i < 4; i++) sH1[(16 * r + 4 * lg + i) * kLdH + 16 * c + lr] = (_Float16)fmaxf(acc[i], 0.f);
    }
  }
  __syncthreads();
  float part[4][4] = {};
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const h8 bA    = ld8(W2T + (16 * c + lr) * kNfHidden + 8 * lg);
    const h8 bB    = ld8(W2T + (16 * c + lr) * kNfHidden + 32 + 8 * lg);
    const float w3 = m.w.W3[16 * c + lr];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
      acc    = mma(ld8(sH1 + (16 * r + lr) * kLdH + 8 * lg), bA, acc);
      acc    = mma(ld8(sH1 + (16 * r + lr) * kLdH + 32 + 8 * lg), bB, acc);
#pragma unroll
      for (int i = 0; i < 4; i++) part[r][i] += (float)(_Float16)fmaxf(acc[i], 0.f) * w3;
    }
  }
  float *sP = reinterpret_cast<float *>(sX);
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int i = 0; i < 4; i++) sP[(16 * r + 4 * lg + i) * kLdP + lr] = part[r][i];
  __syncthreads();
  float y = 0.f;
#pragma unroll
  for (int j = 0; j < 16; j++) y += sP[lane * kLdP + j];
  return y;
}

// ---- training ------------------------------------------------------------------------

__device__ __forceinline__ uint64_t splitmix(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ float unit24(uint64_t h) { return (float)(h >> 40) * (1.f / 16777216.f); }

__global__ void kNfGenerate(const float *__restrict__ i2o, const float *__restrict__ boxLo,
                            const float *__restrict__ boxHi, const int32_t *__restrict__ leafOrigins,
                            uint64_t numLeaves, float leafFrac, uint64_t n, uint64_t seed, float *__restrict__ xyz)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t h0 = splitmix(seed ^ (i * 0x632BE59BD9B4E019ull));
  const uint64_t h1 = splitmix(h0), h2 = splitmix(h1), h3 = splitmix(h2), h4 = splitmix(h3);
  F3 p;
  if (numLeaves && unit24(h0) < leafFrac) {
    const uint64_t leaf = (h1 >> 16) % numLeaves;
    p = {leafOrigins[3 * leaf] + 8.f * unit24(h2), leafOrigins[3 * leaf + 1] + 8.f * unit24(h3),
         leafOrigins[3 * leaf + 2] + 8.f * unit24(h4)};
  } else {
    p = {boxLo[0] + (boxHi[0] - boxLo[0]) * unit24(h2), boxLo[1] + (boxHi[1] - boxLo[1]) * unit24(h3),
         boxLo[2] + (boxHi[2] - boxLo[2]) * unit24(h4)};
  }
  const F3 o = xfmPoint(i2o, p);
  xyz[3 * i]     = o.x;
  xyz[3 * i + 1] = o.y;
  xyz[3 * i + 2] = o.z;
}

hipError_t launchNfGenerate(const float *i2o, const float *boxLo, const float *boxHi, const int32_t *leafOrigins,
                            uint64_t numLeaves, float leafFrac, uint64_t n, uint64_t seed, float *xyz,
                            hipStream_t stream)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(kNfGenerate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, i2o, boxLo, boxHi,
                     leafOrigins, numLeaves, leafFrac, n, seed, xyz);
  return hipGetLastError();
}

// One wavefront per workgroup; persistent over 64-sample tiles. Forward, then
// backward of sum((y - t)^2):
//   dy = 2 (y - t); dW3 = H2^T dy; G2 = dy W3^T o [H2 > 0]; dW2 = H1^T G2;
//   G1 = G2 W2^T o [H1 > 0]; dW1 = X^T G1; dX = G1 W1^T -> table (atomics).
// Weight gradients accumulate in MFMA registers across tiles and are flushed
// once per workgroup.
__global__ __launch_bounds__(kWave) void kNfTrain(NfModel m, uint64_t n, const float *__restrict__ xyz,
                                                   const float *__restrict__ target, NfGrads g)
{
  __shared__ __attribute__((aligned(16))) _Float16 sX[kWave * kLdX];
  __shared__ __attribute__((aligned(16))) _Float16 sH1[kWave * kLdH];
  __shared__ __attribute__((aligned(16))) _Float16 sH2[kWave * kLdH];  // H2, later G1
  __shared__ __attribute__((aligned(16))) _Float16 sG[kWave * kLdH];   // G2, later dX (fp32)
  __shared__ float sDy[kWave];
  const int lane = threadIdx.x, lr = lane & 15, lg = lane >> 4;
  const _Float16 *W1 = reinterpret_cast<const _Float16 *>(m.w.W1);
  const _Float16 *W2 = reinterpret_cast<const _Float16 *>(m.w.W2);
  f4 dW2[4][4], dW1[2][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) dW2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) dW1[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  float dW3 = 0.f, loss = 0.f;
  const uint64_t tiles = (n + kWave - 1) / kWave;
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint64_t s     = t * kWave + lane;
    const bool valid     = s < n;
    const F3 oc          = valid ? F3{xyz[3 * s], xyz[3 * s + 1], xyz[3 * s + 2]} : F3{0.f, 0.f, 0.f};
    const F3 x           = nfUnit(m, oc);
    nfEncode(m, x, sX + lane * kLdX);
    __syncthreads();
    mlpHidden(m, sX, sH1, sH2, lr, lg);
    const float y  = mlpOut(m, sH2, lane);
    // Samples whose exact value is undefined (NaN background taps) carry no loss.
    const float tv = valid ? target[s] : 0.f;
    const float e  = (valid && !isnan(tv)) ? y - tv : 0.f;
    const float dy = 2.f * e;
    loss += e * e;
    sDy[lane] = dy;
    __syncthreads();
    // dW3 (lane = hidden unit j) and G2 (lane = sample row).
    for (int q = 0; q < kWave; q++) dW3 += sDy[q] * (float)sH2[q * kLdH + lane];
    for (int j = 0; j < kNfHidden; j++) {
      const float h = (float)sH2[lane * kLdH + j];
      sG[lane * kLdH + j] = (_Float16)(h > 0.f ? dy * m.w.W3[j] : 0.f);
    }
    __syncthreads();
    // dW2 += H1^T G2 (K = samples).
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      h8 a[4], b[4];
#pragma unroll
      for (int mm = 0; mm < 4; mm++)
#pragma unroll
        for (int q = 0; q < 8; q++) {
          a[mm][q] = sH1[(32 * ks + 8 * lg + q) * kLdH + 16 * mm + lr];
          b[mm][q] = sG[(32 * ks + 8 * lg + q) * kLdH + 16 * mm + lr];
        }
#pragma unroll
      for (int mm = 0; mm < 4; mm++)
#pragma unroll
        for (int nn = 0; nn < 4; nn++) dW2[mm][nn] = mma(a[mm], b[nn], dW2[mm][nn]);
    }
    // G1 = (G2 W2^T) o [H1 > 0] into the sH2 buffer.
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const h8 bA = ld8(W2 + (16 * c + lr) * kNfHidden + 8 * lg);
      const h8 bB = ld8(W2 + (16 * c + lr) * kNfHidden + 32 + 8 * lg);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc    = mma(ld8(sG + (16 * r + lr) * kLdH + 8 * lg), bA, acc);
        acc    = mma(ld8(sG + (16 * r + lr) * kLdH + 32 + 8 * lg), bB, acc);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int row = 16 * r + 4 * lg + i, col = 16 * c + lr;
          sH2[row * kLdH + col] = (_Float16)((float)sH1[row * kLdH + col] > 0.f ? acc[i] : 0.f);
        }
      }
    }
    __syncthreads();
    // dW1 += X^T G1 (K = samples).
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      h8 a[2], b[4];
#pragma unroll
      for (int mm = 0; mm < 2; mm++)
#pragma unroll
        for (int q = 0; q < 8; q++) a[mm][q] = sX[(32 * ks + 8 * lg + q) * kLdX + 16 * mm + lr];
#pragma unroll
      for (int nn = 0; nn < 4; nn++)
#pragma unroll
        for (int q = 0; q < 8; q++) b[nn][q] = sH2[(32 * ks + 8 * lg + q) * kLdH + 16 * nn + lr];
#pragma unroll
      for (int mm = 0; mm < 2; mm++)
#pragma unroll
        for (int nn = 0; nn < 4; nn++) dW1[mm][nn] = mma(a[mm], b[nn], dW1[mm][nn]);
    }
    // dX = G1 W1^T (64 x 32) into sG reinterpreted as fp32 [64][33].
    float *sDX = reinterpret_cast<float *>(sG);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const h8 bA = ld8(W1 + (16 * c + lr) * kNfHidden + 8 * lg);
      const h8 bB = ld8(W1 + (16 * c + lr) * kNfHidden + 32 + 8 * lg);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc    = mma(ld8(sH2 + (16 * r + lr) * kLdH + 8 * lg), bA, acc);
        acc    = mma(ld8(sH2 + (16 * r + lr) * kLdH + 32 + 8 * lg), bB, acc);
#pragma unroll
        for (int i = 0; i < 4; i++) sDX[(16 * r + 4 * lg + i) * 33 + 16 * c + lr] = acc[i];
      }
    }
    __syncthreads();
    // Table gradients: each lane scatters its own sample.
    if (valid) {
#pragma unroll 1
      for (int l = 0; l < kNfLevels; l++) {
        Corners c;
        nfCorners(m, l, x, c);
        const float g0 = sDX[lane * 33 + 2 * l], g1 = sDX[lane * 33 + 2 * l + 1];
        float *gt = g.table + (size_t)l * kNfT * 2;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          atomicAdd(gt + 2 * (size_t)c.idx[k], c.w[k] * g0);
          atomicAdd(gt + 2 * (size_t)c.idx[k] + 1, c.w[k] * g1);
        }
      }
    }
    __syncthreads();
  }
  // Flush register-held weight gradients.
#pragma unroll
  for (int mm = 0; mm < 4; mm++)
#pragma unroll
    for (int nn = 0; nn < 4; nn++)
#pragma unroll
      for (int i = 0; i < 4; i++)
        atomicAdd(g.W2 + (16 * mm + 4 * lg + i) * kNfHidden + 16 * nn + lr, dW2[mm][nn][i]);
#pragma unroll
  for (int mm = 0; mm < 2; mm++)
#pragma unroll
    for (int nn = 0; nn < 4; nn++)
#pragma unroll
      for (int i = 0; i < 4; i++)
        atomicAdd(g.W1 + (16 * mm + 4 * lg + i) * kNfHidden + 16 * nn + lr, dW1[mm][nn][i]);
  atomicAdd(g.W3 + lane, dW3);
  // wave-reduce the loss
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
  if (lane == 0) atomicAdd(g.loss, loss);
}

hipError_t launchNfTrainStep(const NfModel &m, uint64_t n, const float *xyz, const float *target, const NfGrads &g,
                             hipStream_t stream)
{
  if (n == 0) return hipSuccess;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint64_t tiles  = (n + kWave - 1) / kWave;
  // One workgroup per CU, each over several tiles: on a 65 536-sample step
  // (1024 tiles) this measured 7 % faster than one tile per workgroup.
  const unsigned blocks = (unsigned)std::min<uint64_t>(tiles, (uint64_t)cus);
  hipLaunchKernelGGL(kNfTrain, dim3(blocks), dim3(kWave), 0, stream, m, n, xyz, target, g);
  return hipGetLastError();
}

// ---- optimizer ------------------------------------------------------------------------

__device__ __forceinline__ uint16_t f2h(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }

__device__ __forceinline__ float adamStep(float &p, float &mo, float &ve, float gr, float lr, float b1, float b2,
                                          float eps, float c1, float c2)
{
  mo = b1 * mo + (1.f - b1) * gr;
  ve = b2 * ve + (1.f - b2) * gr * gr;
  p -= lr * (mo / c1) / (sqrtf(ve / c2) + eps);
  return p;
}

__global__ void kNfAdamTable(NfAdam a, float *__restrict__ grad, size_t count, float lr, float b1, float b2,
                             float eps, float c1, float c2, float scale)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gr = grad[i] * scale;
  grad[i]        = 0.f;
  if (gr == 0.f && a.mTable[i] == 0.f) return;  // untouched entries stay put
  const float p = adamStep(a.table32[i], a.mTable[i], a.vTable[i], gr, lr, b1, b2, eps, c1, c2);
  a.table16[i]  = f2h(p);
}

__global__ void kNfAdamMlp(NfAdam a, NfGrads g, float lr, float b1, float b2, float eps, float c1, float c2,
                           float scale)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kNfW1) {
    const float gr = g.W1[i] * scale;
    g.W1[i]        = 0.f;
    const float p  = adamStep(a.W1_32[i], a.mW[i], a.vW[i], gr, lr, b1, b2, eps, c1, c2);
    const int e = i / kNfHidden, h = i % kNfHidden;
    a.W1h[i]                 = f2h(p);
    a.W1Th[h * kNfIn + e]    = f2h(p);
  } else if (i < kNfW1 + kNfW2) {
    const int k    = i - kNfW1;
    const float gr = g.W2[k] * scale;
    g.W2[k]        = 0.f;
    const float p  = adamStep(a.W2_32[k], a.mW[i], a.vW[i], gr, lr, b1, b2, eps, c1, c2);
    const int r = k / kNfHidden, c = k % kNfHidden;
    a.W2h[k]                    = f2h(p);
    a.W2Th[c * kNfHidden + r]   = f2h(p);
  } else if (i < kNfW1 + kNfW2 + kNfW3) {
    const int k    = i - kNfW1 - kNfW2;
    const float gr = g.W3[k] * scale;
    g.W3[k]        = 0.f;
    a.W3f[k]       = adamStep(a.W3_32[k], a.mW[i], a.vW[i], gr, lr, b1, b2, eps, c1, c2);
  } else if (i == kNfW1 + kNfW2 + kNfW3) {
    g.loss[0] = 0.f;
  }
}

hipError_t launchNfAdam(const NfAdam &a, const NfGrads &g, float lrTable, float lrMlp, float beta1, float beta2,
                        float eps, uint32_t step, float gradScale, hipStream_t stream)
{
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(kNfAdamTable, dim3((unsigned)((kNfTableEntries + 255) / 256)), dim3(256), 0, stream, a, g.table,
                     kNfTableEntries, lrTable, beta1, beta2, eps, c1, c2, gradScale);
  const int total = kNfW1 + kNfW2 + kNfW3 + 1;
  hipLaunchKernelGGL(kNfAdamMlp, dim3((total + 255) / 256), dim3(256), 0, stream, a, g, lrMlp, beta1, beta2, eps, c1,
                     c2, gradScale);
  return hipGetLastError();
}

// Initial parameters: table U(-1e-4, 1e-4); MLP He-uniform.
__global__ void kNfInit(NfAdam a, uint64_t seed)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kNfTableEntries) {
    const float u = unit24(splitmix(seed ^ (i * 0x9E3779B97F4A7C15ull))) * 2.f - 1.f;
    a.table32[i]  = 1e-4f * u;
    a.table16[i]  = f2h(a.table32[i]);
    a.mTable[i] = a.vTable[i] = 0.f;
  }
  if (i < (size_t)(kNfW1 + kNfW2 + kNfW3)) {
    const float u = unit24(splitmix(~seed ^ (i * 0xD1B54A32D192ED03ull))) * 2.f - 1.f;
    a.mW[i] = a.vW[i] = 0.f;
    if (i < (size_t)kNfW1) {
      const float p = u * sqrtf(6.f / kNfIn);
      a.W1_32[i]    = p;
      a.W1h[i]      = f2h(p);
      a.W1Th[(i % kNfHidden) * kNfIn + i / kNfHidden] = f2h(p);
    } else if (i < (size_t)(kNfW1 + kNfW2)) {
      const size_t k = i - kNfW1;
      const float p  = u * sqrtf(6.f / kNfHidden);
      a.W2_32[k]     = p;
      a.W2h[k]       = f2h(p);
      a.W2Th[(k % kNfHidden) * kNfHidden + k / kNfHidden] = f2h(p);
    } else {
      const size_t k = i - kNfW1 - kNfW2;
      a.W3_32[k]     = u * sqrtf(3.f / kNfHidden);
      a.W3f[k]       = a.W3_32[k];
    }
  }
}

hipError_t launchNfInit(const NfAdam &a, uint64_t seed, hipStream_t stream)
{
  hipLaunchKernelGGL(kNfInit, dim3((unsigned)((kNfTableEntries + 255) / 256)), dim3(256), 0, stream, a, seed);
  return hipGetLastError();
}

// ---- inference -----------------------------------------------------------------------------

// Dense leaf slot of the index-space cell (ix, iy, iz), or -1 (no dense leaf).
// *vox receives the level-2 voxel holding the leaf code; its value range
// (the leaf's majorant over the filter support) bounds every exact sample
// taken in the leaf with the nearest or linear filter.
__device__ __forceinline__ int64_t cellLeaf(const GridView &g, int32_t ix, int32_t iy, int32_t iz, uint64_t *vox)
{
  const uint32_t ox = (uint32_t)(ix - g.rootOrigin[0]), oy = (uint32_t)(iy - g.rootOrigin[1]),
                 oz = (uint32_t)(iz - g.rootOrigin[2]);
  if (!(ox < g.activeSize[0] && oy < g.activeSize[1] && oz < g.activeSize[2])) return -1;
  uint64_t node = 0;
  for (uint32_t l = 0; l < 3; ++l) {
    const uint64_t vo = node * kNumVoxels[l] + offsetToLinear(l, ox, oy, oz);
    const uint64_t v  = g.levels[l].voxels[vo];
    if ((v & 3u) == 2u && l < 2) {
      node = v >> 2;
      continue;
    }
    if (l == 2 && (v & 3u) == 3u && ((v >> 2) & 3u) == 1u) {
      *vox = vo;
      return (int64_t)(g.leafSlot ? g.leafSlot[v >> 6] : (v >> 6));
    }
    return -1;
  }
  return -1;
}

// With useFallback, points whose cell leaf is approved get the network value
// clamped to the leaf's value range when clampBand is set (nearest / linear
// filters, whose exact samples lie in that range); the others are appended to
// `fallback` for the exact pass.
__global__ __launch_bounds__(kWave) void kNfInfer(NfModel m, GridView g, uint32_t attr,
                                                   const uint8_t *__restrict__ leafOk, uint64_t n,
                                                   const float *__restrict__ xyz, float *__restrict__ out,
                                                   uint32_t *__restrict__ fallback,
                                                   uint32_t *__restrict__ numFallback, int useFallback, int clampBand)
{
  __shared__ __attribute__((aligned(16))) _Float16 sX[kWave * kLdX];  // X, later the W3 partial sums
  __shared__ __attribute__((aligned(16))) _Float16 sH1[kWave * kLdH];
  const int lane = threadIdx.x, lr = lane & 15, lg = lane >> 4;
  const uint64_t tiles = (n + kWave - 1) / kWave;
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint64_t s = t * kWave + lane;
    const bool valid = s < n;
    const F3 oc      = valid ? F3{xyz[3 * s], xyz[3 * s + 1], xyz[3 * s + 2]} : F3{0.f, 0.f, 0.f};
    bool exact       = false;
    Range1f band     = {-vinf(), vinf()};
    if (useFallback && valid) {
      const F3 p         = xfmPoint(g.o2i, oc);
      uint64_t vox       = 0;
      const int64_t slot = cellLeaf(g, floorToInt(p.x), floorToInt(p.y), floorToInt(p.z), &vox);
      exact              = slot < 0 || !leafOk[slot];
      if (!exact && clampBand) band = g.levels[2].valueRange[vox * g.numAttributes + attr];
    }
    nfEncode(m, nfUnit(m, oc), sX + lane * kLdX);
    __syncthreads();
    const float y = fminf(fmaxf(mlpForward(m, sX, sH1, lr, lg, lane), band.lower), band.upper);
    // Wave-aggregated compaction of the fallback points.
    const bool need       = valid && exact;
    const uint64_t mask   = __ballot(need);
    const uint32_t below  = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    uint32_t base         = 0;
    if (lane == 0 && mask) base = atomicAdd(numFallback, (uint32_t)__popcll(mask));
    base = __shfl(base, 0);
    if (need) fallback[base + below] = (uint32_t)s;
    if (valid && !exact) out[s] = y;
    __syncthreads();
  }
}

hipError_t launchNfInfer(const NfModel &m, const GridView &g, uint32_t attr, const uint8_t *leafOk, uint64_t n,
                         const float *xyz, float *out, uint32_t *fallback, uint32_t *numFallback, int useFallback,
                         int clampBand, hipStream_t stream)
{
  if (n == 0) return hipSuccess;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint64_t tiles  = (n + kWave - 1) / kWave;
  const unsigned blocks = (unsigned)std::min<uint64_t>(tiles, (uint64_t)cus * 16);
  hipLaunchKernelGGL(kNfInfer, dim3(blocks), dim3(kWave), 0, stream, m, g, attr, leafOk, n, xyz, out, fallback,
                     numFallback, useFallback, clampBand);
  return hipGetLastError();
}

// Exact samples of the fallback points. kObserve: the sampler has a
// LeafNodeAccess observer, so the traversals record their leaves as the
// exact sampler's own kernels do.
// The filter is a template parameter (the same point sampler as the exact
// sampling kernel), so the linear kernel carries no cubic code.
template <typename T, bool kObserve, int kFilter>
__global__ __launch_bounds__(256) void kGatherSample(SamplerView sv, AttrView a, uint32_t attr, const uint32_t *__restrict__ idx,
                              const uint32_t *__restrict__ count, uint64_t maxCount, const float *__restrict__ xyz,
                              float *__restrict__ out)
{
  const uint64_t k   = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t cnt = *count < maxCount ? *count : maxCount;
  if (k >= cnt) return;
  const uint32_t i = idx[k];
  const F3 oc      = {xyz[3 * (uint64_t)i], xyz[3 * (uint64_t)i + 1], xyz[3 * (uint64_t)i + 2]};
  const PointSampler<SparseSource<T, kObserve>, kFilter> smp(sv.grid, a, attr);
  out[i] = smp(oc);
}

namespace {
struct GatherLaunch
{
  const SamplerView &s;
  const AttrView &a;
  uint32_t attr;
  const uint32_t *idx, *count;
  uint64_t maxCount;
  const float *xyz;
  float *out;
  hipStream_t st;
  template <typename T, bool kObserve, int kFilter>
  hipError_t run() const
  {
    const dim3 grid((unsigned)((maxCount + 255) / 256)), block(256);
    hipLaunchKernelGGL((kGatherSample<T, kObserve, kFilter>), grid, block, 0, st, s, a, attr, idx, count, maxCount,
                       xyz, out);
    return hipGetLastError();
  }
  template <typename T, bool kObserve>
  hipError_t byFilter() const
  {
    if (s.filter == kFilterNearest) return run<T, kObserve, kFilterNearest>();
    if (s.filter == kFilterCubic) return run<T, kObserve, kFilterCubic>();
    return run<T, kObserve, kFilterLinear>();
  }
};
}  // namespace

hipError_t launchGatherSample(const SamplerView &s, const AttrView &a, uint32_t attr, const uint32_t *idx,
                              const uint32_t *count, uint64_t maxCount, const float *xyz, float *out,
                              hipStream_t stream)
{
  if (maxCount == 0) return hipSuccess;
  const GatherLaunch g{s, a, attr, idx, count, maxCount, xyz, out, stream};
  const bool observed = s.grid.leafAccess != nullptr;
  if (a.type == kTypeHalf) return observed ? g.byFilter<__half, true>() : g.byFilter<__half, false>();
  return observed ? g.byFilter<float, true>() : g.byFilter<float, false>();
}

// ---- per-leaf error bounds ---------------------------------------------------------------------

__global__ void kNfLeafPoints(const float *__restrict__ i2o, const int32_t *__restrict__ leafOrigins,
                              uint64_t numLeaves, uint32_t perLeaf, uint64_t seed, float *__restrict__ xyz)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numLeaves * perLeaf) return;
  const uint64_t leaf = i / perLeaf;
  const uint64_t h0 = splitmix(seed ^ (i * 0x632BE59BD9B4E019ull)), h1 = splitmix(h0), h2 = splitmix(h1);
  const F3 p = {leafOrigins[3 * leaf] + 8.f * unit24(h0), leafOrigins[3 * leaf + 1] + 8.f * unit24(h1),
                leafOrigins[3 * leaf + 2] + 8.f * unit24(h2)};
  const F3 o = xfmPoint(i2o, p);
  xyz[3 * i]     = o.x;
  xyz[3 * i + 1] = o.y;
  xyz[3 * i + 2] = o.z;
}

hipError_t launchNfLeafPoints(const float *i2o, const int32_t *leafOrigins, uint64_t numLeaves, uint32_t perLeaf,
                              uint64_t seed, float *xyz, hipStream_t stream)
{
  const uint64_t n = numLeaves * perLeaf;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(kNfLeafPoints, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, i2o, leafOrigins,
                     numLeaves, perLeaf, seed, xyz);
  return hipGetLastError();
}

__global__ void kNfLeafError(const float *__restrict__ nf, const float *__restrict__ exact, uint64_t numLeaves,
                             uint32_t perLeaf, float *__restrict__ leafErr)
{
  const uint64_t leaf = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (leaf >= numLeaves) return;
  float e = 0.f;
  for (uint32_t k = 0; k < perLeaf; k++) {
    const float a = nf[leaf * perLeaf + k], b = exact[leaf * perLeaf + k];
    const float d = fabsf(a - b);
    e             = (isnan(b) || isnan(a)) ? vinf() : fmaxf(e, d);
  }
  leafErr[leaf] = e;
}

hipError_t launchNfLeafError(const GridView &, const float *nf, const float *exact, uint64_t numLeaves,
                             uint32_t perLeaf, float *leafErr, hipStream_t stream)
{
  if (numLeaves == 0) return hipSuccess;
  hipLaunchKernelGGL(kNfLeafError, dim3((unsigned)((numLeaves + 255) / 256)), dim3(256), 0, stream, nf, exact,
                     numLeaves, perLeaf, leafErr);
  return hipGetLastError();
}

// A leaf is approved when its measured error times `safety` is within the
// tolerance, or (useBand) when its measured error is finite and the width of
// its value range -- which bounds |clamped network - exact| everywhere in the
// leaf -- is within the tolerance.
__global__ void kNfApprove(GridView g, const int32_t *__restrict__ leafOrigins, const float *__restrict__ leafErr,
                           uint64_t numLeaves, uint32_t attr, float tolerance, float safety, int useBand,
                           uint8_t *__restrict__ leafOk, uint32_t *__restrict__ numOk)
{
  const uint64_t leaf = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (leaf >= numLeaves) return;
  const float e = leafErr[leaf];
  bool ok       = e * safety <= tolerance;
  if (!ok && useBand && e < vinf()) {
    uint64_t vox = 0;
    if (cellLeaf(g, leafOrigins[3 * leaf], leafOrigins[3 * leaf + 1], leafOrigins[3 * leaf + 2], &vox) >= 0) {
      const Range1f r = g.levels[2].valueRange[vox * g.numAttributes + attr];
      ok              = r.upper - r.lower <= tolerance;
    }
  }
  leafOk[leaf] = ok ? 1 : 0;
  if (ok) atomicAdd(numOk, 1u);
}

hipError_t launchNfApprove(const GridView &g, const int32_t *leafOrigins, const float *leafErr, uint64_t numLeaves,
                           uint32_t attr, float tolerance, float safety, int useBand, uint8_t *leafOk,
                           uint32_t *numOk, hipStream_t stream)
{
  if (numLeaves == 0) return hipSuccess;
  hipLaunchKernelGGL(kNfApprove, dim3((unsigned)((numLeaves + 255) / 256)), dim3(256), 0, stream, g, leafOrigins,
                     leafErr, numLeaves, attr, tolerance, safety, useBand, leafOk, numOk);
  return hipGetLastError();
}

}  // namespace vklhip
