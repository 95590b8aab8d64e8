/* Hot-path calls of the HIP device.  Every declaration below replaces the
 * same-named one in openvkl/devices/cpu/include/openvkl/device/openvkl.h
 * (line numbers cited per call) and keeps its signature, argument meaning and
 * default-return error behaviour (NaN sample, null iterator, 0 size, 0 =
 * "no more intervals/hits").
 *
 * Pointer arguments (coordinates, outputs, times) may be host or device
 * memory; host pointers are staged through device buffers. All calls are
 * synchronous, as in the reference. */
#pragma once

#include <stddef.h>
#include "../common.h"
#include "../iterator.h"
#include "../sampler.h"

/* Upper bounds of vklGet*IteratorSize*() on this device (bytes). */
#define VKL_MAX_INTERVAL_ITERATOR_SIZE 512
#define VKL_MAX_INTERVAL_ITERATOR_SIZE_4 (4 * VKL_MAX_INTERVAL_ITERATOR_SIZE)
#define VKL_MAX_INTERVAL_ITERATOR_SIZE_8 (8 * VKL_MAX_INTERVAL_ITERATOR_SIZE)
#define VKL_MAX_INTERVAL_ITERATOR_SIZE_16 (16 * VKL_MAX_INTERVAL_ITERATOR_SIZE)
#define VKL_MAX_HIT_ITERATOR_SIZE 768
#define VKL_MAX_HIT_ITERATOR_SIZE_4 (4 * VKL_MAX_HIT_ITERATOR_SIZE)
#define VKL_MAX_HIT_ITERATOR_SIZE_8 (8 * VKL_MAX_HIT_ITERATOR_SIZE)
#define VKL_MAX_HIT_ITERATOR_SIZE_16 (16 * VKL_MAX_HIT_ITERATOR_SIZE)

#ifdef __cplusplus
extern "C" {
#endif

/* openvkl.h:23 -- registers the "hip" device. */
OPENVKL_INTERFACE void vklInit(void);

/* ---- sampling ----------------------------------------------------------- */

/* openvkl.h:32 */
OPENVKL_INTERFACE float vklComputeSample(const VKLSampler *sampler, const vkl_vec3f *objectCoordinates,
                                         unsigned int attributeIndex VKL_DEFAULT_VAL(= 0), float time VKL_DEFAULT_VAL(= 0));
/* openvkl.h:38 */
OPENVKL_INTERFACE void vklComputeSample4(const int *valid, const VKLSampler *sampler, const vkl_vvec3f4 *objectCoordinates,
                                         float *samples, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0),
                                         const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:46 */
OPENVKL_INTERFACE void vklComputeSample8(const int *valid, const VKLSampler *sampler, const vkl_vvec3f8 *objectCoordinates,
                                         float *samples, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0),
                                         const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:54 */
OPENVKL_INTERFACE void vklComputeSample16(const int *valid, const VKLSampler *sampler, const vkl_vvec3f16 *objectCoordinates,
                                          float *samples, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0),
                                          const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:62 -- the stream call; the HIP device's main entry point. */
OPENVKL_INTERFACE void vklComputeSampleN(const VKLSampler *sampler, unsigned int N, const vkl_vec3f *objectCoordinates,
                                         float *samples, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0),
                                         const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:72 */
OPENVKL_INTERFACE void vklComputeSampleM(const VKLSampler *sampler, const vkl_vec3f *objectCoordinates, float *samples,
                                         unsigned int M, const unsigned int *attributeIndices, float time VKL_DEFAULT_VAL(= 0));
/* openvkl.h:80 */
OPENVKL_INTERFACE void vklComputeSampleM4(const int *valid, const VKLSampler *sampler, const vkl_vvec3f4 *objectCoordinates,
                                          float *samples, unsigned int M, const unsigned int *attributeIndices,
                                          const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:89 */
OPENVKL_INTERFACE void vklComputeSampleM8(const int *valid, const VKLSampler *sampler, const vkl_vvec3f8 *objectCoordinates,
                                          float *samples, unsigned int M, const unsigned int *attributeIndices,
                                          const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:98 */
OPENVKL_INTERFACE void vklComputeSampleM16(const int *valid, const VKLSampler *sampler, const vkl_vvec3f16 *objectCoordinates,
                                           float *samples, unsigned int M, const unsigned int *attributeIndices,
                                           const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:107 -- output layout samples[i * M + a]. */
OPENVKL_INTERFACE void vklComputeSampleMN(const VKLSampler *sampler, unsigned int N, const vkl_vec3f *objectCoordinates,
                                          float *samples, unsigned int M, const unsigned int *attributeIndices,
                                          const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:116 */
OPENVKL_INTERFACE vkl_vec3f vklComputeGradient(const VKLSampler *sampler, const vkl_vec3f *objectCoordinates,
                                               unsigned int attributeIndex VKL_DEFAULT_VAL(= 0), float time VKL_DEFAULT_VAL(= 0));
/* openvkl.h:122 */
OPENVKL_INTERFACE void vklComputeGradient4(const int *valid, const VKLSampler *sampler, const vkl_vvec3f4 *objectCoordinates,
                                           vkl_vvec3f4 *gradients, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0),
                                           const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:130 */
OPENVKL_INTERFACE void vklComputeGradient8(const int *valid, const VKLSampler *sampler, const vkl_vvec3f8 *objectCoordinates,
                                           vkl_vvec3f8 *gradients, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0),
                                           const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:138 */
OPENVKL_INTERFACE void vklComputeGradient16(const int *valid, const VKLSampler *sampler, const vkl_vvec3f16 *objectCoordinates,
                                            vkl_vvec3f16 *gradients, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0),
                                            const float *times VKL_DEFAULT_VAL(= nullptr));
/* openvkl.h:146 */
OPENVKL_INTERFACE void vklComputeGradientN(const VKLSampler *sampler, unsigned int N, const vkl_vec3f *objectCoordinates,
                                           vkl_vec3f *gradients, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0),
                                           const float *times VKL_DEFAULT_VAL(= nullptr));

/* ---- interval iteration -------------------------------------------------- */

struct IntervalIterator;
typedef struct IntervalIterator *VKLIntervalIterator;
struct IntervalIterator4;
typedef struct IntervalIterator4 *VKLIntervalIterator4;
struct IntervalIterator8;
typedef struct IntervalIterator8 *VKLIntervalIterator8;
struct IntervalIterator16;
typedef struct IntervalIterator16 *VKLIntervalIterator16;

/* Layout of openvkl.h:175-180. */
typedef struct
{
  vkl_range1f tRange;
  vkl_range1f valueRange;
  float nominalDeltaT;
} VKLInterval;

typedef struct VKL_ALIGN(16) { vkl_vrange1f4 tRange; vkl_vrange1f4 valueRange; float nominalDeltaT[4]; } VKLInterval4;
typedef struct VKL_ALIGN(32) { vkl_vrange1f8 tRange; vkl_vrange1f8 valueRange; float nominalDeltaT[8]; } VKLInterval8;
typedef struct VKL_ALIGN(64) { vkl_vrange1f16 tRange; vkl_vrange1f16 valueRange; float nominalDeltaT[16]; } VKLInterval16;

/* openvkl.h:208-217 */
OPENVKL_INTERFACE size_t vklGetIntervalIteratorSize(const VKLIntervalIteratorContext *context);
OPENVKL_INTERFACE size_t vklGetIntervalIteratorSize4(const VKLIntervalIteratorContext *context);
OPENVKL_INTERFACE size_t vklGetIntervalIteratorSize8(const VKLIntervalIteratorContext *context);
OPENVKL_INTERFACE size_t vklGetIntervalIteratorSize16(const VKLIntervalIteratorContext *context);

/* openvkl.h:239 -- `buffer` is caller-owned, at least vklGetIntervalIteratorSize() bytes. */
OPENVKL_INTERFACE VKLIntervalIterator vklInitIntervalIterator(const VKLIntervalIteratorContext *context, const vkl_vec3f *origin,
                                                              const vkl_vec3f *direction, const vkl_range1f *tRange, float time,
                                                              void *buffer);
/* openvkl.h:248 */
OPENVKL_INTERFACE VKLIntervalIterator4 vklInitIntervalIterator4(const int *valid, const VKLIntervalIteratorContext *context,
                                                                const vkl_vvec3f4 *origin, const vkl_vvec3f4 *direction,
                                                                const vkl_vrange1f4 *tRange, const float *times, void *buffer);
/* openvkl.h:258 */
OPENVKL_INTERFACE VKLIntervalIterator8 vklInitIntervalIterator8(const int *valid, const VKLIntervalIteratorContext *context,
                                                                const vkl_vvec3f8 *origin, const vkl_vvec3f8 *direction,
                                                                const vkl_vrange1f8 *tRange, const float *times, void *buffer);
/* openvkl.h:268 */
OPENVKL_INTERFACE VKLIntervalIterator16 vklInitIntervalIterator16(const int *valid, const VKLIntervalIteratorContext *context,
                                                                  const vkl_vvec3f16 *origin, const vkl_vvec3f16 *direction,
                                                                  const vkl_vrange1f16 *tRange, const float *times, void *buffer);
/* openvkl.h:279 */
OPENVKL_INTERFACE int vklIterateInterval(VKLIntervalIterator iterator, VKLInterval *interval);
/* openvkl.h:282 */
OPENVKL_INTERFACE void vklIterateInterval4(const int *valid, VKLIntervalIterator4 iterator, VKLInterval4 *interval, int *result);
/* openvkl.h:288 */
OPENVKL_INTERFACE void vklIterateInterval8(const int *valid, VKLIntervalIterator8 iterator, VKLInterval8 *interval, int *result);
/* openvkl.h:294 */
OPENVKL_INTERFACE void vklIterateInterval16(const int *valid, VKLIntervalIterator16 iterator, VKLInterval16 *interval, int *result);

/* ---- hit iteration --------------------------------------------------------- */

struct HitIterator;
typedef struct HitIterator *VKLHitIterator;
struct HitIterator4;
typedef struct HitIterator4 *VKLHitIterator4;
struct HitIterator8;
typedef struct HitIterator8 *VKLHitIterator8;
struct HitIterator16;
typedef struct HitIterator16 *VKLHitIterator16;

/* Layout of openvkl.h:315-320. */
typedef struct
{
  float t;
  float sample;
  float epsilon;
} VKLHit;

typedef struct VKL_ALIGN(16) { float t[4]; float sample[4]; float epsilon[4]; } VKLHit4;
typedef struct VKL_ALIGN(32) { float t[8]; float sample[8]; float epsilon[8]; } VKLHit8;
typedef struct VKL_ALIGN(64) { float t[16]; float sample[16]; float epsilon[16]; } VKLHit16;

/* openvkl.h:348-357 */
OPENVKL_INTERFACE size_t vklGetHitIteratorSize(const VKLHitIteratorContext *context);
OPENVKL_INTERFACE size_t vklGetHitIteratorSize4(const VKLHitIteratorContext *context);
OPENVKL_INTERFACE size_t vklGetHitIteratorSize8(const VKLHitIteratorContext *context);
OPENVKL_INTERFACE size_t vklGetHitIteratorSize16(const VKLHitIteratorContext *context);
/* openvkl.h:380 */
OPENVKL_INTERFACE VKLHitIterator vklInitHitIterator(const VKLHitIteratorContext *context, const vkl_vec3f *origin,
                                                    const vkl_vec3f *direction, const vkl_range1f *tRange, float time,
                                                    void *buffer);
/* openvkl.h:388 */
OPENVKL_INTERFACE VKLHitIterator4 vklInitHitIterator4(const int *valid, const VKLHitIteratorContext *context,
                                                      const vkl_vvec3f4 *origin, const vkl_vvec3f4 *direction,
                                                      const vkl_vrange1f4 *tRange, const float *times, void *buffer);
/* openvkl.h:397 */
OPENVKL_INTERFACE VKLHitIterator8 vklInitHitIterator8(const int *valid, const VKLHitIteratorContext *context,
                                                      const vkl_vvec3f8 *origin, const vkl_vvec3f8 *direction,
                                                      const vkl_vrange1f8 *tRange, const float *times, void *buffer);
/* openvkl.h:406 */
OPENVKL_INTERFACE VKLHitIterator16 vklInitHitIterator16(const int *valid, const VKLHitIteratorContext *context,
                                                        const vkl_vvec3f16 *origin, const vkl_vvec3f16 *direction,
                                                        const vkl_vrange1f16 *tRange, const float *times, void *buffer);
/* openvkl.h:416 */
OPENVKL_INTERFACE int vklIterateHit(VKLHitIterator iterator, VKLHit *hit);
/* openvkl.h:419 */
OPENVKL_INTERFACE void vklIterateHit4(const int *valid, VKLHitIterator4 iterator, VKLHit4 *hit, int *result);
/* openvkl.h:425 */
OPENVKL_INTERFACE void vklIterateHit8(const int *valid, VKLHitIterator8 iterator, VKLHit8 *hit, int *result);
/* openvkl.h:431 */
OPENVKL_INTERFACE void vklIterateHit16(const int *valid, VKLHitIterator16 iterator, VKLHit16 *hit, int *result);

#ifdef __cplusplus
}
#endif
