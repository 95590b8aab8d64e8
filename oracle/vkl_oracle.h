/* CPU parity oracle for the Open VKL sampling / traversal hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a single-threaded (optionally
 * multi-threaded for baseline timing) C++ restatement of the reference
 * algorithms; it is the checker that tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py call. The shipped library never links or calls
 * it.
 *
 * Floating-point convention: every arithmetic expression is evaluated in
 * IEEE binary32 with no FMA contraction (built with -ffp-contract=off), in the
 * operation order of the reference source.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleGrid OracleGrid;

/* structuredRegular (DenseVdbVolume). data[a] holds dims.x*dims.y*dims.z
 * elements of dtypes[a] (VKLDataType code), x fastest. i2o: 3x4 row-major
 * index->object affine (rotation-shear-scale | translation). background may be
 * NULL (= VKL_BACKGROUND_UNDEFINED for every attribute). */
OracleGrid *oracle_dense_new(const int32_t dims[3],
                             const int32_t indexOrigin[3],
                             const float i2o[12],
                             int cellCentered,
                             uint32_t numAttributes,
                             const uint32_t *dtypes,
                             const void *const *data,
                             const float *background);

/* structuredRegular with temporal data: temporalFormat 0 constant, 1
 * structured (data[a] holds numTimesteps values per voxel, time fastest), 2
 * unstructured (tuIndices: voxels + 1 entries, tuTimes: numTimes floats,
 * data[a] holds tuIndices[voxels] values). */
OracleGrid *oracle_dense_new_temporal(const int32_t dims[3],
                                      const int32_t indexOrigin[3],
                                      const float i2o[12],
                                      int cellCentered,
                                      uint32_t numAttributes,
                                      const uint32_t *dtypes,
                                      const void *const *data,
                                      const float *background,
                                      uint32_t temporalFormat,
                                      uint32_t numTimesteps,
                                      const uint64_t *tuIndices,
                                      const float *tuTimes,
                                      uint64_t numTimes);

/* vdb volume. Leaves: level, origin (index space, 3 ints each), format.
 * packed != 0: packedDense[a] holds (#dense leaves)*512 elements and
 * packedTile[a] (#tile leaves) elements (may be NULL when there are no tiles).
 * packed == 0: leafData[leaf*numAttributes + a] points at 512 (dense) or 1
 * (tile) elements. dtypes[a] is VKL_FLOAT or VKL_HALF. clip (6 ints) may be
 * NULL; it is the "indexClippingBounds" parameter. */
OracleGrid *oracle_vdb_new(uint64_t numLeaves,
                           const uint32_t *level,
                           const int32_t *origin,
                           const uint32_t *format,
                           uint32_t numAttributes,
                           const uint32_t *dtypes,
                           int packed,
                           const void *const *packedDense,
                           const void *const *packedTile,
                           const void *const *leafData,
                           const float i2o[12],
                           const float *background,
                           const int32_t *clip);

/* vdb with temporal leaves (leafData only): per leaf its temporal format
 * (0 constant, 1 structured, 2 unstructured), structured time steps,
 * unstructured index array (voxels + 1 uint64) and time array (leafTimeCounts
 * floats); leafData[leaf * numAttributes + a] holds the leaf's samples. */
OracleGrid *oracle_vdb_new_temporal(uint64_t numLeaves,
                                    const uint32_t *level,
                                    const int32_t *origin,
                                    const uint32_t *format,
                                    uint32_t numAttributes,
                                    const uint32_t *dtypes,
                                    int packed,
                                    const void *const *packedDense,
                                    const void *const *packedTile,
                                    const void *const *leafData,
                                    const float i2o[12],
                                    const float *background,
                                    const int32_t *clip,
                                    const uint32_t *leafTF,
                                    const uint32_t *leafSteps,
                                    const uint64_t *const *leafIdx,
                                    const float *const *leafTimes,
                                    const uint64_t *leafTimeCounts);

void oracle_free(OracleGrid *g);
const char *oracle_last_error(void);

void oracle_set_sampler(OracleGrid *g, int filter, int gradientFilter, int maxSamplingDepth);
/* Evaluate the tricubic filter (same matrix and order) in binary64 instead of
 * the reference's binary32; used to bound the binary32 form's rounding. */
void oracle_set_cubic_double(OracleGrid *g, int on);

void oracle_sample(const OracleGrid *g, uint64_t n, const float *xyz, float *out, uint32_t attr);
/* Per-point times (NULL: 0) for temporal volumes; single-threaded. */
void oracle_sample_times(const OracleGrid *g, uint64_t n, const float *xyz, const float *times, float *out,
                         uint32_t attr);
void oracle_gradient_times(const OracleGrid *g, uint64_t n, const float *xyz, const float *times, float *out3,
                           uint32_t attr);
void oracle_sample_mt(const OracleGrid *g, uint64_t n, const float *xyz, float *out, uint32_t attr, int threads);
void oracle_sample_m(const OracleGrid *g, uint64_t n, const float *xyz, float *out, uint32_t m, const uint32_t *attrs);
void oracle_gradient(const OracleGrid *g, uint64_t n, const float *xyz, float *out3, uint32_t attr);

void oracle_bounding_box(const OracleGrid *g, float out[6]);
void oracle_value_range(const OracleGrid *g, uint32_t attr, float out[2]);
uint32_t oracle_num_attributes(const OracleGrid *g);
void oracle_grid_info(const OracleGrid *g, int32_t rootOrigin[3], uint32_t activeSize[3], float domainBox[6], float o2i[12]);

/* Node table of level l (0..2): numNodes, then origin (3 u32 per node),
 * voxel codes (numNodes * res^3), value ranges (2 floats per voxel per attr). */
uint64_t oracle_level_num_nodes(const OracleGrid *g, int l);
void oracle_level_copy(const OracleGrid *g, int l, uint32_t *origin, uint64_t *voxels, float *ranges);

/* LeafNodeAccess: with buffer non-NULL (numLeaves entries), sampling sets
 * buffer[leaf index] = 1 for every traversal that ends on a leaf pointer. */
void oracle_set_leaf_access(OracleGrid *g, uint32_t *buffer);
/* InnerNode observer at maxDepth: returns the node count and writes up to
 * capacity (6 + 2 * numAttributes)-float tuples to out (may be NULL). */
uint64_t oracle_inner_nodes(const OracleGrid *g, int maxDepth, float *out, uint64_t capacity);

/* intervalResolutionHint -> iterator depth (IteratorContext.cpp:49-174). */
int oracle_map_depth(float hint);

/* Run an interval iterator to completion (or maxOut) for one ray. out holds
 * 5 floats per interval: tRange.lower, tRange.upper, valueRange.lower,
 * valueRange.upper, nominalDeltaT. ranges: numRanges (lower, upper) pairs. */
uint32_t oracle_iterate_intervals(const OracleGrid *g, uint32_t attr, int depth,
                                  int numRanges, const float *ranges,
                                  const float org[3], const float dir[3], const float tRange[2],
                                  float *out, uint32_t maxOut);

/* Run a hit iterator to completion (or maxOut) for one ray. out holds 3
 * floats per hit: t, sample, epsilon. depth < 0 selects the reference default
 * (3 for sparse vdb, 2 for structuredRegular). */
uint32_t oracle_iterate_hits(const OracleGrid *g, uint32_t attr, int depth,
                             int numValues, const float *values,
                             const float org[3], const float dir[3], const float tRange[2], float time,
                             float *out, uint32_t maxOut);
void oracle_iterate_many(const OracleGrid *g, uint32_t attr, int depth, int hits, int numValues,
                         const float *values, uint64_t numRays, const float *org, const float *dir,
                         uint32_t maxPerRay, uint32_t threads, uint32_t *counts, float *out);

/* The Lekien-Marsden 64x64 tricubic matrix (row = coefficient index
 * 16i+4j+k, column = constraint index), derived at load time. */
void oracle_tricubic_matrix(int32_t *out64x64);

float oracle_half_to_float(uint16_t h);

/* Woodcock (delta) tracking, density path tracer sampleWoodcock
 * (DensityPathTracer.ispc:61-110): per ray, TEA state rng[2r..2r+1] (updated),
 * tf = tfSize (r, g, b, opacity) entries over [tfLo, tfHi]. eventT: accepted
 * collision t, +inf past tRange upper, NaN after maxCollisions rejections;
 * samples: value at the accepted collision (0 otherwise); count (may be
 * NULL): samples taken. */
void oracle_woodcock(const OracleGrid *g, uint64_t numRays, const float *org, const float *dir, const float *tRange,
                     uint32_t *rng, float sigmaMax, const float *tf, uint32_t tfSize, float tfLo, float tfHi,
                     float time, float shutter, uint32_t maxCollisions, uint32_t attr, float *eventT,
                     float *samples, uint32_t *count);
/* TEA block cipher, `rounds` rounds over v[0..1] with key k[0..3]. */
void oracle_tea(uint32_t *v, const uint32_t *k, uint32_t rounds);
/* The device library's binary32 natural logarithm (atanh series). */
float oracle_flight_log(float x);

/* Interval ray marching with transfer-function compositing, the ray-march
 * iterator renderer (RayMarchIterator.ispc:39-137): intervals as
 * oracle_iterate_intervals (depth, value ranges), sub-intervals of
 * nominalDeltaT / samplingRate sampled at their midpoints at `time`, front to
 * back until opacity 0.99; rgba 4 floats per ray; count (may be NULL). */
void oracle_ray_march_intervals(const OracleGrid *g, uint32_t attr, int depth, int numRanges, const float *ranges,
                                uint64_t numRays, const float *org, const float *dir, const float *tRanges,
                                float time, float samplingRate, const float *tf, uint32_t tfSize, float tfLo,
                                float tfHi, uint32_t maxSamples, float *rgba, uint32_t *count);

/* Hit-iterator renderer (HitIteratorRenderer.ispc:51-150): hits of the iso
 * values (as oracle_iterate_hits with the default depth), gradient normal,
 * ambient 0.1 + two lights with shadow hit queries, transfer-function albedo,
 * surface opacity 0.6 until opacity 0.99, at most maxHits hits; rgba 4 floats
 * per ray; count (may be NULL) = hits shaded. */
void oracle_hit_render(const OracleGrid *g, uint32_t attr, int numValues, const float *values, uint64_t numRays,
                       const float *org, const float *dir, const float *tRanges, float time, const float *tf,
                       uint32_t tfSize, float tfLo, float tfHi, uint32_t maxHits, float *rgba, uint32_t *count);

#ifdef __cplusplus
}
#endif
