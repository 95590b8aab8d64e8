/* Extensions of the HIP (MI355X) device. None of these has a counterpart in
 * the reference API; each one exists because per-ray host calls and
 * synchronous calls cannot feed a GPU. Results equal those of the reference
 * calls they batch (see each declaration). */
#pragma once

#include <stdint.h>
#include "device/openvkl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* vklComputeSampleN without the host synchronisation: device pointers only,
 * enqueued on `stream` (a hipStream_t; NULL = the null stream). Usable under
 * hipGraph capture. Same results as vklComputeSampleN. */
OPENVKL_INTERFACE void vklComputeSampleNAsync(const VKLSampler *sampler, unsigned int N,
                                              const vkl_vec3f *objectCoordinates, float *samples,
                                              unsigned int attributeIndex, void *stream);

/* vklComputeGradientN without the host synchronisation (device pointers). */
OPENVKL_INTERFACE void vklComputeGradientNAsync(const VKLSampler *sampler, unsigned int N,
                                                const vkl_vec3f *objectCoordinates, vkl_vec3f *gradients,
                                                unsigned int attributeIndex, void *stream);

/* vklComputeSampleN that allocates nothing. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and synchronous behaviour as vklComputeSampleN, but
 * objectCoordinates, samples and times must be device memory, and the
 * scratch of the binned sampling kernels comes from a fixed 64 MiB array per
 * GPU (held until the process ends) instead of an allocation. A call whose
 * binned scratch would exceed 64 MiB fails with VKL_INVALID_ARGUMENT. All
 * calls on one GPU share that array, which keeps the last call's contents:
 * calls that overlapped on one GPU would race on it, so the library
 * serialises them (one lock per GPU; each call returns after its kernels
 * have finished). */
OPENVKL_INTERFACE void vklComputeSampleNNoAlloc(const VKLSampler *sampler, unsigned int N,
                                                const vkl_vec3f *objectCoordinates, float *samples,
                                                unsigned int attributeIndex, const float *times);

/* Fixed-step ray marching fused with sampling (a consumer of the sampler, as
 * the reference's ray-march renderer, examples/interactive/renderer/
 * ray_march_iterator/RayMarchIterator.ispc:39-137, but without host round
 * trips): integrals[r] = sum over s < numSteps of sample(origins[r] +
 * t_s directions[r]) with t_s = t0 + (s + 0.5) (t1 - t0) / numSteps,
 * [t0, t1] = tRanges[r], binary32 without contraction, summed in step order.
 * Pointers may be host or device memory. */
OPENVKL_INTERFACE void vklComputeRayMarchN(const VKLSampler *sampler, unsigned int numRays,
                                           const vkl_vec3f *origins, const vkl_vec3f *directions,
                                           const vkl_range1f *tRanges, unsigned int numSteps, float *integrals,
                                           unsigned int attributeIndex);

/* Woodcock (delta) tracking fused with sampling: the free-flight sampler of
 * the reference density path tracer (examples/interactive/renderer/
 * density_path_tracer/DensityPathTracer.ispc:61-110, sampleWoodcock) for
 * numRays rays in one launch. Ray r starts at t = tRanges[r].lower with TEA
 * state (rngStates[2r], rngStates[2r+1]) (the reference seeds it with
 * (frame, pixel index), Random.ih:32-38). Each tentative collision draws two
 * TEA pairs (u0, u1), (u2, -), advances t by -ln(1 - u0) / sigmaMax, samples
 * attribute attributeIndex at origins[r] + t directions[r] at time
 * clamp(time + (u2 - 0.5) shutter, 0, 1), and is accepted when
 * u1 < sigmaMax * opacity(sample) / sigmaMax, opacity being the w channel of
 * the transfer function colorsAndOpacities (numTfValues (r, g, b, a)
 * entries over tfValueRange, TransferFunction.ih:17-47).
 * Writes eventT[r] = t of the accepted collision, +inf when t passes
 * tRanges[r].upper first (transmittance 1) and NaN when maxCollisions
 * tentative collisions were all rejected; eventSamples[r] = the sample at the
 * accepted collision (0 otherwise); numCollisions[r] (may be NULL) = samples
 * taken; rngStates updated in place. sigmaMax <= 0 gives +inf without drawing.
 * Binary32 without contraction; ln is the library's own atanh-series
 * logarithm, so the event sequence is reproducible bit for bit. Pointers may
 * be host or device memory. */
OPENVKL_INTERFACE void vklComputeWoodcockN(const VKLSampler *sampler, unsigned int numRays,
                                           const vkl_vec3f *origins, const vkl_vec3f *directions,
                                           const vkl_range1f *tRanges, unsigned int *rngStates, float sigmaMax,
                                           const float *colorsAndOpacities, unsigned int numTfValues,
                                           vkl_range1f tfValueRange, float time, float shutter,
                                           unsigned int maxCollisions, unsigned int attributeIndex, float *eventT,
                                           float *eventSamples, unsigned int *numCollisions);

/* Interval-driven ray marching with transfer-function compositing, the
 * reference's ray-march iterator renderer (examples/interactive/renderer/
 * ray_march_iterator/RayMarchIterator.ispc:39-137) for numRays rays in one
 * launch: the intervals of the context (as vklIterateInterval returns them)
 * are split into sub-intervals of nominalDeltaT / samplingRate; each is
 * sampled at its midpoint (at `time`), mapped through the transfer function
 * colorsAndOpacities (numTfValues (r, g, b, a) entries over tfValueRange,
 * TransferFunction.ih:17-47), and composited front to back with opacity
 * clamp(a * dt, 0, 1) until the accumulated opacity reaches 0.99.
 * rgba[4r..4r+3] = colour and accumulated opacity of ray r (the renderer
 * writes opacity 1); numSamples[r] (may be NULL) = samples taken, at most
 * maxSamplesPerRay (which bounds rays whose sub-intervals are tiny).
 * tRanges == NULL means [0, inf). Binary32 without contraction, in the
 * renderer's operation order. Pointers may be host or device memory. */
OPENVKL_INTERFACE void vklComputeRayMarchIntervalsN(const VKLIntervalIteratorContext *context, unsigned int numRays,
                                                    const vkl_vec3f *origins, const vkl_vec3f *directions,
                                                    const vkl_range1f *tRanges, float time, float samplingRate,
                                                    const float *colorsAndOpacities, unsigned int numTfValues,
                                                    vkl_range1f tfValueRange, unsigned int maxSamplesPerRay,
                                                    float *rgba, unsigned int *numSamples);

/* The reference's hit-iterator renderer (examples/interactive/renderer/
 * hit_iterator_renderer/HitIteratorRenderer.ispc:51-150) for numRays rays in
 * one launch. The hits of the context's iso values (as vklIterateHit returns
 * them) are shaded with the gradient normal (sampler's gradient filter, at
 * `time`): ambient 0.1 plus two lights, normalize(1, 1, 1) with emission 1
 * and normalize(1, 1, -1) with 0.8, each Lambertian when it lies on the
 * viewer's side of the surface and a shadow hit query from the hit point over
 * [epsilon, inf) finds nothing (unshaded 1 without a normal). The albedo is
 * the transfer function's colour at the hit sample; hits are composited with
 * surface opacity 0.6 until the opacity reaches 0.99. rgba[4r..4r+3] =
 * colour and accumulated opacity (the renderer writes opacity 1);
 * numHits[r] (may be NULL) = hits shaded, at most maxHitsPerRay.
 * normalize(v) = v (1 / sqrt(v . v)). tRanges == NULL means [0, inf).
 * Pointers may be host or device memory. */
OPENVKL_INTERFACE void vklComputeHitRenderN(const VKLHitIteratorContext *context, unsigned int numRays,
                                            const vkl_vec3f *origins, const vkl_vec3f *directions,
                                            const vkl_range1f *tRanges, float time, const float *colorsAndOpacities,
                                            unsigned int numTfValues, vkl_range1f tfValueRange,
                                            unsigned int maxHitsPerRay, float *rgba, unsigned int *numHits);

/* Iterate the intervals of numRays rays in one launch. Ray r writes up to
 * maxIntervalsPerRay intervals to intervals[r * maxIntervalsPerRay + k] and
 * their count to numIntervals[r]; the sequence equals calling
 * vklInitIntervalIterator + vklIterateInterval until it returns 0 (truncated
 * at maxIntervalsPerRay). tRanges == NULL means [0, inf) for every ray.
 * Pointers may be host or device memory. */
OPENVKL_INTERFACE void vklIterateIntervalsBatch(const VKLIntervalIteratorContext *context, unsigned int numRays,
                                                const vkl_vec3f *origins, const vkl_vec3f *directions,
                                                const vkl_range1f *tRanges, const float *times,
                                                VKLInterval *intervals, unsigned int maxIntervalsPerRay,
                                                unsigned int *numIntervals);

/* Hit iteration for numRays rays in one launch, same contract as above with
 * vklInitHitIterator + vklIterateHit. */
OPENVKL_INTERFACE void vklIterateHitsBatch(const VKLHitIteratorContext *context, unsigned int numRays,
                                           const vkl_vec3f *origins, const vkl_vec3f *directions,
                                           const vkl_range1f *tRanges, const float *times, VKLHit *hits,
                                           unsigned int maxHitsPerRay, unsigned int *numHits);

/* Read-only view of a committed volume's node table level (0..2): node
 * count, voxel codes (numNodes * res^3, res = 64/32/16) and value ranges
 * (numNodes * res^3 * numAttributes). The same data the reference exposes
 * through the "InnerNode" observer. Returns 1 on success. */
OPENVKL_INTERFACE int vklHipGetVolumeLevel(VKLVolume volume, unsigned int level, uint64_t *numNodes,
                                           const uint64_t **voxels, const vkl_range1f **valueRanges);

/* ---- batched multi-volume sampling --------------------------------------- */

typedef struct VKLSampleBatchPlan_ *VKLSampleBatchPlan;

/* Plan the sampling of numBatches (sampler, points) pairs:
 * samples[i][k] = sample of attribute attributeIndices[i] (NULL = 0) of
 * samplers[i] at objectCoordinates[i][k], k < counts[i]. Point and sample
 * arrays must be device memory and stay valid while the plan is used; the
 * samplers are retained. Each group of batches (see
 * vklGetSampleBatchPlanLaunchCount) is one persistent kernel over its
 * flattened point space; the launches are captured in one hipGraph.
 * autotune: 0 = fixed shape, batches grouped per voxel source type and
 * filter; 1 = shape and grouping chosen by timing candidates once per
 * workload signature (cached, and persisted to the file
 * OPENVKL_HIP_TUNE_CACHE names); 2 = fixed shape, structuredRegular batches
 * of a filter merged into one group. Results equal one vklComputeSampleN per
 * batch. */
OPENVKL_INTERFACE VKLSampleBatchPlan vklNewSampleBatchPlan(unsigned int numBatches, const VKLSampler *samplers,
                                                           const unsigned int *attributeIndices,
                                                           const unsigned int *counts,
                                                           const vkl_vec3f *const *objectCoordinates,
                                                           float *const *samples, int autotune);
/* Replay the plan's graph on `stream` (hipStream_t; NULL = null stream). */
OPENVKL_INTERFACE void vklExecuteSampleBatchPlan(VKLSampleBatchPlan plan, void *stream);
/* The launch shape the plan uses. */
OPENVKL_INTERFACE void vklGetSampleBatchPlanLaunch(VKLSampleBatchPlan plan, unsigned int *blocks,
                                                   unsigned int *threads, unsigned int *tile);
/* Kernel launches one execution of the plan makes: one per group of batches.
 * Batches are grouped per voxel source type and filter, or, where the
 * autotuner measured it faster, with all structuredRegular volumes of 32-bit
 * voxel offsets and one filter in one group whatever their element types;
 * each tricubic batch on a volume with cubicRecords / cubicStencils is a
 * launch of its own (the stencil-cooperative kernel of vklComputeSampleN). */
OPENVKL_INTERFACE unsigned int vklGetSampleBatchPlanLaunchCount(VKLSampleBatchPlan plan);
OPENVKL_INTERFACE void vklReleaseSampleBatchPlan(VKLSampleBatchPlan plan);

/* ---- ray batch plans ---------------------------------------------------------
 * Interval and hit iteration of numBatches (iterator context, rays) pairs in
 * one graph replay: batch i iterates numRays[i] rays of contexts[i] (an
 * interval or a hit iterator context; the batches may use different volumes)
 * with the contract of vklIterateIntervalsBatch / vklIterateHitsBatch:
 * results[i] receives VKLInterval or VKLHit records, maxPerRay[i] per ray,
 * counts[i] the number per ray; tRanges (or tRanges[i]) NULL means [0, inf),
 * times (or times[i]) NULL means time 0 (hit contexts). Every array must be
 * device memory and stay valid while the plan is used. The batches run on
 * parallel branches of one hipGraph (at most 16, batches spread round-robin);
 * the plan keeps the context state it was built with and retains the
 * contexts. */
typedef struct VKLRayBatchPlan_ *VKLRayBatchPlan;
OPENVKL_INTERFACE VKLRayBatchPlan vklNewRayBatchPlan(unsigned int numBatches, const VKLObject *contexts,
                                                     const unsigned int *numRays, const vkl_vec3f *const *origins,
                                                     const vkl_vec3f *const *directions,
                                                     const vkl_range1f *const *tRanges, const float *const *times,
                                                     void *const *results, const unsigned int *maxPerRay,
                                                     unsigned int *const *counts);
/* Replay the plan's graph on `stream` (hipStream_t; NULL = null stream). */
OPENVKL_INTERFACE void vklExecuteRayBatchPlan(VKLRayBatchPlan plan, void *stream);
/* Kernel launches one execution makes (one per non-empty batch). */
OPENVKL_INTERFACE unsigned int vklGetRayBatchPlanLaunchCount(VKLRayBatchPlan plan);
OPENVKL_INTERFACE void vklReleaseRayBatchPlan(VKLRayBatchPlan plan);

/* ---- neural-field sampler ---------------------------------------------------
 * A multiresolution hash-grid (16 levels x 2 binary16 features, 2^19 entries
 * per level) + 32-64-64-1 ReLU MLP that stands in for a vdb volume's exact
 * sampler. Trained on the GPU (binary16 MFMA forward / backward, binary32
 * accumulation, Adam) on samples of the exact sampler; per-leaf error bounds
 * decide where it may answer, everything else falls back to the exact
 * kernel. */

typedef struct VKLNeuralField_ *VKLNeuralField;

/* Create an untrained field for `exactSampler` (a vdb sampler; its filter is
 * what the field learns). maxResolution: finest grid resolution over the
 * volume's bounding cube (0 = 2048). */
OPENVKL_INTERFACE VKLNeuralField vklNewNeuralField(const VKLSampler *exactSampler, unsigned int attributeIndex,
                                                   unsigned int maxResolution, unsigned int seed);
OPENVKL_INTERFACE void vklReleaseNeuralField(VKLNeuralField field);
/* `steps` Adam steps of `batchSize` fresh exact samples each (a fraction
 * leafFraction drawn inside leaves, the rest in the bounding box). Writes the
 * per-step MSE to lossHistory (host, may be NULL); returns the last one. */
OPENVKL_INTERFACE float vklNeuralFieldTrain(VKLNeuralField field, unsigned int steps, unsigned int batchSize,
                                            float learningRate, float leafFraction, unsigned int seed,
                                            float *lossHistory);
/* Per-leaf error bounds: a leaf is approved when the maximum |field - exact|
 * over pointsPerLeaf random points in it, times `safety`, is <= tolerance.
 * For nearest / linear samplers the field's output in an approved leaf is
 * clamped to the leaf's value range (which holds every exact sample taken in
 * the leaf), so a leaf whose range is no wider than the tolerance is approved
 * with a guaranteed bound. Returns the number of approved leaves. */
OPENVKL_INTERFACE unsigned int vklNeuralFieldFinalize(VKLNeuralField field, float tolerance,
                                                      unsigned int pointsPerLeaf, float safety, unsigned int seed);
/* Sample N object-space points (device pointers). useFallback != 0: points
 * outside approved leaves are sampled by the exact kernel (count written to
 * *numFallback). workspace: NULL or >= 4 * N + 4 bytes of device memory. */
OPENVKL_INTERFACE void vklNeuralFieldSampleN(VKLNeuralField field, unsigned int N, const vkl_vec3f *objectCoordinates,
                                             float *samples, int useFallback, unsigned int *numFallback,
                                             void *workspace);
/* One forward + backward pass (host pointers) without an optimizer step:
 * summed gradients of sum((y - t)^2); returns the loss sum. */
OPENVKL_INTERFACE float vklNeuralFieldGradients(VKLNeuralField field, unsigned int N,
                                                const vkl_vec3f *objectCoordinates, const float *targets,
                                                float *gTable, float *gW1, float *gW2, float *gW3);
/* Host copies of the binary16 parameters and the encoding constants. */
OPENVKL_INTERFACE void vklNeuralFieldGetParams(VKLNeuralField field, uint16_t *table, uint16_t *W1, uint16_t *W2,
                                               float *W3, unsigned int *resolutions, float *boxLo, float *invExtent);
OPENVKL_INTERFACE unsigned int vklNeuralFieldNumApprovedLeaves(VKLNeuralField field, unsigned int *numLeaves);

#ifdef __cplusplus
}
#endif
