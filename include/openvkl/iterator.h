/* Iterator contexts; declarations match openvkl/include/openvkl/iterator.h:30-37.
 * Parameters: "attributeIndex" (int), "valueRanges" (VKL_BOX1F data),
 * "intervalResolutionHint" (float) for interval contexts; "values" (VKL_FLOAT
 * data) for hit contexts. */
#pragma once

#include "common.h"
#include "sampler.h"

#ifdef __cplusplus
struct VKLIntervalIteratorContext : public VKLObject
{
};
struct VKLHitIteratorContext : public VKLObject
{
};
extern "C" {
#else
typedef VKLObject VKLIntervalIteratorContext;
typedef VKLObject VKLHitIteratorContext;
#endif

/* iterator.h:31 */
OPENVKL_INTERFACE VKLIntervalIteratorContext vklNewIntervalIteratorContext(VKLSampler sampler);
/* iterator.h:37 */
OPENVKL_INTERFACE VKLHitIteratorContext vklNewHitIteratorContext(VKLSampler sampler);

#ifdef __cplusplus
}
#endif
