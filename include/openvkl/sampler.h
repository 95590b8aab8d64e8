/* Sampler objects; declarations match openvkl/include/openvkl/sampler.h:30-41. */
#pragma once

#include "common.h"
#include "volume.h"

#ifdef __cplusplus
struct VKLSampler : public VKLObject
{
};
extern "C" {
#else
typedef VKLObject VKLSampler;
#endif

/* sampler.h:30 */
OPENVKL_INTERFACE VKLSampler vklNewSampler(VKLVolume volume);

typedef vkl_uint64 VKLFeatureFlags;
#define VKL_FEATURE_FLAGS_NONE 0
#define VKL_FEATURE_FLAGS_DEFAULT -1

/* sampler.h:41 */
OPENVKL_INTERFACE VKLFeatureFlags vklGetFeatureFlags(VKLSampler sampler);

#ifdef __cplusplus
}
#endif
