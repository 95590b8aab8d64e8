/* Volume objects; declarations match openvkl/include/openvkl/volume.h:53-75.
 * Volume types on the HIP device: "structuredRegular" and "vdb". */
#pragma once

#include "common.h"
#include "device.h"

#ifdef __cplusplus
struct VKLVolume : public VKLObject
{
};
extern "C" {
#else
typedef VKLObject VKLVolume;
#endif

/* volume.h:54 */
OPENVKL_INTERFACE VKLVolume vklNewVolume(VKLDevice device, const char *type);
/* volume.h:58 */
OPENVKL_INTERFACE vkl_box3f vklGetBoundingBox(VKLVolume volume);
/* volume.h:60 */
OPENVKL_INTERFACE unsigned int vklGetNumAttributes(VKLVolume volume);
/* volume.h:62 */
OPENVKL_INTERFACE vkl_range1f vklGetValueRange(VKLVolume volume, unsigned int attributeIndex VKL_DEFAULT_VAL(= 0));
/* volume.h:68 */
OPENVKL_INTERFACE void vklGetBoundingBoxRef(const VKLVolume *volume, vkl_box3f *boundingBox);
/* volume.h:71 */
OPENVKL_INTERFACE unsigned int vklGetNumAttributesRef(const VKLVolume *volume);
/* volume.h:73 */
OPENVKL_INTERFACE void vklGetValueRangeRef(const VKLVolume *volume, unsigned int attributeIndex, vkl_range1f *valueRange);

#ifdef __cplusplus
}
#endif
