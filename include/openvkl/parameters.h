/* Typed parameter setters; declarations match openvkl/include/openvkl/parameters.h:12-32. */
#pragma once

#include "VKLDataType.h"
#include "common.h"
#include "data.h"

#ifdef __cplusplus
extern "C" {
#endif

/* parameters.h:12 */
OPENVKL_INTERFACE void vklSetBool(VKLObject object, const char *name, int b);
/* parameters.h:13 */
OPENVKL_INTERFACE void vklSetFloat(VKLObject object, const char *name, float x);
/* parameters.h:14 */
OPENVKL_INTERFACE void vklSetVec3f(VKLObject object, const char *name, float x, float y, float z);
/* parameters.h:16 */
OPENVKL_INTERFACE void vklSetInt(VKLObject object, const char *name, int x);
/* parameters.h:17 */
OPENVKL_INTERFACE void vklSetVec3i(VKLObject object, const char *name, int x, int y, int z);
/* parameters.h:19 */
OPENVKL_INTERFACE void vklSetData(VKLObject object, const char *name, VKLData data);
/* parameters.h:22 */
OPENVKL_INTERFACE void vklSetString(VKLObject object, const char *name, const char *s);
/* parameters.h:25 */
OPENVKL_INTERFACE void vklSetVoidPtr(VKLObject object, const char *name, void *v);
/* parameters.h:29 -- `mem` points at one element of `dataType`. */
OPENVKL_INTERFACE void vklSetParam(VKLObject object, const char *name, VKLDataType dataType, const void *mem);

#ifdef __cplusplus
}
#endif
