/* Device object API; declarations match openvkl/include/openvkl/device.h:15-53.
 * The HIP device registers under the name "hip" (vklNewDevice("hip")). */
#pragma once

#include "VKLError.h"
#include "common.h"

struct VKLDeviceInternal;
typedef struct VKLDeviceInternal *VKLDevice;

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*VKLLogCallback)(void *userData, const char *message);
typedef void (*VKLErrorCallback)(void *userData, VKLError error, const char *message);

/* device.h:15 */
OPENVKL_INTERFACE VKLDevice vklNewDevice(const char *deviceName);
/* device.h:18 */
OPENVKL_INTERFACE void vklDeviceSetLogCallback(VKLDevice device, VKLLogCallback callback, void *userData);
/* device.h:25 */
OPENVKL_INTERFACE void vklDeviceSetErrorCallback(VKLDevice device, VKLErrorCallback callback, void *userData);
/* device.h:29 */
OPENVKL_INTERFACE void vklDeviceSetInt(VKLDevice device, const char *name, int x);
/* device.h:33 */
OPENVKL_INTERFACE void vklDeviceSetString(VKLDevice device, const char *name, const char *s);
/* device.h:37 */
OPENVKL_INTERFACE void vklDeviceSetVoidPtr(VKLDevice device, const char *name, const void *a);
/* device.h:41 */
OPENVKL_INTERFACE void vklCommitDevice(VKLDevice device);
/* device.h:43 */
OPENVKL_INTERFACE VKLError vklDeviceGetLastErrorCode(VKLDevice device);
/* device.h:45 */
OPENVKL_INTERFACE const char *vklDeviceGetLastErrorMsg(VKLDevice device);
/* device.h:47 -- the HIP device reports the wavefront width (64). */
OPENVKL_INTERFACE int vklGetNativeSIMDWidth(VKLDevice device);
/* device.h:49 */
OPENVKL_INTERFACE void vklCommit(VKLObject object);
/* device.h:51 */
OPENVKL_INTERFACE void vklRelease(VKLObject object);
/* device.h:53 */
OPENVKL_INTERFACE void vklReleaseDevice(VKLDevice device);

#ifdef __cplusplus
}
#endif
