/* Data arrays; declarations match openvkl/include/openvkl/data.h:16-44.
 * VKL_DATA_DEFAULT copies (compacting strides); VKL_DATA_SHARED_BUFFER keeps
 * the caller's pointer, which on the HIP device may be host or device memory. */
#pragma once

#include <stddef.h>
#include "VKLDataType.h"
#include "device.h"

typedef enum
#if __cplusplus >= 201103L
    : uint32_t
#endif
{
  VKL_DATA_DEFAULT = 0,
  VKL_DATA_SHARED_BUFFER = (1 << 0),
} VKLDataCreationFlags;

#ifdef __cplusplus
struct VKLData : public VKLObject
{
};
extern "C" {
#else
typedef VKLObject VKLData;
#endif

/* data.h:38 */
OPENVKL_INTERFACE VKLData vklNewData(VKLDevice device,
                                     size_t numItems,
                                     VKLDataType dataType,
                                     const void *source,
                                     VKLDataCreationFlags dataCreationFlags VKL_DEFAULT_VAL(= VKL_DATA_DEFAULT),
                                     size_t byteStride VKL_DEFAULT_VAL(= 0));

#ifdef __cplusplus
}
#endif
