/* VDB leaf formats; values fixed by openvkl/include/openvkl/VKLFormat.h:13-25. */
#pragma once
#include "types.h"

enum VKLFormat
#if __cplusplus >= 201103L
    : vkl_uint32
#endif
{
  VKL_FORMAT_TILE = 0,      /* one value for the whole node */
  VKL_FORMAT_DENSE_ZYX,     /* dense voxels, z varies fastest */
  VKL_FORMAT_INVALID = 100
};
#ifndef __cplusplus
typedef enum VKLFormat VKLFormat;
#endif
