/* Temporal formats; values fixed by openvkl/include/openvkl/VKLTemporalFormat.h.
 * On the HIP device structuredRegular volumes and vdb nodes given through
 * node.data take all three formats. */
#pragma once
#include "types.h"

enum VKLTemporalFormat
#if __cplusplus >= 201103L
    : vkl_uint32
#endif
{
  VKL_TEMPORAL_FORMAT_CONSTANT = 0,
  VKL_TEMPORAL_FORMAT_STRUCTURED,
  VKL_TEMPORAL_FORMAT_UNSTRUCTURED,
  VKL_TEMPORAL_FORMAT_INVALID = 100
};
#ifndef __cplusplus
typedef enum VKLTemporalFormat VKLTemporalFormat;
#endif
