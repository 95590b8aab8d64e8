/* Error codes; values fixed by openvkl/include/openvkl/VKLError.h. */
#pragma once
#include <stdint.h>

typedef enum
#if __cplusplus >= 201103L
    : uint32_t
#endif
{
  VKL_NO_ERROR = 0,
  VKL_UNKNOWN_ERROR = 1,
  VKL_INVALID_ARGUMENT = 2,
  VKL_INVALID_OPERATION = 3,
  VKL_OUT_OF_MEMORY = 4,
  VKL_UNSUPPORTED_CPU = 5,
} VKLError;
