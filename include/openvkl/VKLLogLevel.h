/* Log levels; values fixed by openvkl/include/openvkl/VKLLogLevel.h. */
#pragma once
#include <stdint.h>

typedef enum
#if __cplusplus >= 201103L
    : uint32_t
#endif
{
  VKL_LOG_DEBUG = 1,
  VKL_LOG_INFO = 2,
  VKL_LOG_WARNING = 3,
  VKL_LOG_ERROR = 4,
  VKL_LOG_NONE = 5,
} VKLLogLevel;
