/* Shared macros and SIMD-wide structs; mirrors openvkl/include/openvkl/common.h
 * (VKL_ALIGN, OPENVKL_INTERFACE, vkl_vvec3f{4,8,16}, vkl_vrange1f{4,8,16}). */
#pragma once

#include "types.h"

#define VKL_ALIGN(...) __attribute__((aligned(__VA_ARGS__)))
#define OPENVKL_INTERFACE __attribute__((visibility("default")))

#ifdef __cplusplus
#define VKL_DEFAULT_VAL(a) a
#else
#define VKL_DEFAULT_VAL(a)
#endif

/* Structure-of-arrays layouts used by the width-4/8/16 entry points. */
#define VKL_DEFINE_WIDE_TYPES(W, ALIGN)                                   \
  typedef struct VKL_ALIGN(ALIGN) { vkl_float x[W], y[W], z[W]; }         \
      vkl_vvec3f##W;                                                      \
  typedef struct VKL_ALIGN(ALIGN) { vkl_float lower[W], upper[W]; }       \
      vkl_vrange1f##W;
VKL_DEFINE_WIDE_TYPES(4, 16)
VKL_DEFINE_WIDE_TYPES(8, 32)
VKL_DEFINE_WIDE_TYPES(16, 64)
#undef VKL_DEFINE_WIDE_TYPES

#define NOWARN_C_LINKAGE_PUSH                                   \
  _Pragma("clang diagnostic push")                              \
  _Pragma("clang diagnostic ignored \"-Wreturn-type-c-linkage\"")
#define NOWARN_C_LINKAGE_POP _Pragma("clang diagnostic pop")
