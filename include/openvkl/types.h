/* Open VKL public types for the HIP (MI355X) device.
 * Layout-compatible with openvkl/include/openvkl/types.h (handle at :56-69,
 * vector / box / affine types below it): same names, field order and sizes so
 * that applications compiled against the reference headers link unchanged. */
#pragma once

#ifdef __cplusplus
#include <cstdint>
#else
#include <stdbool.h>
#include <stdint.h>
#endif

typedef bool vkl_bool;
typedef int8_t vkl_int8;
typedef int16_t vkl_int16;
typedef int32_t vkl_int32;
typedef int64_t vkl_int64;
typedef uint8_t vkl_uint8;
typedef uint16_t vkl_uint16;
typedef uint32_t vkl_uint32;
typedef uint64_t vkl_uint64;

/* Every API object is a pair: `host` is the library-side object, `device`
 * the device-resident view kernels read (null for objects without one). */
typedef struct __VKLObject
{
#ifdef __cplusplus
  __VKLObject() : host(nullptr), device(nullptr) {}
  operator bool() const { return host != nullptr; }
#endif
  void *host;
  void *device;
} VKLObject;

#define VKL_DEFINE_VEC_TYPES(SCALAR, SUFFIX)                  \
  typedef struct { SCALAR x, y; } vkl_vec2##SUFFIX;           \
  typedef struct { SCALAR x, y, z; } vkl_vec3##SUFFIX;        \
  typedef struct { SCALAR x, y, z, w; } vkl_vec4##SUFFIX;

typedef vkl_int8 vkl_char;
typedef vkl_uint8 vkl_uchar;
typedef vkl_uint8 vkl_byte;
typedef vkl_uint8 vkl_raw;
typedef vkl_int16 vkl_short;
typedef vkl_uint16 vkl_ushort;
typedef vkl_int32 vkl_int;
typedef vkl_uint32 vkl_uint;
typedef vkl_int64 vkl_long;
typedef vkl_uint64 vkl_ulong;
typedef vkl_uint16 vkl_half; /* IEEE 754 binary16 bit pattern */
typedef float vkl_float;
typedef double vkl_double;

VKL_DEFINE_VEC_TYPES(vkl_char, c)
VKL_DEFINE_VEC_TYPES(vkl_uchar, uc)
VKL_DEFINE_VEC_TYPES(vkl_short, s)
VKL_DEFINE_VEC_TYPES(vkl_ushort, us)
VKL_DEFINE_VEC_TYPES(vkl_int, i)
VKL_DEFINE_VEC_TYPES(vkl_uint, ui)
VKL_DEFINE_VEC_TYPES(vkl_long, l)
VKL_DEFINE_VEC_TYPES(vkl_ulong, ul)
VKL_DEFINE_VEC_TYPES(vkl_half, h)
VKL_DEFINE_VEC_TYPES(vkl_float, f)
VKL_DEFINE_VEC_TYPES(vkl_double, d)
#undef VKL_DEFINE_VEC_TYPES

typedef struct { vkl_int lower, upper; } vkl_box1i;
typedef struct { vkl_vec2i lower, upper; } vkl_box2i;
typedef struct { vkl_vec3i lower, upper; } vkl_box3i;
typedef struct { vkl_vec4i lower, upper; } vkl_box4i;
typedef struct { vkl_float lower, upper; } vkl_box1f;
typedef struct { vkl_vec2f lower, upper; } vkl_box2f;
typedef struct { vkl_vec3f lower, upper; } vkl_box3f;
typedef struct { vkl_vec4f lower, upper; } vkl_box4f;
typedef vkl_box1f vkl_range1f;
typedef vkl_box2f vkl_range2f;
typedef vkl_box3f vkl_range3f;
typedef vkl_box4f vkl_range4f;

typedef struct { vkl_vec2f vx, vy; } vkl_linear2f;
typedef struct { vkl_vec3f vx, vy, vz; } vkl_linear3f;
typedef struct { vkl_linear2f l; vkl_vec2f p; } vkl_affine2f;
typedef struct { vkl_linear3f l; vkl_vec3f p; } vkl_affine3f;
