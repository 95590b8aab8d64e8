/* Default background: the quiet NaN with bits 0xFFC068B5
 * (openvkl/include/openvkl/VKLBackgroundUndefined.h:30). */
#pragma once
#include <stdint.h>
#include <string.h>

static inline float vkl_floatbits(uint32_t bits)
{
  float f;
  memcpy(&f, &bits, sizeof(f));
  return f;
}

#define VKL_BACKGROUND_UNDEFINED vkl_floatbits(0xFFC068B5u)
