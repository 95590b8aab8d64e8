/* Reconstruction filters; values fixed by openvkl/include/openvkl/VKLFilter.h:11-24. */
#pragma once
#include "types.h"

enum VKLFilter
#if __cplusplus >= 201103L
    : vkl_uint32
#endif
{
  VKL_FILTER_NEAREST = 0, /* one voxel, piecewise constant */
  VKL_FILTER_LINEAR = 100, /* 2x2x2 trilinear */
  VKL_FILTER_CUBIC = 200, /* 4x4x4 tricubic (central-difference constraints) */
};
#ifndef __cplusplus
typedef enum VKLFilter VKLFilter;
#endif
