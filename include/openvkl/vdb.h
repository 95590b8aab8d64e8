/* VDB tree topology used by the "vdb" and "structuredRegular" volumes.
 * The reference generates these constants at build time from
 * VKL_VDB_LOG_RESOLUTION_{0..3} = 6/5/4/3 (openvkl/CMakeLists.txt:37-49,
 * cmake/openvkl_macros.cmake:239-323); the HIP device fixes the same values.
 * Level 0 is the root, level 3 the leaf level; RES_l is the edge length of a
 * level-l node in leaf voxels. */
#pragma once

#define VKL_VDB_NUM_LEVELS 4

#define VKL_VDB_LOG_RES_0 6
#define VKL_VDB_LOG_RES_1 5
#define VKL_VDB_LOG_RES_2 4
#define VKL_VDB_LOG_RES_3 3

#define VKL_VDB_TOTAL_LOG_RES_0 18
#define VKL_VDB_TOTAL_LOG_RES_1 12
#define VKL_VDB_TOTAL_LOG_RES_2 7
#define VKL_VDB_TOTAL_LOG_RES_3 3
#define VKL_VDB_TOTAL_LOG_RES_4 0

#define VKL_VDB_RES_0 262144
#define VKL_VDB_RES_1 4096
#define VKL_VDB_RES_2 128
#define VKL_VDB_RES_3 8
#define VKL_VDB_RES_4 1

#define VKL_VDB_LEAF_LEVEL 3
#define VKL_VDB_RES_LEAF 8
#define VKL_VDB_LOG_RES_LEAF 3
#define VKL_VDB_NUM_VOXELS_LEAF 512
