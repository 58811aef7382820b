/*
 * kfx_types.h — plain-old-data types shared by the C ABI, the HIP kernels,
 * the host drop-in layer and the CPU oracle.
 *
 * Every struct here is byte-for-byte the layout of the reference type it
 * stands for, so a caller of the reference host API can hand its buffers and
 * parameter blocks to this library unchanged.  Layouts are pinned with
 * static asserts at the bottom of the file.
 *
 * Reference layouts (dbolshak/3deye):
 *   point_3f / point_3i      common/inc/elementary_points.h:35-48 (3 x 4 B unions)
 *   point_2us                common/inc/elementary_points.h:64-69 (2 x unsigned int)
 *   matrix_3f                common/inc/types/elementary_matrix_template.h:1-3
 *                            (3 row vectors; multiply() is row . vector)
 *   intrinsics               common/inc/types/intrinsics.h:8-26 (24 B)
 *   round_buffer             common/inc/types/round_buffer.h:25-63 (72 B)
 *   Tsdf {short value, weight} common/inc/types/base_points.h:75-85 (4 B)
 *   Color {uchar r,g,b,w}    common/inc/types/base_points.h:87-99 (4 B)
 *   XYZ (Point_XYZ / Point_NXYZ, float x3)  base_points.h:172-186 (12 B)
 *   XYZI (float4 on the device)             base_points.h:135-158 (16 B)
 *   Depth {unsigned short}   base_points.h:17-31 (2 B, millimetres)
 */
#ifndef KFX_TYPES_H
#define KFX_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kfx_point_3f { float x, y, z; } kfx_point_3f;
typedef struct kfx_point_3i { int32_t x, y, z; } kfx_point_3i;
typedef struct kfx_point_2us { uint32_t x, y; } kfx_point_2us;

/* Row-major 3x3: data[r] is row r.  v' = M v means v'.r = dot(data[r], v). */
typedef struct kfx_matrix_3f { kfx_point_3f data[3]; } kfx_matrix_3f;

/* Pinhole camera.  Level l of the pyramid divides fx, fy, cx, cy by 2^l and
 * shifts cols, rows right by l (common/src/intrinsics.cpp:7-18). */
typedef struct kfx_intrinsics {
    float fx, fy, cx, cy;
    int32_t cols, rows;
} kfx_intrinsics;

/* Cyclic ("round") TSDF buffer state, passed by value to every volume kernel.
 * A local grid coordinate (x, y, z) lives at the linear voxel address
 *   ((x+ox)%Nx) + Nx*((y+oy)%Ny) + Nx*Ny*((z+oz)%Nz) + start_index
 * wrapped modulo buffer_length when it exceeds it (raycast.cl:29-47). */
typedef struct kfx_round_buffer {
    int32_t start_index;
    int32_t end_index;
    int32_t buffer_length;
    kfx_point_3i origin_GRID;        /* cyclic offset, each axis in [0, N) */
    kfx_point_3f origin_GRID_global; /* unwrapped grid offset (voxels, as float) */
    kfx_point_3f origin_metric;      /* world position of local voxel (0,0,0), metres */
    kfx_point_3f volume_size;        /* metres per axis */
    kfx_point_3i voxels_size;        /* voxels per axis */
} kfx_round_buffer;

/* Short-quantised TSDF voxel: value = clamp(rtz(tsdf * 32767), +-32767),
 * weight in [0, 128] (raycast.cl:9-27, common_defines.h:6-10). */
typedef struct kfx_tsdf { int16_t value; int16_t weight; } kfx_tsdf;

typedef struct kfx_color { uint8_t r, g, b, w; } kfx_color;

/* Vertex / normal map element (array of structures, NaN = invalid). */
typedef struct kfx_xyz { float x, y, z; } kfx_xyz;

/* Point of an extracted slice: local grid coordinates as floats, w = 1. */
typedef struct kfx_xyzi { float x, y, z, w; } kfx_xyzi;

/* Record of an extracted slice point: its voxel's global grid coordinates
 * (local coordinates plus the cube's global voxel offset, origin_GRID_global,
 * so records of successive shifts share one grid; world position = the
 * volume's world origin at the first frame + (g + 1/2) * cell) and the
 * voxel's stored tsdf value and weight. */
typedef struct kfx_slice_record { int32_t x, y, z; int16_t tsdf, weight; } kfx_slice_record;

/* Constants of the reference (common/inc/common_defines.h). */
#define KFX_MAX_DEPTH_VALUE 32767          /* TSDF quantisation divisor */
#define KFX_MAX_TSDF_WEIGHT 128            /* MAX_TSDF_WEIGHT (1<<7) */
#define KFX_DEFAULT_SIGMA_SPACE 4.5f       /* bilateral, pixels */
#define KFX_DEFAULT_SIGMA_COLOR 30.0f      /* bilateral, millimetres */
#define KFX_DEFAULT_TRUNCATION 0.03f       /* metres */
#define KFX_DEFAULT_DIST_THRESHOLD 0.1f    /* ICP correspondence distance, metres */
#define KFX_DEFAULT_SHIFT_THRESHOLD 1.5f   /* cyclic shift trigger, metres */
#define KFX_DEFAULT_COLOR_MAX_WEIGHT 1     /* DEFAULT_MAX_WEIGHT, colour volume */
#define KFX_ICP_LEVELS 3
#define KFX_ICP_SUMS 27                    /* upper triangle of [A | b], 6x7 */

#ifdef __cplusplus
} /* extern "C" */

static_assert(sizeof(kfx_point_3f) == 12, "point_3f layout");
static_assert(sizeof(kfx_point_3i) == 12, "point_3i layout");
static_assert(sizeof(kfx_point_2us) == 8, "point_2us layout");
static_assert(sizeof(kfx_matrix_3f) == 36, "matrix_3f layout");
static_assert(sizeof(kfx_intrinsics) == 24, "intrinsics layout");
static_assert(sizeof(kfx_round_buffer) == 72, "round_buffer layout");
static_assert(offsetof(kfx_round_buffer, origin_GRID) == 12, "round_buffer.origin_GRID");
static_assert(offsetof(kfx_round_buffer, voxels_size) == 60, "round_buffer.voxels_size");
static_assert(sizeof(kfx_tsdf) == 4, "Tsdf layout");
static_assert(sizeof(kfx_color) == 4, "Color layout");
static_assert(sizeof(kfx_xyz) == 12, "XYZ layout");
static_assert(sizeof(kfx_xyzi) == 16, "XYZI layout");
static_assert(sizeof(kfx_slice_record) == 16, "slice record layout");
#endif

#endif /* KFX_TYPES_H */
