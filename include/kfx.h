/*
 * kfx.h — C ABI of the MI355X-native KinectFusion per-frame pipeline.
 *
 * Plain pointers and sizes only.  Every function returns a status code
 * (KFX_OK == 0); kfx_last_error() holds the message of the most recent
 * failure on the calling thread.  Device pointers are HIP device memory of
 * the GPU selected by kfx_init().  Work is enqueued on the library stream and
 * is asynchronous unless the comment says otherwise; kfx_synchronize() waits
 * for it.
 *
 * Each entry point names the reference interface it replaces
 * (dbolshak/3deye, common/inc/opencl_wrappers.h and common/src/...).
 */
#ifndef KFX_H
#define KFX_H

#include "kfx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Rigid camera transform in double precision: p_out = R p_cam + t (R row-major). */
typedef struct kfx_pose { double R[9]; double t[3]; } kfx_pose;

enum kfx_status {
    KFX_OK = 0,
    KFX_ERR_INVALID_ARG = 1,
    KFX_ERR_HIP = 2,
    KFX_ERR_CAPACITY = 3,       /* output buffer too small; required size reported */
    KFX_ERR_NOT_INITIALISED = 4,
    KFX_ERR_SINGULAR = 5,       /* ICP normal equations not positive definite */
    KFX_ERR_UNSUPPORTED = 6
};

/* ------------------------------------------------------------------------
 * Runtime.  Replaces the OpenCL singleton and buffer layer:
 * opencl_helper::Get_Instance / init_opencl (common/src/opencl_helper.cpp:49-72,169-174),
 * gpu_array_base::dev_alloc / dev_free / memcpy_* (common/inc/gpu_array_base.h:17-123,
 * common/src/opencl_helper.cpp:351-462,524-534).
 * ------------------------------------------------------------------------ */

/* Select the GPU and create the library stream.  Idempotent per device. */
int kfx_init(int device);
/* Message of the last failure on this thread ("" if none). */
const char* kfx_last_error(void);
/* Number of visible GPUs. */
int kfx_device_count(int* count);
/* Wait for all work enqueued on the library stream. */
int kfx_synchronize(void);
/* Diagnostics: the number of positive floats with exponent in [exp_lo,
 * exp_hi] (within [-126, 127]) whose device reciprocal of the kernels'
 * division-free form (v_rcp_f32 and one Newton step) differs from the IEEE
 * quotient; 0 over [-126, 125] on MI355X.  Blocking.  No reference
 * counterpart. */
int kfx_check_reciprocal(int exp_lo, int exp_hi, uint64_t* mismatches);
/* The library stream as a hipStream_t (for callers that interleave work). */
int kfx_get_stream(void** hip_stream);

int kfx_malloc(void** dev_ptr, size_t bytes);
int kfx_free(void* dev_ptr);
/* Blocking copies, as the reference's CL_TRUE transfers. */
int kfx_memcpy_htod(void* dev_dst, const void* host_src, size_t bytes);
int kfx_memcpy_dtoh(void* host_dst, const void* dev_src, size_t bytes);
/* Asynchronous device-to-device copy and byte fill on the library stream. */
int kfx_memcpy_dtod(void* dev_dst, const void* dev_src, size_t bytes);
int kfx_fill_u8(void* dev_dst, uint8_t value, size_t bytes);
/* Device-to-device copy by a streaming kernel (16-byte aligned pointers, size
 * a multiple of 16): the achievable-bandwidth reference for the roofline and
 * the calibration of the HBM traffic counters. */
int kfx_copy_kernel(void* dev_dst, const void* dev_src, size_t bytes);

/* Events on the library stream, for timing (replaces the OpenCL event timing
 * of opencl_helper::execute_kernel, common/src/opencl_helper.cpp:675-685). */
int kfx_event_create(void** event);
int kfx_event_destroy(void* event);
int kfx_event_record(void* event);
int kfx_event_elapsed_ms(void* start, void* stop, float* ms);


/* Per-kernel timing on the library stream (replaces the per-kernel OpenCL
 * event printout of opencl_helper::execute_kernel, opencl_helper.cpp:675-685).
 * While enabled, the library brackets the launches of its named kernels with
 * HIP events on its own stream; kfx_profile_query sums their durations.
 * Names: "integrate", "scale_depth", "raycast", ... */
int kfx_profile_enable(int on);
int kfx_profile_reset(void);
int kfx_profile_query(const char* kernel, double* total_ms, uint64_t* launches);

/* ------------------------------------------------------------------------
 * Cyclic TSDF volume.
 * ------------------------------------------------------------------------ */

/* Linear voxel address of local grid coordinates (x, y, z), computed on the
 * host with the device's formula (GPU/opencl/raycast.cl:29-47). */
uint32_t kfx_cyclic_address(int x, int y, int z, const kfx_round_buffer* rb);

/* Device-side evaluation of the same formula for n coordinate triples
 * (xyz: 3*n ints on the device, out: n uint32 on the device). */
int kfx_cyclic_address_batch(const int32_t* xyz, uint32_t* out, size_t n,
                             const kfx_round_buffer* rb);

/* Reset every voxel to {value 0, weight 0}: the reference's init_tsdf kernel
 * (listed in common/inc/constants.h:168), for buffer_length voxels. */
int kfx_init_tsdf(kfx_tsdf* volume, const kfx_round_buffer* rb);

/* Integrate_Tsdf_Data (common/inc/opencl_wrappers.h:10-21,
 * common/src/opencl_wrappers.cpp:21-64): scale_depth (uint16 mm -> metres
 * along the pixel ray) into depth_scaled (cols*rows floats, device), then the
 * projective TSDF update of every voxel column.  rotation_inverse maps world
 * (volume) coordinates to camera coordinates; translation is the camera
 * position in volume coordinates. */
int kfx_integrate_tsdf(const uint16_t* depth, float* depth_scaled, kfx_tsdf* volume,
                       const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                       const kfx_point_3f* translation, float truncation_distance,
                       const kfx_point_3f* cell_size, const kfx_round_buffer* rb);

/* kfx_integrate_tsdf that allocates nothing: no reference counterpart; for
 * callers that cannot call hipMalloc (code under graph capture, tight
 * serving loops, harnesses that forbid allocation).  Same arguments and the
 * same result as kfx_integrate_tsdf.  Its depth-bound pyramid (the only
 * device memory the integration uses beyond its arguments) lives in a fixed
 * 4 MiB array per device, set up by kfx_init: images up to 2048 x 1536
 * (levels 1..5 of the pyramid within 2^20 floats); larger images fail with
 * KFX_ERR_INVALID_ARG.  Every call on a device shares that array, so calls
 * that overlap on one device (from several threads) race; each call writes
 * every texel it reads, so nothing carries over between calls. */
int kfx_integrate_tsdf_noalloc(const uint16_t* depth, float* depth_scaled, kfx_tsdf* volume,
                               const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                               const kfx_point_3f* translation, float truncation_distance,
                               const kfx_point_3f* cell_size, const kfx_round_buffer* rb);

/* Number of voxels the update rule of kfx_integrate_tsdf changes for this
 * depth frame and pose (in-image projection, valid depth, sdf >= -trunc).
 * The rule does not read the volume, so the count is independent of its
 * state.  Read-only pass; blocking; depth_scaled is written as in
 * kfx_integrate_tsdf. */
int kfx_integrate_count_updates(const uint16_t* depth, float* depth_scaled, const kfx_intrinsics* intr,
                                const kfx_matrix_3f* rotation_inverse, const kfx_point_3f* translation,
                                float truncation_distance, const kfx_point_3f* cell_size,
                                const kfx_round_buffer* rb, uint64_t* count);

/* kfx_integrate_tsdf with the pose read from device memory (12 floats: R
 * camera -> volume row-major, then t), e.g. the fused ICP's estimate, so the
 * call can follow ICP without a host round trip (no reference counterpart).
 * When dev_prev_pose_inv (R^T, t of the previous frame) is given and
 * motion_threshold > 0, the update is skipped if the camera moved less than
 * the threshold ((rotation angle + translation) / 2, the KinectFusion rule).
 * When dev_skip is given and *dev_skip != 0 at run time, the update is
 * skipped (a decision taken earlier on the device, e.g. a fallback gate). */
int kfx_integrate_tsdf_posed(const uint16_t* depth, float* depth_scaled, kfx_tsdf* volume,
                             const kfx_intrinsics* intr, const float* dev_pose, const float* dev_prev_pose_inv,
                             float motion_threshold, float truncation_distance, const kfx_point_3f* cell_size,
                             const kfx_round_buffer* rb, const int32_t* dev_skip);

/* Brick flags of the skipping raycast (no reference counterpart: an
 * acceleration structure of this library).  One byte per 16^3 brick of
 * local voxel coordinates, x fastest: non-zero when the brick may hold a
 * negative TSDF value.  kfx_brick_flags_size gives the bytes for a volume
 * (0 when an axis is not a multiple of 16: no skipping);
 * kfx_brick_flags_build writes the exact flags of a volume (enqueued);
 * kfx_raycast_bricks is kfx_raycast with the skipping march, whose outputs
 * equal kfx_raycast's bit for bit whenever every brick holding a negative
 * value is flagged (flags may over-state; they may not miss one).  A
 * pipeline keeps its own flags: its integrations set them and every shift
 * rebuilds them. */
int kfx_brick_flags_size(const kfx_round_buffer* rb, size_t* bytes);
int kfx_brick_flags_build(const kfx_tsdf* volume, const kfx_round_buffer* rb, uint8_t* bricks);
int kfx_raycast_bricks(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                       const kfx_round_buffer* rb, const kfx_point_3f* translation, const kfx_matrix_3f* rotation,
                       float truncation_distance, const kfx_intrinsics* intr, const kfx_point_3f* cell_size,
                       const uint8_t* bricks);

/* kfx_raycast that also returns S, the voxel reads its algorithm makes (first
 * sample, every march sample inside the volume, 8 per in-bounds trilinear
 * evaluation): the raycast's algorithmic bytes are 4 S + 24 W H.  Same
 * outputs as kfx_raycast; blocking; for measurement (no reference
 * counterpart). */
int kfx_raycast_count_samples(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                              const kfx_round_buffer* rb, const kfx_point_3f* translation,
                              const kfx_matrix_3f* rotation, float truncation_distance, const kfx_intrinsics* intr,
                              const kfx_point_3f* cell_size, uint64_t* samples);

/* Colour volume (Integrate_Colors, common/src/opencl_wrappers.cpp:181-214,
 * declared in common/inc/opencl_wrappers.h:59-69; its kernel
 * integrate_color_volume is absent from the reference).  Every voxel of the
 * cyclic colour volume whose centre g projects (camera c = R^-1 (g - t),
 * pixel rint(f c/c.z + c0), as in kfx_integrate_tsdf) onto a pixel whose
 * raycast vertex p (volume coordinates, NaN = none) lies closer than
 * truncation_distance to g takes the running average of the frame's colour:
 * rgb = round_half_even((rgb * w + rgb_frame) / (w + 1)) per channel in FP32,
 * w = min(w + 1, max_weight), max_weight in [1, 255] (the reference's
 * default is 1, common/inc/common_defines.h:26).  rgb_frame: cols x rows
 * kfx_color, the w byte ignored (common/src/Capture.cpp:115-123). */
int kfx_integrate_colors(kfx_color* color_volume, const kfx_color* rgb_frame, const kfx_xyz* vertex_map,
                         const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                         const kfx_point_3f* translation, float truncation_distance, int max_weight,
                         const kfx_point_3f* cell_size, const kfx_round_buffer* rb);
/* kfx_integrate_colors with the pose read from device memory (12 floats, R
 * camera -> volume then t), skipped when dev_skip is given and non-zero at
 * run time (no reference counterpart). */
int kfx_integrate_colors_posed(kfx_color* color_volume, const kfx_color* rgb_frame, const kfx_xyz* vertex_map,
                               const kfx_intrinsics* intr, const float* dev_pose, float truncation_distance,
                               int max_weight, const kfx_point_3f* cell_size, const kfx_round_buffer* rb,
                               const int32_t* dev_skip);
/* Zero colour volume (colour 0, weight 0: "not set" for the slice scan). */
int kfx_init_colors(kfx_color* volume, const kfx_round_buffer* rb);

/* Raycast (common/inc/opencl_wrappers.h:23-33, GPU/opencl/raycast.cl:182-405):
 * per-pixel march through the cyclic TSDF; writes the surface vertex and the
 * normal (NaN where there is none) in volume coordinates, cols*rows XYZ each. */
int kfx_raycast(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                const kfx_round_buffer* rb, const kfx_point_3f* translation,
                const kfx_matrix_3f* rotation, float truncation_distance,
                const kfx_intrinsics* intr, const kfx_point_3f* cell_size);
/* kfx_raycast with the pose read from device memory (12 floats, as above);
 * nothing is written when dev_skip is given and *dev_skip != 0. */
int kfx_raycast_posed(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                      const kfx_round_buffer* rb, const float* dev_pose, float truncation_distance,
                      const kfx_intrinsics* intr, const kfx_point_3f* cell_size, const int32_t* dev_skip);


/* ------------------------------------------------------------------------
 * Depth preprocessing.  Images are row-major; depth is uint16 millimetres;
 * vertex/normal maps are cols*rows kfx_xyz, NaN x = invalid.
 * ------------------------------------------------------------------------ */

/* Apply_Billateral_Filtration (common/inc/opencl_wrappers.h:72-77,
 * common/src/opencl_wrappers.cpp:216-235, GPU/opencl/compute_bilateral_filtering.cl:8-57):
 * 13x13 bilateral filter, output clamp(rint(sum(w d)/sum(w)), 0, 32767), zero
 * above max_distance_mm (pass 0xFFFFFFFF for none: the reference wrapper binds
 * the float truncation distance's bits to this argument, so it never fires). */
int kfx_bilateral_filter(const uint16_t* src, uint16_t* dst, int cols, int rows, float sigma_space,
                         float sigma_color, uint32_t max_distance_mm);
/* One level of Create_Pyramid (opencl_wrappers.h:79-83, opencl_wrappers.cpp:244-271):
 * dst is (src_cols/2) x (src_rows/2). */
int kfx_pyr_down(const uint16_t* src, int src_cols, int src_rows, uint16_t* dst, float sigma_color);
/* One level of Create_Vertex_Maps (opencl_wrappers.h:85-90, opencl_wrappers.cpp:281-305);
 * intr is the level's intrinsics (intrinsics::operator(), intrinsics.cpp:7-18). */
int kfx_vertex_map(const uint16_t* depth, kfx_xyz* vmap, const kfx_intrinsics* intr);
/* One level of Create_Normal_Maps (opencl_wrappers.h:92-97, opencl_wrappers.cpp:315-343). */
int kfx_normal_map(const kfx_xyz* vmap, kfx_xyz* nmap, int cols, int rows);
/* Transform_Maps (opencl_wrappers.h:49-57, opencl_wrappers.cpp:148-179,
 * GPU/opencl/transform_map.cl:6-65): v' = R v + t, n' = R n. */
int kfx_transform_maps(const kfx_xyz* vmap_src, const kfx_xyz* nmap_src, const kfx_matrix_3f* rotation,
                       const kfx_point_3f* translation, int cols, int rows, kfx_xyz* vmap_dst, kfx_xyz* nmap_dst);
/* Resize_Vertex_Map (normalize = 0) / Resize_Normal_Map (normalize = 1)
 * (opencl_wrappers.h:37-47, opencl_wrappers.cpp:97-146): 2x2 mean, NaN if any
 * of the four is invalid; out is (in_cols/2) x (in_rows/2). */
int kfx_resize_map(const kfx_xyz* in, int in_cols, int in_rows, kfx_xyz* out, int normalize);
/* Vertex and normal maps of `levels` pyramid levels (at most 3) in one launch:
 * for each level l, kfx_vertex_map(depth[l], vmap[l], &intr[l]) then
 * kfx_normal_map(vmap[l], nmap[l], ...), bit for bit (no reference
 * counterpart: one launch instead of the 2 x levels of Create_Vertex_Maps and
 * Create_Normal_Maps). */
int kfx_maps_pyramid(const uint16_t* const* depth, kfx_xyz* const* vmap, kfx_xyz* const* nmap,
                     const kfx_intrinsics* intr, int levels);
/* kfx_maps_pyramid from level-0 depth alone: level l >= 1's depth is
 * kfx_pyr_down of level l - 1 (sigma_color as there), derived inside the same
 * launch, and every level's maps equal those of kfx_pyr_down followed by
 * kfx_maps_pyramid, bit for bit.  Level l must be exactly half of level
 * l - 1 (intr[l].cols == intr[l-1].cols / 2, rows likewise).  depth_out may
 * be NULL; otherwise depth_out[l] (l >= 1; depth_out[0] is not read) receives
 * level l's depth, or is NULL to skip it (no reference counterpart: one
 * launch instead of compute_depth_pyramid's levels plus the maps). */
int kfx_maps_pyramid_from_depth(const uint16_t* depth, uint16_t* const* depth_out, kfx_xyz* const* vmap,
                                kfx_xyz* const* nmap, const kfx_intrinsics* intr, int levels, float sigma_color);
/* Levels 1 and 2 of a vertex map and a normal map from level 0 (cols x rows,
 * multiples of 4) in one launch: kfx_resize_map twice per map, bit for bit
 * (no reference counterpart). */
int kfx_resize_pyramid(const kfx_xyz* vmap0, const kfx_xyz* nmap0, int cols, int rows, kfx_xyz* vmap1,
                       kfx_xyz* nmap1, kfx_xyz* vmap2, kfx_xyz* nmap2);

/* ------------------------------------------------------------------------
 * ICP normal equations and pose update.
 * ------------------------------------------------------------------------ */

/* Estimate_Combined (common/inc/opencl_wrappers.h:105-122,
 * common/src/opencl_wrappers.cpp:364-438): projective point-to-plane
 * correspondences between the current level's maps (camera coordinates) and
 * the previous raycast maps (volume coordinates), per-32x8-block partial sums
 * of the 27 entries of [A | b] into gbuf (entry k at gbuf[k * gbuf_stride +
 * block]), the 256-lane reduction of transform_estimation.cl into mbuf (27
 * floats, device), then a blocking download and the unpacking of
 * opencl_wrappers.cpp:426-437 into A (6x6, row-major, symmetric) and b (6).
 * gbuf_stride >= DIVUP(cols,32)*DIVUP(rows,8) (the reference hard-codes 1200). */
int kfx_estimate_combined(const kfx_matrix_3f* rotation_current, const kfx_point_3f* translation_current,
                          const kfx_xyz* vmap_current, const kfx_xyz* nmap_current,
                          const kfx_matrix_3f* rotation_previous_inverse, const kfx_point_3f* translation_previous,
                          const kfx_intrinsics* intr, const kfx_xyz* vmap_previous, const kfx_xyz* nmap_previous,
                          float distance_threshold, float angle_threshold, float* gbuf, int gbuf_stride,
                          float* mbuf, double* A, double* b);

/* Clean_Buff (opencl_wrappers.h:222-224, GPU/opencl/clean_buff.cl:4-11):
 * zero gbuf[27][width]. */
int kfx_clean_buff(float* gbuf, int width);

/* Fused ICP (no reference counterpart as one call: it replaces the
 * Estimate_Combined + host solve + pose update loop of the absent fusion
 * application).  The handle owns the device-resident pose (double) and the
 * per-workgroup partial sums. */
typedef struct kfx_icp kfx_icp;
typedef struct kfx_icp_result {
    kfx_pose pose;         /* current estimate, camera -> volume */
    double A[36], b[6];    /* normal equations of the last solve */
    double x[6];           /* last increment: rx, ry, rz, tx, ty, tz */
    int32_t status;        /* KFX_OK or KFX_ERR_SINGULAR (last solve) */
    int32_t singular_count;
    int32_t iterations;    /* solves performed since kfx_icp_reset */
    int32_t converged_tag; /* level tag that met the convergence test, -1 none */
} kfx_icp_result;

/* Precision of the per-workgroup sums of the fused iteration; the
 * cross-workgroup sum, the 6x6 solve and the pose update are FP64 in every
 * mode.  FP32 Jacobian rows and products in every mode.  FP64 is the default
 * (of kfx_icp_create and kfx_pipeline_default_config): the most accurate and,
 * on gfx950, the fastest (a wave's FP64 sum tree costs one half-rate add per
 * step where a compensated FP32 pair merge costs about ten FP32 operations). */
#define KFX_ICP_SUM_FP32_KAHAN 0 /* compensated (TwoSum) FP32 sums */
#define KFX_ICP_SUM_FP32 1       /* plain FP32 sums */
#define KFX_ICP_SUM_FP64 2       /* FP64 sums (default) */

int kfx_icp_create(int max_pixels, kfx_icp** out);
int kfx_icp_destroy(kfx_icp* icp);
/* Set the estimate (asynchronous, on the library stream). */
int kfx_icp_reset(kfx_icp* icp, const kfx_pose* pose);
/* Selects KFX_ICP_SUM_* for later iterations (host-side setting). */
int kfx_icp_set_sum_mode(kfx_icp* icp, int mode);
/* How iterations are launched (host-side setting; the states are
 * bit-identical in every mode).  KFX_ICP_LAUNCH_FRAME (default):
 * kfx_icp_frame runs a frame's iterations in one launch when the device
 * holds all of that kernel's workgroups at once (kfx_icp_frame_resident),
 * and kfx_icp_iterate runs as KFX_ICP_LAUNCH_TWO.  KFX_ICP_LAUNCH_TWO: an
 * iteration is the sums launch, then the cross-workgroup sum, solve and
 * pose update as a second, one-workgroup launch.  KFX_ICP_LAUNCH_ONE: the
 * workgroup of the sums launch that finishes last does the second half
 * (faster for eager launches: 320x240 15.2 against 11.9 + 11.5 us, HIP
 * events; slower under graph replay: the 512^3 walk 1751 against 1875
 * frames/s).  In the last two modes kfx_icp_frame enqueues the frame as
 * kfx_icp_begin_frame and kfx_icp_iterate calls.  No reference
 * counterpart. */
#define KFX_ICP_LAUNCH_FRAME 0
#define KFX_ICP_LAUNCH_TWO 1
#define KFX_ICP_LAUNCH_ONE 2
int kfx_icp_set_launch_mode(kfx_icp* icp, int mode);
/* One iteration (one or two launches, kfx_icp_set_launch_mode): FP32
 * Jacobian rows and products, workgroup sums in the handle's KFX_ICP_SUM_*
 * precision (fixed order), FP64 cross-workgroup sum in one workgroup, FP64
 * 6x6 solve (Gauss-Jordan over one wave) and pose composition on the GPU.  A singular system leaves the pose
 * unchanged and records KFX_ERR_SINGULAR.  When level_tag >= 0 and an
 * iteration's increment is below (converge_rotation rad, converge_translation
 * m), later iterations with the same tag return immediately (early exit
 * without a host round trip); converge_rotation <= 0 disables the test. */
int kfx_icp_iterate(kfx_icp* icp, const kfx_xyz* vmap_current, const kfx_xyz* nmap_current,
                    const kfx_xyz* vmap_previous, const kfx_xyz* nmap_previous,
                    const kfx_matrix_3f* rotation_previous_inverse, const kfx_point_3f* translation_previous,
                    const kfx_intrinsics* intr, float distance_threshold, float angle_threshold, int level_tag,
                    double converge_rotation, double converge_translation);
/* With rotation_previous_inverse and translation_previous both NULL,
 * kfx_icp_iterate takes the previous pose from the state (set by
 * kfx_icp_begin_frame), so a frame's iterations need no host arguments. */
/* Start a frame: the current estimate becomes the previous pose; counters reset. */
int kfx_icp_begin_frame(kfx_icp* icp);
/* One pyramid level of a frame's ICP: its maps, camera and iteration count. */
typedef struct kfx_icp_level {
    const kfx_xyz* vmap_current;
    const kfx_xyz* nmap_current;
    const kfx_xyz* vmap_previous;
    const kfx_xyz* nmap_previous;
    kfx_intrinsics intr;
    int32_t iterations;
} kfx_icp_level;
/* A frame's ICP: with begin_frame, kfx_icp_begin_frame first; then
 * levels[l].iterations iterations at each level l from KFX_ICP_LEVELS - 1
 * down to 0, as kfx_icp_iterate with the previous pose from the state and
 * level tag l when tagged (else -1).  The state equals that sequence's bit
 * for bit.  In KFX_ICP_LAUNCH_FRAME mode on a device that holds all its
 * workgroups at once it is one launch: 256 workgroups, each the sums of
 * the two-launch form's workgroups w, w + 256, ..., joined by a grid
 * barrier per iteration, after which every workgroup solves the same
 * system; a barrier wait that runs out (workgroups not resident) makes the
 * state invalid, which kfx_icp_read and kfx_icp_read_published report as
 * KFX_ERR_HIP.  A pending kfx_icp_publish_next applies to the frame's last
 * iteration.  The reference runs the frame as Estimate_Combined calls with
 * a host solve per iteration (common/src/opencl_wrappers.cpp:364-438). */
int kfx_icp_frame(kfx_icp* icp, const kfx_icp_level levels[KFX_ICP_LEVELS], int begin_frame,
                  float distance_threshold, float angle_threshold, int tagged, double converge_rotation,
                  double converge_translation);
/* 1 when kfx_icp_frame runs as one launch on this device (KFX_ICP_LAUNCH_FRAME mode). */
int kfx_icp_frame_resident(kfx_icp* icp, int* resident);
/* One solve and pose update from 27 sums given by the caller (host memory,
 * in the unpacking order of common/src/opencl_wrappers.cpp:422-437: row k of
 * the upper triangle of [A | b]), as an iteration applies them: x from
 * A x = b, R <- R_inc(x) R, t <- R_inc(x) t + x[3..5]; a system that is not
 * positive definite leaves the pose and counts as singular.  No reference
 * counterpart: the reference hands A and b to its caller
 * (opencl_wrappers.cpp:364-438, kfx_estimate_combined here), which solves on
 * the host; this applies the GPU's solve to them.  Enqueued on the stream. */
int kfx_icp_solve_sums(kfx_icp* icp, const double sums[27]);
/* Add (dx, dy, dz) to the estimate's translation (re-expression after a
 * cyclic shift of the volume). */
int kfx_icp_translate(kfx_icp* icp, double dx, double dy, double dz);
/* Blocking read of the state. */
int kfx_icp_read(kfx_icp* icp, kfx_icp_result* out);
/* The next kfx_icp_iterate's solve also writes the resulting state to the
 * handle's pinned host copy (also when that iteration is skipped by the
 * early exit), in a graph replay too.  kfx_icp_read_published then
 * synchronises the stream and reads that copy without a copy command.  It
 * fails (KFX_ERR_INVALID_ARG) when no publishing iteration was enqueued
 * since the last kfx_icp_reset or kfx_icp_translate; a graph replay of a
 * captured publishing iteration counts as enqueued.  A publish request
 * applies to the next kfx_icp_iterate call only, also when that call fails
 * its argument checks.  No reference counterpart: it saves the per-frame
 * device-to-host copy of the fused pipeline. */
int kfx_icp_publish_next(kfx_icp* icp);
int kfx_icp_read_published(kfx_icp* icp, kfx_icp_result* out);
/* Device addresses of the single-precision current pose (R row-major camera
 * -> volume, then t: 12 floats) and of the previous pose (R^T then t), for
 * kernels that consume the estimate without a host round trip. */
int kfx_icp_pose_device(kfx_icp* icp, const float** dev_pose, const float** dev_prev_pose_inv);

/* ------------------------------------------------------------------------
 * RGBD photometric odometry (Estimate_Isometry_By_RGBD,
 * common/src/opencl_wrappers.cpp:1328-1466; kernels
 * scan_for_interest_point.cl, scan_list_of_valid_correspondence.cl,
 * compute_least_squares_equation.cl).  The reference's path is not
 * functional; this is the dense RGBD odometry it follows (OpenCV rgbd
 * module): Sobel-textured pixels of the current frame paired with the
 * previous frame by projection and a depth test, weights 1/(sigma + |diff|),
 * Sobel scale 1/8, normal equations in double precision, Cholesky solve with
 * the |det| >= 1e-6 test, estimate <- exp(twist) * estimate, levels coarsest
 * first.  Frames: RGB (kfx_color) and depth (uint16 mm) of cols x rows.
 * ------------------------------------------------------------------------ */
typedef struct kfx_rgbd kfx_rgbd;
typedef struct kfx_rgbd_params {
    float max_depth_diff;   /* m: RGBD_Depth_Threshold (0.07 in OpenCV) */
    float min_gradient;     /* Sobel magnitude of a textured pixel (10) */
    int32_t levels;         /* pyramid levels used (<= the handle's) */
    int32_t iterations[4];  /* per level, finest first (7, 7, 7, 10) */
} kfx_rgbd_params;
typedef struct kfx_rgbd_result {
    kfx_pose pose;            /* previous camera -> current camera coordinates */
    double A[36], b[6];       /* last normal equations */
    double sigma;             /* last residual RMS, intensity units */
    int32_t correspondences;  /* of the last iteration */
    int32_t solves;
    int32_t ok;               /* a solve succeeded (the reference's return value) */
    int32_t reserved;
} kfx_rgbd_result;
/* levels 1..4; cols and rows at least 8 << (levels - 1). */
int kfx_rgbd_create(int cols, int rows, int levels, kfx_rgbd** out);
int kfx_rgbd_destroy(kfx_rgbd* rgbd);
/* Intensity, metric depth and Sobel pyramids of a frame into slot 0
 * (previous) or 1 (current); enqueues only. */
int kfx_rgbd_set_frame(kfx_rgbd* rgbd, int slot, const kfx_color* rgb, const uint16_t* depth);
/* The current frame becomes the previous one (no copy). */
int kfx_rgbd_swap(kfx_rgbd* rgbd);
/* Blocking copy of a prepared level (any output may be NULL). */
int kfx_rgbd_download_level(kfx_rgbd* rgbd, int slot, int level, float* intensity, float* depth, float* dx,
                            float* dy);
/* The motion from the previous to the current frame, starting at init;
 * blocking.  out->ok == 0 when no level produced a solution. */
int kfx_rgbd_estimate(kfx_rgbd* rgbd, const kfx_intrinsics* intr, const kfx_rgbd_params* params, const kfx_pose* init,
                      kfx_rgbd_result* out);

/* ------------------------------------------------------------------------
 * Cyclic shift: black-zone clean and slice extraction.  A local voxel
 * (x, y, z) is in the black zone when, on any axis, x + origin_GRID.x lies in
 * [min_bounds.x, max_bounds.x] (inclusive, unwrapped; clean_color_slice.cl:29-40).
 * ------------------------------------------------------------------------ */

/* determine_extract_step (common/src/mem_utils.cpp:43-85, divisor 4):
 * chunk depth depth/8 (512 -> 64, 256 -> 32, 128 -> 16), extended to any
 * depth divisible by 8 (one chunk otherwise; the reference throws). */
int kfx_determine_extract_step(int depth, int* step, int* num_of_steps);
/* Clean_Tsdf_Slice (opencl_wrappers.h:180-185, opencl_wrappers.cpp:835-854):
 * reset black-zone voxels to {0, 0}. */
int kfx_clean_tsdf_slice(const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds, const kfx_round_buffer* rb,
                         kfx_tsdf* volume);
/* Clean_Color_Slice (opencl_wrappers.h:186-191, opencl_wrappers.cpp:814-833,
 * GPU/opencl/clean_color_slice.cl:41-67). */
int kfx_clean_color_slice(const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds,
                          const kfx_round_buffer* rb, kfx_color* volume);
/* Reusable scratch for slice extraction of one volume size. */
typedef struct kfx_slicer kfx_slicer;
int kfx_slicer_create(const kfx_round_buffer* rb, kfx_slicer** out);
int kfx_slicer_destroy(kfx_slicer* slicer);
/* Extract_Slice (V + C overload; opencl_wrappers.h:200-211,
 * opencl_wrappers.cpp:856-964, GPU/opencl/scan_slice_of_both_volumes.cl:76-208,
 * GPU/opencl/extract_slice_as_cloud.cl:72-202): black-zone zero crossings
 * (x/y/z neighbour, both weights != 0, F != 1, sign change; the z neighbour
 * only inside the current chunk) and set colour voxels, as XYZI points (local
 * grid coordinates of the neighbour, w = 1) with their colours, in the
 * reference's order.  color_volume may be NULL (no colour volume: plane 3 is
 * empty and colours are zero).  Blocking: *count receives the number of
 * points; if it exceeds capacity the call returns KFX_ERR_CAPACITY and
 * writes nothing. */
int kfx_extract_slice(kfx_slicer* slicer, const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds,
                      const kfx_round_buffer* rb, const kfx_tsdf* tsdf_volume, const kfx_color* color_volume,
                      kfx_xyzi* points, kfx_color* colors, uint64_t capacity, uint64_t* count);
/* The same points as kfx_extract_slice, in the same order, as records: the
 * point's voxel in global grid coordinates (local + round_buffer
 * origin_GRID_global) with its stored tsdf value and weight, the
 * (x, y, z, tsdf, weight) form of a world-model slice (no reference
 * counterpart: the reference emits grid positions with w = 1).  colors may
 * be NULL. */
int kfx_extract_slice_records(kfx_slicer* slicer, const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds,
                              const kfx_round_buffer* rb, const kfx_tsdf* tsdf_volume,
                              const kfx_color* color_volume, kfx_slice_record* records, kfx_color* colors,
                              uint64_t capacity, uint64_t* count);

/* Extract_Color_Volume (common/inc/opencl_wrappers.h:131-138,
 * common/src/opencl_wrappers.cpp:546-590; kernel extract_colors absent): for
 * each point (volume coordinates in metres, device array), the colour of the
 * voxel floor(p / cell) of the cyclic colour volume; (0,0,0,0) outside the
 * cube.  Enqueues only. */
int kfx_extract_colors(const kfx_color* color_volume, const kfx_round_buffer* rb, const kfx_xyzi* points,
                       uint64_t n, const kfx_point_3f* cell_size, kfx_color* colors);

/* ------------------------------------------------------------------------
 * Learned depth denoiser (no reference counterpart; it stands in for the
 * bilateral filter of Apply_Billateral_Filtration).  Residual network of four
 * 3x3 convolutions, 2 -> 16 -> 16 -> 16 -> 1 channels (input: depth in
 * metres and the valid mask; output: depth + correction on valid pixels,
 * invalid pixels stay 0), 5,089 fp32 master parameters (per layer weights
 * [co][ci][ky][kx] then biases), used in bfloat16 with fp32 accumulation on
 * the matrix cores.  Trained by the kernels below with Adam.
 * ------------------------------------------------------------------------ */
typedef struct kfx_denoiser kfx_denoiser;
int kfx_denoiser_param_count(void);
/* All parameters start at zero (the identity map); load trained ones with
 * kfx_denoiser_set_params. */
int kfx_denoiser_create(int cols, int rows, kfx_denoiser** out);
int kfx_denoiser_destroy(kfx_denoiser* dn);
/* Host <-> device copies of the master parameters (blocking); setting them
 * also resets the Adam state. */
int kfx_denoiser_set_params(kfx_denoiser* dn, const float* host_params);
int kfx_denoiser_get_params(kfx_denoiser* dn, float* host_params);
/* Forward pass on a device depth frame: out_depth (uint16 mm) and/or
 * out_metres (float) may be NULL. */
int kfx_denoiser_forward(kfx_denoiser* dn, const uint16_t* depth, uint16_t* out_depth, float* out_metres);
/* The same forward pass layer by layer (four launches, the hidden
 * activations through HBM, as training runs it): kfx_denoiser_forward's
 * output bit for bit, for checking the one-launch inference kernel. */
int kfx_denoiser_forward_layers(kfx_denoiser* dn, const uint16_t* depth, uint16_t* out_depth, float* out_metres);
/* Forward + backward on one (raw depth, target depth in metres) pair: adds
 * the gradient of the frame's mean squared error (pixels with a valid input,
 * a target > 0 and |input - target| < outlier_m) to the gradient buffer.
 * loss / count (optional, blocking when given) receive the frame's mean
 * squared error and pixel count. */
int kfx_denoiser_accumulate(kfx_denoiser* dn, const uint16_t* depth, const float* target_m, float outlier_m,
                            double* loss, uint64_t* count);
int kfx_denoiser_get_grad(kfx_denoiser* dn, float* host_grad);
int kfx_denoiser_zero_grad(kfx_denoiser* dn);
/* Adam step with the accumulated gradient divided by `frames`; clears it. */
int kfx_denoiser_adam_step(kfx_denoiser* dn, float lr, int frames);
/* Training target: z-depth (metres) in the camera at pose_volume of each
 * raycast vertex (volume coordinates); 0 where there is no surface. */
int kfx_model_depth(const kfx_xyz* vmap_model, const kfx_pose* pose_volume, int cols, int rows, float* out_m);
/* Fallback gate: raycast vertices (volume coordinates, previous frame)
 * projected into the camera at the device pose (12 floats, R camera ->
 * volume then t).  dev_stats (4 doubles on the device) receives {sum of the
 * squared depth differences over the pixels where the frame's depth is within
 * max_diff_m of the vertex depth, the number of those pixels, the number of
 * vertices landing on a pixel with depth, the number of valid vertices},
 * summed in a fixed order. */
int kfx_depth_consistency(const kfx_xyz* vmap_model, const float* dev_pose, const uint16_t* depth,
                          const kfx_intrinsics* intr, float max_diff_m, double* dev_stats);

/* ------------------------------------------------------------------------
 * Per-frame pipeline: the driver of the reference's absent fusion
 * application (Makefile:3,19) — preprocessing, coarse-to-fine ICP,
 * integration, raycast and cyclic shift with slice extraction.
 * ------------------------------------------------------------------------ */
enum kfx_pipeline_mode {
    KFX_PIPELINE_PLAIN = 0, /* reference control flow: sync per step, host solve per ICP iteration */
    KFX_PIPELINE_FUSED = 1  /* device-resident pose, fused ICP iterations, optional graph replay */
};
typedef struct kfx_pipeline_config {
    kfx_intrinsics intr;          /* level-0 camera; cols, rows multiples of 4 */
    kfx_point_3i voxels;          /* voxels per axis */
    kfx_point_3f volume_size;     /* metres per axis */
    float truncation;             /* 0.03 m */
    float sigma_space, sigma_color; /* bilateral: 4.5 px, 30 mm */
    float distance_threshold;     /* ICP correspondence distance, 0.1 m */
    float angle_threshold;        /* ICP normal angle, sin(20 deg) */
    int32_t icp_iterations[3];    /* level 0 (finest), 1, 2: 10, 5, 4 */
    int32_t mode;                 /* kfx_pipeline_mode */
    int32_t use_graph;            /* fused mode: replay each frame as one HIP graph */
    int32_t denoiser;             /* 1: learned denoiser instead of the bilateral filter */
    float fallback_rms_mm;        /* denoiser fallback threshold, 15 mm */
    float shift_threshold;        /* cyclic shift trigger: target point more than this from the cube centre, 1.5 m */
    float integrate_threshold;    /* integrate only after (angle + translation) / 2 >= this, 0.001 */
    double converge_rotation;     /* fused early exit per level (rad); 0 disables */
    double converge_translation;  /* fused early exit per level (m) */
    int32_t extract_on_shift;     /* extract the leaving slab before clearing it */
    int32_t icp_sum_mode;         /* fused mode: KFX_ICP_SUM_* workgroup sums (default FP64) */
    double shift_target_distance; /* target point = camera + this along the optical axis (m); 0: camera centre */
    int32_t colors;               /* 1: keep a colour volume; frames then come with RGB (kfx_pipeline_process_rgbd) */
    int32_t color_max_weight;     /* colour running-average weight cap, 1..255 (reference default 1) */
    int32_t extract_records;      /* 1: shifts extract kfx_slice_record (x, y, z, tsdf, weight); 0: XYZI points */
    int32_t icp_launch;           /* fused mode: KFX_ICP_LAUNCH_* (0: a frame's ICP in one launch where possible) */
} kfx_pipeline_config;

typedef struct kfx_frame_info {
    kfx_pose pose_world;          /* camera -> world after this frame */
    kfx_pose pose_volume;         /* camera -> volume used to integrate this frame (before any shift) */
    int32_t icp_iterations;       /* ICP solves performed */
    int32_t icp_singular;         /* solves rejected as singular */
    int32_t integrated;           /* the frame was integrated */
    int32_t shifted;              /* a cyclic shift followed this frame */
    uint64_t extracted_points;    /* points extracted by that shift */
    int32_t fallback;             /* denoiser fell back to the bilateral filter */
    float denoise_rms_mm;         /* denoiser output vs raycast model, RMS */
} kfx_frame_info;

typedef struct kfx_pipeline_view {
    kfx_tsdf* volume;
    kfx_round_buffer round_buffer;
    kfx_point_3f cell_size;
    double origin_world[3];       /* world position of local voxel (0,0,0)'s corner */
    const uint16_t* depth_raw;    /* last input frame as read (device: the caller's buffer, or the pipeline's
                                     copy with cfg.use_graph) */
    uint16_t* depth_filtered;     /* level-0 filtered depth (device) */
    kfx_xyz* vmap_model;          /* last raycast, volume coordinates (device) */
    kfx_xyz* nmap_model;
    kfx_xyz* vmap_current;        /* level-0 maps of the last frame, camera coordinates */
    kfx_xyz* nmap_current;
    kfx_pose pose_volume;         /* camera -> volume */
    int64_t frames;
    int32_t shifts;
    int32_t fallbacks;            /* frames that fell back from the denoiser to the bilateral filter */
    uint64_t extracted_points;    /* total over all shifts */
    kfx_xyzi* last_points;        /* last shift's extracted points (device; NULL with extract_records) */
    kfx_color* last_colors;
    kfx_color* color_volume;      /* colour volume (device), NULL unless cfg.colors */
    kfx_color* rgb;               /* last input colour frame (device), NULL unless cfg.colors */
    kfx_slice_record* last_records; /* last shift's records (device), NULL unless cfg.extract_records */
    uint8_t* brick_flags;         /* the skipping raycast's brick flags (device, kfx_brick_flags_size bytes),
                                     NULL in plain mode or when an axis is not a multiple of 16 */
    kfx_xyz* vmap_model_levels[KFX_ICP_LEVELS]; /* the model maps of every ICP level (device; [0] is vmap_model) */
    kfx_xyz* nmap_model_levels[KFX_ICP_LEVELS];
    float* depth_scaled;          /* the last frame's scaled depth (device, cols*rows floats) */
    float* depth_bound;           /* the integration's depth-bound pyramid (device): levels 1..5 in order, level l
                                     ceil(cols / 2^l) x ceil(rows / 2^l) texels, each the largest scaled depth
                                     of its 2^l x 2^l pixels; NULL when the workspace holds none */
} kfx_pipeline_view;

typedef struct kfx_pipeline kfx_pipeline;
/* Defaults of the 512^3, 640x480 configuration (fx = fy = 525, 3.072 m). */
int kfx_pipeline_default_config(kfx_pipeline_config* cfg);
/* volume_origin_world: world position of local voxel (0,0,0)'s corner;
 * initial_pose_world: the first frame's camera -> world pose. */
int kfx_pipeline_create(const kfx_pipeline_config* cfg, const kfx_pose* initial_pose_world,
                        const double volume_origin_world[3], kfx_pipeline** out);
int kfx_pipeline_destroy(kfx_pipeline* pipeline);
/* Process one depth frame (device pointer, cols x rows uint16 mm).  Returns
 * when the frame's pose is known on the host: in fused mode as soon as the
 * frame's ICP launch has published it (the host waits on a flag that launch
 * raises), while the frame's integration, raycast and map update may still
 * run on the library stream; every later library call on that stream is
 * ordered after them (kfx_synchronize waits for them; kfx_pipeline_state
 * does too).  The depth buffer is read on the library stream (by the
 * frame's kernels, or by a copy into the pipeline's own buffer when
 * cfg.use_graph is set): overwrite it through the library (kfx_memcpy_htod)
 * or after kfx_synchronize.  One synchronisation per step in plain mode. */
int kfx_pipeline_process(kfx_pipeline* pipeline, const uint16_t* depth, kfx_frame_info* info);
/* Process one depth frame with its registered colour frame (device pointer,
 * cols x rows kfx_color): as kfx_pipeline_process, then the colour volume
 * takes the frame's colours at the raycast of the new pose (the reference's
 * Integrate_Colors step).  Needs cfg.colors; kfx_pipeline_process then
 * fails for that pipeline. */
int kfx_pipeline_process_rgbd(kfx_pipeline* pipeline, const uint16_t* depth, const kfx_color* rgb,
                              kfx_frame_info* info);
/* The pipeline's buffers and counters, after the work enqueued so far has
 * completed (it synchronises the library stream). */
int kfx_pipeline_state(kfx_pipeline* pipeline, kfx_pipeline_view* view);
/* Attach a trained denoiser (fused mode; used when cfg.denoiser is set).  The
 * pipeline does not own it. */
int kfx_pipeline_set_denoiser(kfx_pipeline* pipeline, kfx_denoiser* denoiser);

/* ------------------------------------------------------------------------
 * Synthetic input (host only, deterministic): the seeded analytic scene,
 * camera trajectories and Kinect-like depth frames every test, benchmark and
 * training run uses.  The reference reads a Kinect through OpenNI
 * (common/src/Capture.cpp); this generator replaces the sensor.
 * World frame: y points down (gravity), the ground is the plane y = ground_y.
 * ------------------------------------------------------------------------ */
#define KFX_SYNTH_MAX_OBJECTS 64
enum kfx_synth_shape { KFX_SYNTH_SPHERE = 0, KFX_SYNTH_BOX = 1 };
typedef struct kfx_synth_object {
    int32_t shape;     /* kfx_synth_shape */
    int32_t reserved;
    double center[3];  /* metres */
    double half[3];    /* box half extents; sphere radius in half[0] */
} kfx_synth_object;
typedef struct kfx_synth_scene {
    double ground_y;
    int32_t n_objects;
    int32_t reserved;
    kfx_synth_object objects[KFX_SYNTH_MAX_OBJECTS];
} kfx_synth_scene;

uint64_t kfx_splitmix64(uint64_t x);
/* n_boxes boxes and n_spheres spheres resting on the ground, sizes and radii
 * uniform in [0.1, 0.5] m, centres uniform in region {x0, x1, z0, z1}. */
int kfx_synth_make_scene(uint64_t seed, int n_boxes, int n_spheres, const double region[4],
                         double ground_y, kfx_synth_scene* out);
/* n poses on a circle of `radius` around `center`, height offset dy (negative
 * is up), starting at angle0 and advancing dangle radians per frame, each
 * looking at `center`. */
int kfx_synth_orbit(const double center[3], double radius, double dy, double angle0, double dangle,
                    int n, kfx_pose* out);
/* Seeded smooth random walk: `step` metres per frame from `start`, heading
 * wandering within +-0.3 rad of +x, looking ahead and `pitch` radians down. */
int kfx_synth_walk(uint64_t seed, const double start[3], double step, double pitch, int n, kfx_pose* out);
/* Render one depth frame (uint16 mm, row-major cols x rows).  flags bit 0:
 * Kinect noise sigma(z) = 1.2 + 1.9 (z - 0.4)^2 mm, 1 mm quantisation, 2 %
 * dropout, seeded by splitmix64(seed, frame); without it the exact depth is
 * rounded to the millimetre.  Depth outside [0.4, 6.0] m reads 0.  clean_m
 * (optional, cols*rows floats) receives the exact z-depth in metres (0 = no
 * hit).  threads <= 0 uses every host core. */
int kfx_synth_render(const kfx_synth_scene* scene, const kfx_pose* pose, const kfx_intrinsics* intr,
                     uint64_t seed, int frame, int flags, int threads, uint16_t* depth_mm, float* clean_m);
#define KFX_SYNTH_NOISE 1
/* Colour frame of the same view (noise-free): per pixel the albedo of the
 * surface hit (objects: a fixed colour each; ground: 0.5 m checkerboard of two
 * greys; no hit: black), w = 1. */
int kfx_synth_render_rgb(const kfx_synth_scene* scene, const kfx_pose* pose, const kfx_intrinsics* intr,
                         int threads, kfx_color* rgb);

/* ------------------------------------------------------------------------
 * World-model files (host arrays in, files out).
 * ------------------------------------------------------------------------ */
/* matrix<T>::Save_To_File text format (common/inc/impl/matrix.hpp:411-476)
 * of an extracted slice as a 1-D matrix: "n 1 1 <type>", then every element
 * and a space on one line; XYZI elements as x, y, z, w joined by 'x'
 * (base_points.cpp:570-574), Color elements as "r.g.b.w " (base_points.cpp:278-282). */
int kfx_save_points_text(const char* path, const kfx_xyzi* points, uint64_t n);
int kfx_save_colors_text(const char* path, const kfx_color* colors, uint64_t n);
/* PCD v0.7 file of the cloud the reference hands to PCL
 * (convert_to_point_cloud, common/src/math_utils.cpp:216-328): x y z rgb
 * (pcl::PointXYZRGB, the reference's packing b << 16 | g << 8 | r) when
 * colors is given, else x y z intensity (pcl::PointXYZI, intensity = w);
 * binary != 0 for DATA binary, else DATA ascii. */
int kfx_save_pcd(const char* path, const kfx_xyzi* points, const kfx_color* colors, uint64_t n, int binary);
/* Binary slice records: magic "KFXREC1\0", uint64 count, then the
 * kfx_slice_record array (no reference counterpart). */
int kfx_save_records(const char* path, const kfx_slice_record* records, uint64_t n);

/* Trajectory text file in the reference's format (write_isometry_to_file,
 * common/src/string_utils.cpp:26-74): per pose its index, the three rows of
 * R and then t, each as "axbxc " in single precision. */
int kfx_trajectory_write(const char* path, const kfx_pose* poses, int n);
int kfx_trajectory_read(const char* path, kfx_pose* poses, int max_poses, int* n);
/* ATE (RMSE of positions, common first frame) and RPE between consecutive
 * frames (RMSE of relative translation error in m, mean relative rotation
 * error in rad); any output may be NULL. */
int kfx_trajectory_errors(const kfx_pose* est, const kfx_pose* gt, int n, double* ate, double* rpe_trans,
                          double* rpe_rot);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* KFX_H */
