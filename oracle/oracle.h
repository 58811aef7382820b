/*
 * oracle.h — CPU restatement of the reference's per-frame pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the CPU baseline; the product path (libkfx) never calls it.
 *
 * Two precisions per step where the build needs both:
 *   *_f32 — single precision in the reference kernel's operation order, no
 *           floating-point contraction (built with -ffp-contract=off), so the
 *           GPU kernels can be compared bit for bit;
 *   *_f64 — the same algorithm in double precision ("truth"), used to report
 *           the error of the single-precision pipeline.
 * Loops over columns and pixels are parallelised with OpenMP when the caller
 * asks for more than one thread (orc_set_threads); results do not depend on
 * the thread count.
 *
 * Each function cites the reference file:line it follows.  Functions whose
 * reference kernel is absent from dbolshak/3deye follow the KinectFusion
 * convention the reference's call sites imply; how each one is pinned is
 * stated in its comment.
 */
#ifndef KFX_ORACLE_H
#define KFX_ORACLE_H

#include "../include/kfx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

void orc_set_threads(int n);
int orc_get_threads(void);

/* raycast.cl:29-47 (calculate_address), including the '>' fold. */
uint32_t orc_cyclic_address(int x, int y, int z, const kfx_round_buffer* rb);
/* raycast.cl:9-27 (get_tsdf / set_tsdf). */
float orc_tsdf_decode(int16_t v);
int16_t orc_tsdf_encode(float tsdf);

/* scale_depth (call site opencl_wrappers.cpp:37-43; kernel absent):
 * D = d_mm * sqrt(((u-cx)/fx)^2 + ((v-cy)/fy)^2 + 1) / 1000. */
void orc_scale_depth_f32(const uint16_t* depth, float* scaled, const kfx_intrinsics* intr);
void orc_scale_depth_f64(const uint16_t* depth, double* scaled, const kfx_intrinsics* intr);

/* integrate_tsdf_volume (call site opencl_wrappers.cpp:47-61; kernel absent):
 * projective TSDF update of every voxel.  Returns the number of voxels the
 * update rule touched (N_upd). */
uint64_t orc_integrate_f32(const float* scaled, kfx_tsdf* volume, const kfx_intrinsics* intr,
                           const kfx_matrix_3f* rotation_inverse, const kfx_point_3f* translation,
                           float trunc, const kfx_point_3f* cell, const kfx_round_buffer* rb);
/* Double-precision update of the same rule; rotation/translation in double. */
uint64_t orc_integrate_f64(const double* scaled, kfx_tsdf* volume, const kfx_intrinsics* intr,
                           const double* rotation_inverse9, const double* translation3,
                           double trunc, const double* cell3, const kfx_round_buffer* rb);

/* Colour volume update (Integrate_Colors, call site opencl_wrappers.cpp:181-214;
 * kernel absent): the KinectFusion colour-volume rule.  Returns the number of
 * voxels touched. */
uint64_t orc_integrate_colors(kfx_color* volume, const kfx_color* rgb, const kfx_xyz* vmap,
                              const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                              const kfx_point_3f* translation, float trunc, int max_weight,
                              const kfx_point_3f* cell, const kfx_round_buffer* rb);

/* raycast.cl:182-405.  Returns S, the voxel reads of the algorithm (first
 * sample, march samples inside the volume, 8 per in-bounds trilinear
 * evaluation). */
uint64_t orc_raycast_f32(const kfx_tsdf* volume, kfx_xyz* vmap, kfx_xyz* nmap, const kfx_round_buffer* rb,
                     const kfx_point_3f* translation, const kfx_matrix_3f* rotation, float trunc,
                     const kfx_intrinsics* intr, const kfx_point_3f* cell);
/* raycast.cl:95-145 (interpolateTrilineary), for direct checks. */
float orc_interpolate_trilinear_f32(const kfx_tsdf* volume, float px, float py, float pz,
                                    const kfx_point_3f* cell, const kfx_round_buffer* rb);

/* compute_bilateral_filtering.cl:8-57, intended normalised output. */
void orc_bilateral(const uint16_t* src, uint16_t* dst, int cols, int rows, float sigma_space, float sigma_color,
                   uint32_t max_distance);
/* compute_depth_pyramid (absent): KinectFusion pyrDown. */
void orc_pyr_down(const uint16_t* src, int scols, int srows, uint16_t* dst, float sigma_color);
/* compute_vertex_map (absent). */
void orc_vertex_map(const uint16_t* depth, kfx_xyz* vmap, const kfx_intrinsics* intr);
/* compute_normal_map_eigen (absent), double precision, Jacobi eigensolver. */
void orc_normal_map_f64(const kfx_xyz* vmap, kfx_xyz* nmap, int cols, int rows);
/* transform_map.cl:6-65. */
void orc_transform_maps(const kfx_xyz* vs, const kfx_xyz* ns, const kfx_matrix_3f* R, const kfx_point_3f* t,
                        int cols, int rows, kfx_xyz* vd, kfx_xyz* nd);
/* resize_vertex_map / resize_normal_map (absent). */
void orc_resize_map(const kfx_xyz* in, int icols, int irows, kfx_xyz* out, int normalize);

/* ICP sums of the 27 products (precision 0: single-precision rows as the GPU,
 * double accumulation; 1: double throughout).  Returns the correspondences. */
int orc_icp_sums(const kfx_xyz* vc, const kfx_xyz* nc, const kfx_xyz* vp, const kfx_xyz* np, const double* Rc,
                 const double* tc, const double* Rprev_inv, const double* tprev, const kfx_intrinsics* in,
                 double dthr, double athr, int precision, double out[27]);
/* Reference reduction path (estimate_combined + transform_estimation.cl), bit for bit. */
void orc_estimate_combined_f32(const kfx_xyz* vc, const kfx_xyz* nc, const kfx_xyz* vp, const kfx_xyz* np,
                               const float* Rc, const float* tc, const float* Rpi, const float* tp,
                               const kfx_intrinsics* in, float dthr, float athr, float mbuf[27]);
/* Cholesky solve + KinectFusion pose increment in double; 1 if singular. */
int orc_icp_solve_update(const double sums[27], double* R, double* t, double x_out[6]);

/* RGBD photometric odometry (Estimate_Isometry_By_RGBD,
 * common/src/opencl_wrappers.cpp:1328-1466, as the dense RGBD odometry of
 * the OpenCV rgbd module that the reference's kernels follow; the reference
 * path itself is not functional).  Pinned by: the twist exponential against
 * scipy's matrix exponential; identity and pure-translation correspondence
 * cases; recovery of the rendered motion between two frames.
 *   orc_rgbd_prepare: level arrays of one frame, per level l (cols >> l by
 *     rows >> l, n_l pixels) packed as [I, D, dx, dy] one after the other.
 *   orc_rgbd_corr: per previous-frame pixel of a level, the current-frame
 *     pixel paired with it (-1 none), for the estimate's inverse (Rinv, tinv)
 *     in single precision; returns the number of pairs.
 *   orc_rgbd_sums: sigma and the 27 double normal-equation sums (21 upper A
 *     entries row by row, then b) for the pairs and the estimate (R, t).
 *   orc_rgbd_twist_exp: R (9) and t (3) of exp of the twist ksi (6).
 *   orc_rgbd_estimate: the whole estimate; returns ok (a solve succeeded). */
void orc_rgbd_prepare(const kfx_color* rgb, const uint16_t* depth, int cols, int rows, int levels, float* out);
int orc_rgbd_corr(const float* i_prev, const float* d_prev, const float* d_curr, const float* dx_curr,
                  const float* dy_curr, const kfx_intrinsics* intr, const float* Rinv, const float* tinv,
                  float min_gradient_sq, float max_depth_diff, int32_t* corr);
int orc_rgbd_sums(const float* i_prev, const float* d_prev, const float* i_curr, const float* dx_curr,
                  const float* dy_curr, const kfx_intrinsics* intr, const int32_t* corr, const double* R,
                  const double* t, double* sigma, double sums[27]);
void orc_rgbd_twist_exp(const double ksi[6], double R[9], double t[3]);
int orc_rgbd_estimate(const kfx_color* rgb_prev, const uint16_t* depth_prev, const kfx_color* rgb_curr,
                      const uint16_t* depth_curr, const kfx_intrinsics* intr, int levels, const int* iterations,
                      float min_gradient, float max_depth_diff, const double* R0, const double* t0, double* R,
                      double* t, int* solves);

/* clean_color_slice.cl:41-67 pattern (TSDF or colour volume, elem_bytes 4). */
void orc_clean_slice(const kfx_point_3i* mn, const kfx_point_3i* mx, const kfx_round_buffer* rb, void* vol,
                     int elem_bytes);
/* Extract_Slice V+C (opencl_wrappers.cpp:856-964 with its two kernels). */
uint64_t orc_extract_slice(const kfx_point_3i* mn, const kfx_point_3i* mx, const kfx_round_buffer* rb,
                           const kfx_tsdf* vol, const kfx_color* col, kfx_xyzi* pts, kfx_color* cols,
                           uint64_t capacity);

/* Learned denoiser (no reference counterpart): forward and loss gradient in
 * double precision, optionally emulating the GPU's bfloat16 storage. */
int orc_dn_param_count(void);
void orc_dn_forward(const uint16_t* depth, int W, int H, const float* params, int bf16, double* y, double* hidden);
double orc_dn_loss_grad(const uint16_t* depth, const float* target_m, int W, int H, const float* params, int bf16,
                        double outlier_m, double* grad, int* count);

#ifdef __cplusplus
}
#endif

#endif
