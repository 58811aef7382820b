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

/* raycast.cl:182-405. */
void orc_raycast_f32(const kfx_tsdf* volume, kfx_xyz* vmap, kfx_xyz* nmap, const kfx_round_buffer* rb,
                     const kfx_point_3f* translation, const kfx_matrix_3f* rotation, float trunc,
                     const kfx_intrinsics* intr, const kfx_point_3f* cell);
/* raycast.cl:95-145 (interpolateTrilineary), for direct checks. */
float orc_interpolate_trilinear_f32(const kfx_tsdf* volume, float px, float py, float pz,
                                    const kfx_point_3f* cell, const kfx_round_buffer* rb);

#ifdef __cplusplus
}
#endif

#endif
