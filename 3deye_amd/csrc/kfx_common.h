// kfx_common.h — internal helpers shared by the HIP translation units of
// libkfx: error reporting, the library stream, launch checks and the device
// functions for the cyclic TSDF volume (address and voxel codec).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/kfx.h"

// Propagates a non-zero status.
#define KFX_TRY(expr)                 \
    do {                              \
        int kfx_rc_ = (expr);         \
        if (kfx_rc_) return kfx_rc_;  \
    } while (0)

namespace kfx {

// Device buffers that kernels use beyond their inputs and outputs.  A
// pipeline owns one for its lifetime, so the kernel arguments its captured
// graph holds stay valid; the standalone entry points share the library's
// (library_workspace), which grows on demand and never frees a buffer it has
// outgrown before the process ends, since a graph captured earlier may still
// point at it.
struct Workspace {
    float* bound = nullptr;    // depth-bound pyramid, levels 1..5 (floats)
    size_t bound_cap = 0;
    float2* cbound = nullptr;  // colour-bound pyramid, levels 0..5 (texels)
    size_t cbound_cap = 0;
    double* cons = nullptr;    // depth-consistency gate: 4 doubles per 256-pixel workgroup
    size_t cons_cap = 0;       // workgroups
    void* terms = nullptr;     // one frame's pose terms and skip flag (integration with a device pose)
    uint8_t* flags = nullptr;  // one flag per integration workgroup (frustum test)
    size_t flags_cap = 0;
};
// Sizes each buffer needs for images of cols x rows.
size_t bound_floats(int cols, int rows);
size_t cbound_texels(int cols, int rows);
size_t cons_blocks(int cols, int rows);
// Allocates the buffers of ws (the colour pyramid only with colors) for
// images of cols x rows; a buffer already large enough is kept.  Buffers
// replaced here are freed (owned workspaces are never captured mid-growth).
int workspace_reserve(Workspace& ws, int cols, int rows, bool colors);
void workspace_free(Workspace& ws);
// Integration workgroups of a volume (the flags the workspace must hold).
size_t integrate_blocks(const kfx_point_3i& voxels);
// Allocates the frame terms and the workgroup flags of ws for a volume.
int workspace_reserve_volume(Workspace& ws, const kfx_point_3i& voxels);
// Grows the library workspace to cols x rows; outgrown buffers are kept
// (not freed) until the process ends.  Returns an error inside a stream
// capture when it would have to allocate.
int library_workspace_reserve(int cols, int rows, bool colors, bool gate, const kfx_point_3i* voxels = nullptr);
Workspace& library_workspace();

// The integrations and the fallback gate over an explicit workspace (the
// public entry points pass the library's, a pipeline its own).  A workspace
// too small for the image runs the integrations unculled (same result).
int integrate_tsdf_ws(const Workspace& ws, bool profile, const uint16_t* depth, float* depth_scaled,
                      kfx_tsdf* volume, const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                      const kfx_point_3f* translation, float truncation_distance, const kfx_point_3f* cell_size,
                      const kfx_round_buffer* rb, uint8_t* bricks = nullptr);
// bound_ready: depth_scaled and ws's depth-bound pyramid already hold this
// frame's values (built by maps_pyramid_launch with depth_bound_job's job),
// so the pre-pass computes only the pose terms and workgroup flags.
int integrate_tsdf_posed_ws(const Workspace& ws, const uint16_t* depth, float* depth_scaled, kfx_tsdf* volume,
                            const kfx_intrinsics* intr, const float* dev_pose, const float* dev_prev_pose_inv,
                            float motion_threshold, float truncation_distance, const kfx_point_3f* cell_size,
                            const kfx_round_buffer* rb, const int32_t* dev_skip, uint8_t* bricks = nullptr,
                            bool bound_ready = false);

// Coarsest level of the integration's depth-bound pyramid (texels of 32 x 32
// pixels).
constexpr int kBoundLevels = 5;
constexpr int kBoundTile = 1 << kBoundLevels;
// The integration pre-pass's pose-independent half -- the scaled depth of
// every pixel and the depth-bound pyramid's levels 1..5, one 32 x 32 tile per
// workgroup -- as extra workgroups of the maps launch.
struct DepthBoundJob {
    const uint16_t* depth;
    float* scaled;
    float* bound;
    kfx_intrinsics intr;
    int bound_off[kBoundLevels + 1];
    int tiles_x, tiles;
};
// The job for images of *intr into ws's pyramid; false when ws holds no
// pyramid of that size (the pre-pass then builds both halves itself).
bool depth_bound_job(const Workspace& ws, const uint16_t* depth, float* depth_scaled, const kfx_intrinsics* intr,
                     DepthBoundJob* job);
// kfx_maps_pyramid_from_depth with an optional depth-bound job in the same
// launch.
int maps_pyramid_launch(const uint16_t* depth, uint16_t* const* depth_out, kfx_xyz* const* vmap, kfx_xyz* const* nmap,
                        const kfx_intrinsics* intr, int levels, float sigma_color, const DepthBoundJob* job);
// Brick flags of the raycast's skipping march (kfx_raycast.hip): bytes for
// a volume (0 when an axis is not a multiple of 16, which turns the skipping
// off); the exact flags of a volume (non-zero: the brick holds a negative
// value); the raycast at a device pose with the skipping march, which also
// builds levels 1 and 2 of the maps when pyramid = {v1, n1, v2, n2} (the
// separate resize's results, bit for bit).
size_t brick_flag_bytes(const kfx_point_3i& voxels);
int brick_flags_build(const kfx_tsdf* volume, const kfx_round_buffer* rb, uint8_t* bricks);
int raycast_posed_bricks(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                         const kfx_round_buffer* rb, const float* dev_pose, float truncation_distance,
                         const kfx_intrinsics* intr, const kfx_point_3f* cell_size, const int32_t* dev_skip,
                         const uint8_t* bricks, kfx_xyz* const* pyramid = nullptr);
int integrate_colors_ws(const Workspace& ws, kfx_color* color_volume, const kfx_color* rgb_frame,
                        const kfx_xyz* vertex_map, const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                        const kfx_point_3f* translation, float truncation_distance, int max_weight,
                        const kfx_point_3f* cell_size, const kfx_round_buffer* rb);
int integrate_colors_posed_ws(const Workspace& ws, kfx_color* color_volume, const kfx_color* rgb_frame,
                              const kfx_xyz* vertex_map, const kfx_intrinsics* intr, const float* dev_pose,
                              float truncation_distance, int max_weight, const kfx_point_3f* cell_size,
                              const kfx_round_buffer* rb, const int32_t* dev_skip);
// Lane exchanges of the wave butterflies (ICP sums, fallback gate).  Partner l ^ 32 and l ^ 16:
// v_permlane32_swap / v_permlane16_swap of (a, b) return, in a lane whose
// bit (5 or 4) is clear, (its own a, the partner's a), and where it is set
// (the partner's b, its own b).  Partner l ^ 8, 4, 2, 1 (within rows of 16
// lanes): DPP row_ror:8; quad_perm [3,2,1,0] then row_half_mirror
// (l ^ 3 ^ 7); quad_perm [2,3,0,1]; quad_perm [1,0,3,2].  Every lane of the
// wave takes part.
template <int kXor>
__device__ __forceinline__ unsigned dpp_xor_u32(unsigned v) {
    int x = (int)v;
    if constexpr (kXor == 8) {
        x = __builtin_amdgcn_mov_dpp(x, 0x128, 0xf, 0xf, false);
    } else if constexpr (kXor == 4) {
        x = __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x1B, 0xf, 0xf, false), 0x141, 0xf, 0xf, false);
    } else if constexpr (kXor == 2) {
        x = __builtin_amdgcn_mov_dpp(x, 0x4E, 0xf, 0xf, false);
    } else {
        static_assert(kXor == 1, "DPP exchange for xor 8, 4, 2, 1");
        x = __builtin_amdgcn_mov_dpp(x, 0xB1, 0xf, 0xf, false);
    }
    return (unsigned)x;
}

template <int kXor>
__device__ __forceinline__ float dpp_xor(float v) {
    return __uint_as_float(dpp_xor_u32<kXor>(__float_as_uint(v)));
}
template <int kXor>
__device__ __forceinline__ double dpp_xor(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint64_t lo = dpp_xor_u32<kXor>((unsigned)b), hi = dpp_xor_u32<kXor>((unsigned)(b >> 32));
    return __longlong_as_double((long long)((hi << 32) | lo));
}

// (kept, received) in either order for the partner l ^ (2 kHalf), kHalf =
// 16 or 8, from a = the low-half value and b = the high-half value.
template <int kHalf>
__device__ __forceinline__ void swap_halves(float a, float b, float& x, float& y) {
    static_assert(kHalf == 16 || kHalf == 8, "permlane swap distances");
    const auto r = kHalf == 16 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false)
                               : __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
}
template <int kHalf>
__device__ __forceinline__ void swap_halves(double a, double b, double& x, double& y) {
    const uint64_t ua = (uint64_t)__double_as_longlong(a), ub = (uint64_t)__double_as_longlong(b);
    float xl, yl, xh, yh;
    swap_halves<kHalf>(__uint_as_float((unsigned)ua), __uint_as_float((unsigned)ub), xl, yl);
    swap_halves<kHalf>(__uint_as_float((unsigned)(ua >> 32)), __uint_as_float((unsigned)(ub >> 32)), xh, yh);
    x = __longlong_as_double((long long)(((uint64_t)__float_as_uint(xh) << 32) | __float_as_uint(xl)));
    y = __longlong_as_double((long long)(((uint64_t)__float_as_uint(yh) << 32) | __float_as_uint(yl)));
}

// Sum over the wave by the xor butterfly (offsets 32, 16, 8, 4, 2, 1): each
// lane adds its partner's value to its own at every step, so every lane ends
// with the same total, in the association of `v += shfl_xor(v, off)`.  Lane
// 0's total is also that of the last six levels of a halving tree (lane
// t < h adds lane t + h, h = 32 .. 1): for t < h the partner t ^ h is t + h.
__device__ __forceinline__ double wave_sum_xor(double v, int lane) {
    double x, y;
    swap_halves<16>(v, v, x, y);
    v = v + ((lane & 32) ? x : y);
    swap_halves<8>(v, v, x, y);
    v = v + ((lane & 16) ? x : y);
    v = v + dpp_xor<8>(v);
    v = v + dpp_xor<4>(v);
    v = v + dpp_xor<2>(v);
    v = v + dpp_xor<1>(v);
    return v;
}

// Per-pixel terms of the fallback gate: the model vertex of pixel p
// (volume coordinates) seen from the camera at `pose` (single precision, R
// row-major camera -> volume then t), projected to the nearest pixel of
// `depth`; v = (squared depth difference if below max_diff, 1 if so,
// 1 if the projection finds a depth, 1 if the vertex exists).  In two parts:
// gate_project (v[3]; the pixel index idx the vertex projects to, -1 when
// none, and its camera depth cz), then gate_depth with d = depth[idx].
__device__ __forceinline__ void gate_project(const kfx_xyz& x, const float* pose, const kfx_intrinsics& in,
                                             double (&v)[4], int& idx, float& cz) {
    v[0] = v[1] = v[2] = v[3] = 0.0;
    idx = -1;
    cz = 0.0f;
    if (isnan(x.x)) return;
    v[3] = 1.0;
    const float dx = x.x - pose[9], dy = x.y - pose[10], dz = x.z - pose[11];
    const float cx = pose[0] * dx + pose[3] * dy + pose[6] * dz;
    const float cy = pose[1] * dx + pose[4] * dy + pose[7] * dz;
    cz = pose[2] * dx + pose[5] * dy + pose[8] * dz;
    if (!(cz > 0.0f)) return;
    const float fu = rintf(cx * in.fx / cz + in.cx), fv = rintf(cy * in.fy / cz + in.cy);
    if (!(fu >= 0.0f && fv >= 0.0f && fu < (float)in.cols && fv < (float)in.rows)) return;
    idx = (int)fv * in.cols + (int)fu;
}
__device__ __forceinline__ void gate_depth(uint16_t d, float cz, float max_diff, double (&v)[4]) {
    if (!d) return;
    v[2] = 1.0;
    const float e = (float)d * 0.001f - cz;
    if (fabsf(e) < max_diff) {
        v[0] = (double)e * e;
        v[1] = 1.0;
    }
}
__device__ __forceinline__ void gate_terms(const kfx_xyz* __restrict__ vmap, const float* pose,
                                           const uint16_t* __restrict__ depth, const kfx_intrinsics& in,
                                           float max_diff, int p, double (&v)[4]) {
    const kfx_xyz x = p < in.cols * in.rows ? vmap[p] : kfx_xyz{NAN, NAN, NAN};
    int idx;
    float cz;
    gate_project(x, pose, in, v, idx, cz);
    if (idx >= 0) gate_depth(depth[idx], cz, max_diff, v);
}

// The gate's decision from its totals: fall back with at least 100 valid
// model vertices when the RMS difference of the matched pixels exceeds
// fallback_m or fewer than half the vertices match.
__device__ __forceinline__ int32_t gate_decision(double sse, double cnt, double n_valid, double fallback_m) {
    const double rms = cnt > 0.0 ? sqrt(sse / cnt) : 0.0;
    return (fallback_m > 0.0 && n_valid >= 100.0 && (rms > fallback_m || cnt < 0.5 * n_valid)) ? 1 : 0;
}

// The fallback gate of the learned pipeline (depth_consistency's inputs and
// outputs), evaluated at the end of a one-launch frame ICP at its final pose.
struct FrameGate {
    const kfx_xyz* vmap_model;
    const uint16_t* depth;
    kfx_intrinsics intr;
    float max_diff;
    double fallback_m;
    double* stats;        // [4]: squared error, count, matched, valid model vertices
    int32_t* flag;        // 1: fall back
    int32_t* host_flag;   // copy of the flag, or null
};
// kfx_icp_frame, with the gate (when not null) after the last iteration:
// in the same launch when the frame runs as one, else as depth_consistency
// after the separate launches; the results are the same either way.
int icp_frame_gated(kfx_icp* h, const kfx_icp_level* levels, int begin_frame, float distance_threshold,
                    float angle_threshold, int tagged, double converge_rotation, double converge_translation,
                    const FrameGate* gate, const Workspace* gate_ws);
// A publishing frame launch (kfx_icp_frame as one launch after
// kfx_icp_publish_next) raises a pinned completion flag right after its host
// copy of the state and the gate's host outputs.  icp_clear_done lowers it
// (a host store; call it before enqueueing or replaying such a launch, with
// none in flight).  icp_read_published_early returns the published state as
// soon as the flag is up, without waiting for the work enqueued after the
// launch; when the last publishing launch was not a frame launch it is
// kfx_icp_read_published.
void icp_clear_done(kfx_icp* h);
int icp_read_published_early(kfx_icp* h, kfx_icp_result* out);
// The fallback gate (kfx_denoise.hip); ws.cons must hold cons_blocks().
int depth_consistency(const Workspace& ws, const kfx_xyz* vmap_model, const float* dev_pose, const uint16_t* depth,
                      const kfx_intrinsics* intr, float max_diff, double* dev_stats, double fallback_m,
                      int32_t* dev_flag, int32_t* host_flag);

// Resolves, for the current device, the device address of the fixed scratch
// array of kfx_integrate_tsdf_noalloc (kfx_volume.hip); called by kfx_init.
int noalloc_setup(int device);

// Library stream (created by kfx_init, default priority) on which every
// entry point enqueues its work.
hipStream_t stream();
int current_device();
bool initialised();

// Records a formatted message for kfx_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...);

// Brackets a kernel launch with HIP events on the library stream while
// kfx_profile_enable(1) is in force (no-op otherwise, and for a null name).
struct ProfScope {
    explicit ProfScope(const char* name);
    ~ProfScope();
    int slot;
};

}  // namespace kfx

#define KFX_CHECK_HIP(expr)                                                    \
    do {                                                                       \
        hipError_t kfx_e_ = (expr);                                            \
        if (kfx_e_ != hipSuccess)                                              \
            return ::kfx::set_error(KFX_ERR_HIP, "%s failed: %s", #expr,       \
                                    hipGetErrorString(kfx_e_));                \
    } while (0)

#define KFX_CHECK_LAUNCH() KFX_CHECK_HIP(hipGetLastError())

#define KFX_REQUIRE(cond, ...)                                                 \
    do {                                                                       \
        if (!(cond)) return ::kfx::set_error(KFX_ERR_INVALID_ARG, __VA_ARGS__); \
    } while (0)

#define KFX_REQUIRE_INIT()                                                     \
    do {                                                                       \
        if (!::kfx::initialised())                                             \
            return ::kfx::set_error(KFX_ERR_NOT_INITIALISED,                   \
                                    "kfx_init has not been called");           \
    } while (0)

namespace kfx {

// Single-precision pose passed by value to kernels.
struct kfx_pose_f {
    float R[9];
    float t[3];
};

// Resize of a vertex or normal map (KinectFusion resizeMap): the mean of a
// 2x2 block when all four are valid, NaN otherwise; normals renormalised.
// The scaled depth of one pixel (scale_depth_kernel's arithmetic): depth in
// metres times the length of the pixel's ray at unit depth.
__device__ __forceinline__ float scaled_depth(const uint16_t* __restrict__ depth, const kfx_intrinsics& intr, int u,
                                              int v) {
    const float xl = ((float)u - intr.cx) / intr.fx;
    const float yl = ((float)v - intr.cy) / intr.fy;
    const float lambda = sqrtf(xl * xl + yl * yl + 1.0f);
    return (float)depth[v * intr.cols + u] * lambda / 1000.0f;
}

__device__ __forceinline__ kfx_xyz resize4(kfx_xyz a, kfx_xyz b, kfx_xyz c, kfx_xyz d, bool normalize) {
    kfx_xyz o{NAN, NAN, NAN};
    if (!(isnan(a.x) || isnan(b.x) || isnan(c.x) || isnan(d.x))) {
        o.x = (a.x + b.x + c.x + d.x) / 4.0f;
        o.y = (a.y + b.y + c.y + d.y) / 4.0f;
        o.z = (a.z + b.z + c.z + d.z) / 4.0f;
        if (normalize) {
            const float s = 1.0f / sqrtf(o.x * o.x + o.y * o.y + o.z * o.z);
            o.x *= s; o.y *= s; o.z *= s;
        }
    }
    return o;
}

// ---------------------------------------------------------------------------
// Cyclic volume addressing (reference: GPU/opencl/raycast.cl:29-47).
// Per-axis (coordinate + origin) modulo the axis size, linearised x-fastest,
// plus start_index, folded modulo buffer_length when strictly greater than it
// (the reference's `>` test is kept as is).
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t cyclic_address(int x, int y, int z,
                                                            const kfx_round_buffer& rb) {
    const int rx = (x + rb.origin_GRID.x) % rb.voxels_size.x;
    const int ry = (y + rb.origin_GRID.y) % rb.voxels_size.y;
    const int rz = (z + rb.origin_GRID.z) % rb.voxels_size.z;
    uint64_t a = (uint64_t)(uint32_t)(rb.voxels_size.x * rb.voxels_size.y * rz) +
                 (uint64_t)(uint32_t)(rb.voxels_size.x * ry) + (uint64_t)(uint32_t)rx;
    a += (uint64_t)(int64_t)rb.start_index;
    if (a > (uint64_t)(int64_t)rb.buffer_length) a %= (uint64_t)(int64_t)rb.buffer_length;
    return (uint32_t)a;
}

// (v + o) wrapped once into [0, n); valid when 0 <= v + o < 2n, which holds
// for local coordinates in [0, n) and origins normalised into [0, n).
__host__ __device__ __forceinline__ int wrap_once(int v, int n) { return v >= n ? v - n : v; }

// cyclic_address for local coordinates in [0, N) per axis with the volume
// kernels' preconditions (origin_GRID in [0, N), start_index == 0, checked on
// the host): equal to cyclic_address there, without integer division.
__host__ __device__ __forceinline__ uint32_t cyclic_address_in_range(int x, int y, int z,
                                                                     const kfx_round_buffer& rb) {
    const uint32_t rx = (uint32_t)wrap_once(x + rb.origin_GRID.x, rb.voxels_size.x);
    const uint32_t ry = (uint32_t)wrap_once(y + rb.origin_GRID.y, rb.voxels_size.y);
    const uint32_t rz = (uint32_t)wrap_once(z + rb.origin_GRID.z, rb.voxels_size.z);
    return ((uint32_t)rb.voxels_size.y * rz + ry) * (uint32_t)rb.voxels_size.x + rx;
}

// Correctly rounded square root (IEEE sqrtf) without the library's range
// scaling: the hardware v_sqrt_f32 (within one ulp) corrected by the signs of
// its neighbours' residuals x - s' s, which picks the correctly rounded root
// for every x in [2^-100, 2^100) whatever the starting value within an ulp
// (checked exhaustively on the CPU by tests/cpp/exact_arith_check.c).  Other
// x (zero, tiny, huge, NaN) take sqrtf.
__device__ __forceinline__ float sqrt_rn(float x) {
    if (!(x >= 0x1p-100f && x < 0x1p100f)) return sqrtf(x);
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sdn = __uint_as_float(__float_as_uint(s) - 1u), sup = __uint_as_float(__float_as_uint(s) + 1u);
    const float rdn = fmaf(-sdn, s, x), rup = fmaf(-sup, s, x);
    float r = s;
    if (rdn <= 0.0f) r = sdn;
    if (rup > 0.0f) r = sup;
    return r;
}

// sqrt_rn for arguments known to be non-negative and finite, without its
// range branch: x is first clamped into [2^-100, 2^100).  The clamp changes
// the root only for x < 2^-100 (a root below 2^-50) and x >= 2^100; the
// integration subtracts the root from a scaled depth of at least 1 mm, where
// a root below 2^-50 is below half an ulp either way, and its squared
// distances stay far below 2^100.
__device__ __forceinline__ float sqrt_rn_in_range(float x) {
    const float xc = __builtin_amdgcn_fmed3f(x, 0x1p-100f, 0x1.fffffep99f);
    const float s = __builtin_amdgcn_sqrtf(xc);
    const float sdn = __uint_as_float(__float_as_uint(s) - 1u), sup = __uint_as_float(__float_as_uint(s) + 1u);
    const float rdn = fmaf(-sdn, s, xc), rup = fmaf(-sup, s, xc);
    float r = s;
    if (rdn <= 0.0f) r = sdn;
    if (rup > 0.0f) r = sup;
    return r;
}

// Correctly rounded 1/x for x in [2^-126, 2^126): v_rcp_f32 (not correctly
// rounded by itself) and one Newton step with FMAs.  Equal to the IEEE
// quotient for every float of that range on MI355X, checked exhaustively on
// the device (kfx_check_reciprocal, run by the GPU tests); outside it (zero,
// subnormals, the top binade, infinities) the result may differ, so callers
// keep their operands inside.
__device__ __forceinline__ float rcp_rn(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
}
constexpr float kRcpMin = 0x1p-126f, kRcpMax = 0x1p126f;

// The quotient a / (w + 1) of the TSDF running average, for weights w in
// [0, 128], without a division: y = RN(1/b) from v_rcp_f32 (within one ulp)
// and one Newton step with FMAs (correct from every starting value within an
// ulp, for every b in [1, 129]), then Markstein's correction q0 = a y,
// q = q0 + (a - q0 b) y with the exact FMA residual.  Equal to the IEEE
// quotient for every |a| in [2^-100, 129] and b in [2, 129] (b = 1 is
// exact); for |a| < 2^-100 both quotients are below 1/32767 in magnitude,
// so tsdf_encode maps either to 0.  Checked exhaustively on the CPU by
// tests/cpp/exact_arith_check.c.
__device__ __forceinline__ float tsdf_average_quotient(float a, int b) {
    const float fb = (float)b;
    const float y0 = __builtin_amdgcn_rcpf(fb);
    const float y = fmaf(fmaf(-fb, y0, 1.0f), y0, y0);
    const float q0 = a * y;
    return fmaf(fmaf(-q0, fb, a), y, q0);
}

// Voxel codec (raycast.cl:9-27): decode value / 32767, encode rtz then clamp.
// The decode is the correctly rounded quotient, formed without a division:
// q0 = v * RN(1/32767), then one FMA correction with the exact residual
// (Markstein's theorem; checked for all 65536 int16 values).
__device__ __forceinline__ float tsdf_decode(int16_t v) {
    constexpr float kInv = 1.0f / (float)KFX_MAX_DEPTH_VALUE;
    const float a = (float)v;
    const float q0 = a * kInv;
    return fmaf(fmaf(-q0, (float)KFX_MAX_DEPTH_VALUE, a), kInv, q0);
}
__device__ __forceinline__ int16_t tsdf_encode(float tsdf) {
    // rtz conversion of a finite float; saturate before the int cast so the
    // conversion is defined for every input.
    float s = tsdf * (float)KFX_MAX_DEPTH_VALUE;
    s = fminf(fmaxf(s, -2147483520.0f), 2147483520.0f);
    int fixedp = (int)s;  // truncation toward zero
    fixedp = max(-KFX_MAX_DEPTH_VALUE, min(KFX_MAX_DEPTH_VALUE, fixedp));
    return (int16_t)fixedp;
}

// tsdf_encode for |tsdf| < 2^16, the range of every running average (a
// weighted mean of values within [-1, 1] and one sdf / trunc >= -1 - 2^-23):
// there the product cannot reach the saturation that tsdf_encode guards.
__device__ __forceinline__ int16_t tsdf_encode_unit(float tsdf) {
    const int fixedp = (int)(tsdf * (float)KFX_MAX_DEPTH_VALUE);  // truncation toward zero
    return (int16_t)max(-KFX_MAX_DEPTH_VALUE, min(KFX_MAX_DEPTH_VALUE, fixedp));
}

__host__ __device__ __forceinline__ unsigned divup(unsigned a, unsigned b) { return (a + b - 1) / b; }

}  // namespace kfx
