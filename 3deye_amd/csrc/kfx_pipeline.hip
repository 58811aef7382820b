// kfx_pipeline.hip — the per-frame KinectFusion driver (the reference's
// absent fusion application, Makefile:3,19): preprocessing, coarse-to-fine
// ICP, integration, raycast and the cyclic volume shift, on one GPU.
//
// Two modes over the same kernels:
//   KFX_PIPELINE_PLAIN — the reference's control flow: every step waits for
//     the GPU (its opencl_kernel::execute ends in clFinish), each ICP
//     iteration downloads the 27 sums (Estimate_Combined) and the host solves
//     and updates the pose in double precision; fixed iteration counts.
//   KFX_PIPELINE_FUSED — the pose lives on the GPU: one fused launch per ICP
//     iteration (FP64 sums of the FP32 products, solve and
//     pose update on the GPU), optional early exit per level,
//     integration and raycast read the pose from device memory, and the whole
//     frame can be replayed as one HIP graph; the host reads the pose once per
//     frame to decide the cyclic shift.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "kfx_common.h"

struct kfx_pipeline {
    kfx_pipeline_config cfg;
    kfx_round_buffer rb;
    kfx_point_3f cell;
    double origin_world[3];
    kfx_intrinsics lvl[KFX_ICP_LEVELS];
    kfx_tsdf* vol = nullptr;
    kfx_color* cvol = nullptr;            // colour volume (cfg.colors)
    kfx_color* rgb = nullptr;             // last colour frame (cfg.colors)
    uint16_t* raw = nullptr;
    uint16_t* dep[KFX_ICP_LEVELS] = {};
    float* scaled = nullptr;
    kfx_xyz *vc[KFX_ICP_LEVELS] = {}, *nc[KFX_ICP_LEVELS] = {};
    kfx_xyz *vp[KFX_ICP_LEVELS] = {}, *np[KFX_ICP_LEVELS] = {};
    float* gbuf = nullptr;
    float* mbuf = nullptr;
    int gbuf_stride = 0;
    kfx_icp* icp = nullptr;
    kfx_slicer* slicer = nullptr;
    kfx_xyzi* ext_pts = nullptr;
    kfx_color* ext_col = nullptr;
    uint64_t ext_cap = 0;
    double R[9], t[3], Rprev[9], tprev[3];
    long frame = 0;
    hipGraph_t graph = nullptr;           // whole frame (bilateral) or ICP part (denoiser)
    hipGraphExec_t graph_exec = nullptr;
    kfx_denoiser* denoiser = nullptr;
    double* hstats = nullptr;             // gate statistics (pinned host memory the gate kernel writes)
    int32_t* dgate = nullptr;             // fallback decision of the gate (device)
    int32_t* hgate = nullptr;             // host mirror of dgate (pinned, written by the gate kernel)
    uint8_t* bricks = nullptr;            // brick flags of the skipping raycast (fused mode; null: full march)
    kfx::Workspace ws;                    // depth- and colour-bound pyramids, gate partials: this pipeline's own,
                                          // so its captured graph never points at another's buffers
    uint64_t total_extracted = 0;
    int shifts = 0;
    int fallbacks = 0;
    bool icp_in_flight = false;           // a publishing ICP launch was enqueued and its result not yet read
    bool bound_ready = false;             // the maps launch built this frame's scaled depth and depth-bound pyramid
    const uint16_t* in = nullptr;         // the frame's raw depth: the caller's buffer, or raw (graph replay)
};

using namespace kfx;

namespace {

void mat_T(const double* a, double* out) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[3 * r + c] = a[3 * c + r];
}

kfx_matrix_3f to_m3f(const double* a) {
    kfx_matrix_3f m;
    for (int r = 0; r < 3; ++r) m.data[r] = kfx_point_3f{(float)a[3 * r], (float)a[3 * r + 1], (float)a[3 * r + 2]};
    return m;
}

kfx_point_3f to_p3f(const double* a) { return kfx_point_3f{(float)a[0], (float)a[1], (float)a[2]}; }

// Host Cholesky solve and KinectFusion increment (the plain mode's host step
// after Estimate_Combined).  Returns false when A is not positive definite.
bool host_solve_update(const double* A, const double* b, double* R, double* t) {
    double L[36] = {0};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = A[i * 6 + j];
            for (int p = 0; p < j; ++p) s -= L[i * 6 + p] * L[j * 6 + p];
            if (i == j) {
                if (!(s > 1e-300) || !std::isfinite(s)) return false;
                L[i * 6 + i] = sqrt(s);
            } else {
                L[i * 6 + j] = s / L[j * 6 + j];
            }
        }
    double y[6], x[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int p = 0; p < i; ++p) s -= L[i * 6 + p] * y[p];
        y[i] = s / L[i * 6 + i];
    }
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int p = i + 1; p < 6; ++p) s -= L[p * 6 + i] * x[p];
        x[i] = s / L[i * 6 + i];
    }
    const double ca = cos(x[0]), sa = sin(x[0]), cb = cos(x[1]), sb = sin(x[1]), cg = cos(x[2]), sg = sin(x[2]);
    const double Ri[9] = {cg * cb, cg * sb * sa - sg * ca, cg * sb * ca + sg * sa,
                          sg * cb, sg * sb * sa + cg * ca, sg * sb * ca - cg * sa,
                          -sb, cb * sa, cb * ca};
    double Rn[9], tn[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Rn[3 * r + c] = Ri[3 * r] * R[c] + Ri[3 * r + 1] * R[3 + c] + Ri[3 * r + 2] * R[6 + c];
        tn[r] = Ri[3 * r] * t[0] + Ri[3 * r + 1] * t[1] + Ri[3 * r + 2] * t[2] + x[3 + r];
    }
    memcpy(R, Rn, sizeof(Rn));
    memcpy(t, tn, sizeof(tn));
    return true;
}

// KinectFusion integration metric: (rotation angle + translation) / 2.
double motion(const double* R, const double* t, const double* Rp, const double* tp) {
    double tr = 0;
    for (int i = 0; i < 9; ++i) tr += R[i] * Rp[i];
    const double ang = acos(std::min(1.0, std::max(-1.0, 0.5 * (tr - 1.0))));
    const double d = sqrt((t[0] - tp[0]) * (t[0] - tp[0]) + (t[1] - tp[1]) * (t[1] - tp[1]) +
                          (t[2] - tp[2]) * (t[2] - tp[2]));
    return 0.5 * (ang + d);
}

// The reference's execute() ends every kernel with clFinish; plain mode
// mirrors that.
int step_sync(const kfx_pipeline* p) {
    if (p->cfg.mode == KFX_PIPELINE_PLAIN) return kfx_synchronize();
    return KFX_OK;
}

bool use_denoiser(const kfx_pipeline* p) { return p->cfg.denoiser && p->denoiser != nullptr; }

// Level-0 filtered depth from the raw frame (learned denoiser or bilateral
// filter), then the depth pyramid and the vertex/normal maps of every level.
int preprocess(kfx_pipeline* p, bool learned) {
    const kfx_intrinsics& in = p->lvl[0];
    if (learned)
        KFX_TRY(kfx_denoiser_forward(p->denoiser, p->in, p->dep[0], nullptr));
    else
        KFX_TRY(kfx_bilateral_filter(p->in, p->dep[0], in.cols, in.rows, p->cfg.sigma_space, p->cfg.sigma_color,
                                     0xFFFFFFFFu));
    KFX_TRY(step_sync(p));
    p->bound_ready = false;
    if (p->cfg.mode == KFX_PIPELINE_FUSED) {
        // The depth pyramid and every level's vertex and normal maps in one
        // launch (same values as the steps below; levels halve exactly),
        // with the integration pre-pass's scaled depth and depth-bound
        // pyramid of the raw frame as extra workgroups.
        DepthBoundJob job;
        const bool built = depth_bound_job(p->ws, p->in, p->scaled, &p->lvl[0], &job);
        KFX_TRY(maps_pyramid_launch(p->dep[0], p->dep, p->vc, p->nc, p->lvl, KFX_ICP_LEVELS, p->cfg.sigma_color,
                                    built ? &job : nullptr));
        p->bound_ready = built;
        return KFX_OK;
    }
    for (int l = 1; l < KFX_ICP_LEVELS; ++l) {
        KFX_TRY(kfx_pyr_down(p->dep[l - 1], p->lvl[l - 1].cols, p->lvl[l - 1].rows, p->dep[l], p->cfg.sigma_color));
        KFX_TRY(step_sync(p));
    }
    for (int l = 0; l < KFX_ICP_LEVELS; ++l) {
        KFX_TRY(kfx_vertex_map(p->dep[l], p->vc[l], &p->lvl[l]));
        KFX_TRY(step_sync(p));
    }
    for (int l = 0; l < KFX_ICP_LEVELS; ++l) {
        KFX_TRY(kfx_normal_map(p->vc[l], p->nc[l], p->lvl[l].cols, p->lvl[l].rows));
        KFX_TRY(step_sync(p));
    }
    return KFX_OK;
}

int resize_prev(kfx_pipeline* p) {
    static_assert(KFX_ICP_LEVELS == 3, "the one-launch resize builds levels 1 and 2");
    if (p->cfg.mode == KFX_PIPELINE_FUSED)
        return kfx_resize_pyramid(p->vp[0], p->np[0], p->lvl[0].cols, p->lvl[0].rows, p->vp[1], p->np[1], p->vp[2],
                                  p->np[2]);
    for (int l = 1; l < KFX_ICP_LEVELS; ++l) {
        KFX_TRY(kfx_resize_map(p->vp[l - 1], p->lvl[l - 1].cols, p->lvl[l - 1].rows, p->vp[l], 0));
        KFX_TRY(step_sync(p));
        KFX_TRY(kfx_resize_map(p->np[l - 1], p->lvl[l - 1].cols, p->lvl[l - 1].rows, p->np[l], 1));
        KFX_TRY(step_sync(p));
    }
    return KFX_OK;
}

int integrate_host_pose(kfx_pipeline* p) {
    double Ri[9];
    mat_T(p->R, Ri);
    const kfx_matrix_3f m = to_m3f(Ri);
    const kfx_point_3f tt = to_p3f(p->t);
    KFX_TRY(integrate_tsdf_ws(p->ws, true, p->in, p->scaled, p->vol, &p->lvl[0], &m, &tt, p->cfg.truncation,
                              &p->cell, &p->rb, p->bricks));
    return step_sync(p);
}

int raycast_host_pose(kfx_pipeline* p) {
    const kfx_matrix_3f m = to_m3f(p->R);
    const kfx_point_3f tt = to_p3f(p->t);
    if (p->bricks)
        KFX_TRY(kfx_raycast_bricks(p->vol, p->vp[0], p->np[0], &p->rb, &tt, &m, p->cfg.truncation, &p->lvl[0],
                                   &p->cell, p->bricks));
    else
        KFX_TRY(kfx_raycast(p->vol, p->vp[0], p->np[0], &p->rb, &tt, &m, p->cfg.truncation, &p->lvl[0], &p->cell));
    KFX_TRY(step_sync(p));
    return resize_prev(p);
}

// Colour integration at the host pose from the raycast just made (vp[0]).
int colors_host_pose(kfx_pipeline* p) {
    if (!p->cvol) return KFX_OK;
    double Ri[9];
    mat_T(p->R, Ri);
    const kfx_matrix_3f m = to_m3f(Ri);
    const kfx_point_3f tt = to_p3f(p->t);
    KFX_TRY(integrate_colors_ws(p->ws, p->cvol, p->rgb, p->vp[0], &p->lvl[0], &m, &tt, p->cfg.truncation,
                                p->cfg.color_max_weight, &p->cell, &p->rb));
    return step_sync(p);
}

int icp_plain(kfx_pipeline* p, kfx_frame_info* info) {
    double Rpi[9];
    mat_T(p->Rprev, Rpi);
    const kfx_matrix_3f mpi = to_m3f(Rpi);
    const kfx_point_3f tp = to_p3f(p->tprev);
    for (int l = KFX_ICP_LEVELS - 1; l >= 0; --l)
        for (int it = 0; it < p->cfg.icp_iterations[l]; ++it) {
            const kfx_matrix_3f mc = to_m3f(p->R);
            const kfx_point_3f tc = to_p3f(p->t);
            double A[36], b[6];
            KFX_TRY(kfx_estimate_combined(&mc, &tc, p->vc[l], p->nc[l], &mpi, &tp, &p->lvl[l], p->vp[l], p->np[l],
                                          p->cfg.distance_threshold, p->cfg.angle_threshold, p->gbuf, p->gbuf_stride,
                                          p->mbuf, A, b));
            if (info) info->icp_iterations += 1;
            if (!host_solve_update(A, b, p->R, p->t)) {
                if (info) info->icp_singular += 1;
            }
        }
    return KFX_OK;
}

// Fused ICP from the state's previous pose (one launch where the device
// allows it).  Enqueues only.  The last iteration publishes the state to
// pinned host memory, which the host reads after the frame
// (kfx_icp_read_published).  Level tags only with the convergence test on:
// an untagged iteration skips the read of the convergence state.
int icp_part(kfx_pipeline* p, bool begin, const kfx::FrameGate* gate = nullptr) {
    kfx_icp_level lv[KFX_ICP_LEVELS];
    for (int l = 0; l < KFX_ICP_LEVELS; ++l) {
        lv[l].vmap_current = p->vc[l];
        lv[l].nmap_current = p->nc[l];
        lv[l].vmap_previous = p->vp[l];
        lv[l].nmap_previous = p->np[l];
        lv[l].intr = p->lvl[l];
        lv[l].iterations = p->cfg.icp_iterations[l];
    }
    KFX_TRY(kfx_icp_publish_next(p->icp));
    return kfx::icp_frame_gated(p->icp, lv, begin ? 1 : 0, p->cfg.distance_threshold, p->cfg.angle_threshold,
                                p->cfg.converge_rotation > 0.0 ? 1 : 0, p->cfg.converge_rotation,
                                p->cfg.converge_translation, gate, &p->ws);
}

// Integration and raycast at the device pose, skipped on the device when
// skip points at a non-zero flag.  Enqueues only.
int map_part_skip(kfx_pipeline* p, const int32_t* skip) {
    const float *dpose = nullptr, *dprev = nullptr;
    KFX_TRY(kfx_icp_pose_device(p->icp, &dpose, &dprev));
    KFX_TRY(integrate_tsdf_posed_ws(p->ws, p->in, p->scaled, p->vol, &p->lvl[0], dpose, dprev,
                                    p->cfg.integrate_threshold, p->cfg.truncation, &p->cell, &p->rb, skip,
                                    p->bricks, p->bound_ready));
    // The skipping raycast builds levels 1 and 2 itself (no resize launch).
    kfx_xyz* const pyramid[4] = {p->vp[1], p->np[1], p->vp[2], p->np[2]};
    const bool fused_pyramid = p->bricks != nullptr && p->lvl[1].cols == p->lvl[0].cols / 2 &&
                               p->lvl[1].rows == p->lvl[0].rows / 2 && p->lvl[2].cols == p->lvl[1].cols / 2 &&
                               p->lvl[2].rows == p->lvl[1].rows / 2;
    if (p->bricks)
        KFX_TRY(raycast_posed_bricks(p->vol, p->vp[0], p->np[0], &p->rb, dpose, p->cfg.truncation, &p->lvl[0],
                                     &p->cell, skip, p->bricks, fused_pyramid ? pyramid : nullptr));
    else
        KFX_TRY(kfx_raycast_posed(p->vol, p->vp[0], p->np[0], &p->rb, dpose, p->cfg.truncation, &p->lvl[0],
                                  &p->cell, skip));
    if (p->cvol)
        KFX_TRY(integrate_colors_posed_ws(p->ws, p->cvol, p->rgb, p->vp[0], &p->lvl[0], dpose, p->cfg.truncation,
                                          p->cfg.color_max_weight, &p->cell, &p->rb, skip));
    return fused_pyramid ? KFX_OK : resize_prev(p);
}

int map_part(kfx_pipeline* p) { return map_part_skip(p, nullptr); }

// The whole learned frame, enqueued at once: denoiser, pyramid, maps, ICP;
// the consistency of the denoised frame with the model at the new pose
// (fallback gate), decided on the device; integration and raycast, which
// the device skips when the gate falls back (the host then redoes the frame
// with the bilateral filter).  No host round trip inside the frame.
// The state after a frame's ICP: the copy its last iteration published, or
// a read when the configuration runs no ICP iteration.
int icp_read_frame(kfx_pipeline* p, kfx_icp_result* r) {
    const int n = p->cfg.icp_iterations[0] + p->cfg.icp_iterations[1] + p->cfg.icp_iterations[2];
    KFX_TRY(n > 0 ? icp_read_published_early(p->icp, r) : kfx_icp_read(p->icp, r));
    p->icp_in_flight = false;
    return KFX_OK;
}

// Before a fused frame's ICP is enqueued (or replayed): the completion flag
// of its launch lowered, once no earlier publishing launch can still raise
// it (after a frame that ended in an error before its result was read, the
// stream is drained first).
int arm_icp_read(kfx_pipeline* p) {
    if (p->icp_in_flight) KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    icp_clear_done(p->icp);
    p->icp_in_flight = true;
    return KFX_OK;
}

int learned_frame(kfx_pipeline* p) {
    KFX_TRY(preprocess(p, true));
    // ICP, then the gate at the new pose (in the ICP launch where the frame
    // runs as one).  The gate's statistics go straight to pinned host
    // memory and its flag both to the device (read by the skips below) and
    // to the host: no copy commands after the graph.
    kfx::FrameGate gate;
    gate.vmap_model = p->vp[0];
    gate.depth = p->dep[0];
    gate.intr = p->lvl[0];
    gate.max_diff = 0.1f;
    gate.fallback_m = 1e-3 * (double)p->cfg.fallback_rms_mm;
    gate.stats = p->hstats;
    gate.flag = p->dgate;
    gate.host_flag = p->hgate;
    KFX_TRY(icp_part(p, true, &gate));
    // The raycast below overwrites vp[0]: the gate has read it already.
    return map_part_skip(p, p->dgate);
}

// The whole fused frame with the bilateral filter.
int bilateral_frame(kfx_pipeline* p) {
    KFX_TRY(preprocess(p, false));
    KFX_TRY(icp_part(p, true));
    return map_part(p);
}

void drop_graph(kfx_pipeline* p) {
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
}

int capture(kfx_pipeline* p, int (*body)(kfx_pipeline*), hipGraph_t* graph, hipGraphExec_t* exec) {
    KFX_CHECK_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    int rc = body(p);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream(), &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    KFX_CHECK_HIP(e);
    *graph = g;
    KFX_CHECK_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return KFX_OK;
}

// Runs body eagerly, or as a graph captured on first use.
int run_part(kfx_pipeline* p, int (*body)(kfx_pipeline*), hipGraph_t* graph, hipGraphExec_t* exec) {
    if (!p->cfg.use_graph) return body(p);
    if (!*exec) KFX_TRY(capture(p, body, graph, exec));
    KFX_CHECK_HIP(hipGraphLaunch(*exec, stream()));
    return KFX_OK;
}

// Cyclic shift (KinectFusion large-scale convention): the tracked point is
// the camera's target, shift_target_distance metres along its optical axis
// (the camera centre itself when 0).  When it is more than shift_threshold
// from the cube's centre, the cube moves by whole voxels so that its centre
// comes to the target.  The slab leaving the cube (black zone, in unwrapped
// grid coordinates) is extracted, then cleared; origin_GRID advances modulo
// N; the pose and the world origin are re-expressed.  Voxels keep their
// memory cells.
int maybe_shift(kfx_pipeline* p, kfx_frame_info* info) {
    const double cx = 0.5 * p->rb.volume_size.x, cy = 0.5 * p->rb.volume_size.y, cz = 0.5 * p->rb.volume_size.z;
    const double f = p->cfg.shift_target_distance;
    const double tx = p->t[0] + f * p->R[2], ty = p->t[1] + f * p->R[5], tz = p->t[2] + f * p->R[8];
    const double dx = tx - cx, dy = ty - cy, dz = tz - cz;
    if (sqrt(dx * dx + dy * dy + dz * dz) <= p->cfg.shift_threshold) return KFX_OK;
    const int n[3] = {p->rb.voxels_size.x, p->rb.voxels_size.y, p->rb.voxels_size.z};
    const double cs[3] = {p->cell.x, p->cell.y, p->cell.z};
    const double d[3] = {dx, dy, dz};
    int s[3];
    for (int k = 0; k < 3; ++k) s[k] = std::max(-n[k], std::min(n[k], (int)trunc(d[k] / cs[k])));
    const int o[3] = {p->rb.origin_GRID.x, p->rb.origin_GRID.y, p->rb.origin_GRID.z};
    int mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        if (s[k] > 0) { mn[k] = o[k]; mx[k] = o[k] + s[k] - 1; }
        else if (s[k] < 0) { mn[k] = o[k] + n[k] + s[k]; mx[k] = o[k] + n[k] - 1; }
        else { mn[k] = 1; mx[k] = 0; }
    }
    const kfx_point_3i pmn{mn[0], mn[1], mn[2]}, pmx{mx[0], mx[1], mx[2]};
    if (p->cfg.extract_on_shift) {
        uint64_t count = 0;
        // Records and XYZI points are both 16 B: one buffer serves either.
        static_assert(sizeof(kfx_slice_record) == sizeof(kfx_xyzi), "record size");
        auto run = [&]() {
            if (p->cfg.extract_records)
                return kfx_extract_slice_records(p->slicer, &pmn, &pmx, &p->rb, p->vol, p->cvol,
                                                 reinterpret_cast<kfx_slice_record*>(p->ext_pts), p->ext_col,
                                                 p->ext_cap, &count);
            return kfx_extract_slice(p->slicer, &pmn, &pmx, &p->rb, p->vol, p->cvol, p->ext_pts, p->ext_col,
                                     p->ext_cap, &count);
        };
        int rc = run();
        if (rc == KFX_ERR_CAPACITY) {
            (void)hipFree(p->ext_pts);
            (void)hipFree(p->ext_col);
            p->ext_cap = count + count / 4 + 1024;
            KFX_CHECK_HIP(hipMalloc((void**)&p->ext_pts, sizeof(kfx_xyzi) * p->ext_cap));
            KFX_CHECK_HIP(hipMalloc((void**)&p->ext_col, sizeof(kfx_color) * p->ext_cap));
            rc = run();
        }
        KFX_TRY(rc);
        p->total_extracted += count;
        if (info) info->extracted_points = count;
    }
    KFX_TRY(kfx_clean_tsdf_slice(&pmn, &pmx, &p->rb, p->vol));
    if (p->cvol) KFX_TRY(kfx_clean_color_slice(&pmn, &pmx, &p->rb, p->cvol));
    KFX_TRY(step_sync(p));
    int* og[3] = {&p->rb.origin_GRID.x, &p->rb.origin_GRID.y, &p->rb.origin_GRID.z};
    float* ogg[3] = {&p->rb.origin_GRID_global.x, &p->rb.origin_GRID_global.y, &p->rb.origin_GRID_global.z};
    float* om[3] = {&p->rb.origin_metric.x, &p->rb.origin_metric.y, &p->rb.origin_metric.z};
    for (int k = 0; k < 3; ++k) {
        *og[k] = ((o[k] + s[k]) % n[k] + n[k]) % n[k];
        *ogg[k] += (float)s[k];
        p->origin_world[k] += s[k] * cs[k];
        *om[k] = (float)p->origin_world[k];
        p->t[k] -= s[k] * cs[k];
    }
    if (p->cfg.mode == KFX_PIPELINE_FUSED)
        KFX_TRY(kfx_icp_translate(p->icp, -s[0] * cs[0], -s[1] * cs[1], -s[2] * cs[2]));
    // The brick flags are in local voxel coordinates, which moved with the
    // origin: rebuilt from the (cleaned) volume.
    if (p->bricks) KFX_TRY(brick_flags_build(p->vol, &p->rb, p->bricks));
    // The previous-frame maps must be in the new volume coordinates.
    KFX_TRY(raycast_host_pose(p));
    drop_graph(p);
    p->shifts += 1;
    if (info) info->shifted = 1;
    return KFX_OK;
}

}  // namespace

extern "C" {

int kfx_pipeline_default_config(kfx_pipeline_config* c) {
    KFX_REQUIRE(c != nullptr, "kfx_pipeline_default_config: null");
    memset(c, 0, sizeof(*c));
    c->intr = kfx_intrinsics{525.0f, 525.0f, 319.5f, 239.5f, 640, 480};
    c->voxels = kfx_point_3i{512, 512, 512};
    c->volume_size = kfx_point_3f{3.072f, 3.072f, 3.072f};
    c->truncation = KFX_DEFAULT_TRUNCATION;
    c->sigma_space = KFX_DEFAULT_SIGMA_SPACE;
    c->sigma_color = KFX_DEFAULT_SIGMA_COLOR;
    c->distance_threshold = KFX_DEFAULT_DIST_THRESHOLD;
    c->angle_threshold = (float)sin(20.0 * M_PI / 180.0);
    c->icp_iterations[0] = 10;
    c->icp_iterations[1] = 5;
    c->icp_iterations[2] = 4;
    c->mode = KFX_PIPELINE_PLAIN;
    c->use_graph = 0;
    c->denoiser = 0;
    c->fallback_rms_mm = 15.0f;
    c->shift_threshold = KFX_DEFAULT_SHIFT_THRESHOLD;
    c->shift_target_distance = 1.5;
    c->integrate_threshold = 0.001f;
    c->converge_rotation = 0.0;
    c->converge_translation = 0.0;
    c->extract_on_shift = 1;
    c->colors = 0;
    c->color_max_weight = KFX_DEFAULT_COLOR_MAX_WEIGHT;
    c->icp_launch = KFX_ICP_LAUNCH_FRAME;
    c->icp_sum_mode = KFX_ICP_SUM_FP64;
    return KFX_OK;
}

int kfx_pipeline_destroy(kfx_pipeline* p) {
    if (!p) return KFX_OK;
    KFX_REQUIRE_INIT();
    (void)hipStreamSynchronize(stream());
    drop_graph(p);
    (void)hipFree(p->vol);
    (void)hipFree(p->cvol);
    (void)hipFree(p->rgb);
    (void)hipFree(p->raw);
    for (int l = 0; l < KFX_ICP_LEVELS; ++l) {
        (void)hipFree(p->dep[l]);
        (void)hipFree(p->vc[l]);
        (void)hipFree(p->nc[l]);
        (void)hipFree(p->vp[l]);
        (void)hipFree(p->np[l]);
    }
    (void)hipFree(p->scaled);
    (void)hipFree(p->gbuf);
    (void)hipFree(p->mbuf);
    (void)hipFree(p->ext_pts);
    (void)hipFree(p->ext_col);
    if (p->hstats) (void)hipHostFree(p->hstats);
    (void)hipFree(p->dgate);
    if (p->hgate) (void)hipHostFree(p->hgate);
    (void)hipFree(p->bricks);
    kfx_icp_destroy(p->icp);
    kfx_slicer_destroy(p->slicer);
    workspace_free(p->ws);
    delete p;
    return KFX_OK;
}

int kfx_pipeline_create(const kfx_pipeline_config* cfg, const kfx_pose* initial_pose_world,
                        const double volume_origin_world[3], kfx_pipeline** out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(cfg && initial_pose_world && volume_origin_world && out, "kfx_pipeline_create: null pointer");
    KFX_REQUIRE(cfg->mode == KFX_PIPELINE_PLAIN || cfg->mode == KFX_PIPELINE_FUSED, "kfx_pipeline_create: bad mode");
    KFX_REQUIRE(cfg->intr.cols >= 8 && cfg->intr.rows >= 8 && cfg->intr.cols % 4 == 0 && cfg->intr.rows % 4 == 0,
                "kfx_pipeline_create: image size must be a multiple of 4 (3 pyramid levels)");
    KFX_REQUIRE(cfg->voxels.x > 0 && cfg->voxels.y > 0 && cfg->voxels.z > 0 &&
                    (int64_t)cfg->voxels.x * cfg->voxels.y * cfg->voxels.z < (int64_t)1 << 31,
                "kfx_pipeline_create: voxel counts must be positive, below 2^31 in total");
    KFX_REQUIRE(!(cfg->mode == KFX_PIPELINE_PLAIN && cfg->use_graph), "kfx_pipeline_create: graphs need fused mode");
    KFX_REQUIRE(!cfg->colors || (cfg->color_max_weight >= 1 && cfg->color_max_weight <= 255),
                "kfx_pipeline_create: color_max_weight %d outside [1, 255]", cfg->color_max_weight);
    kfx_pipeline* p = new kfx_pipeline();
    p->cfg = *cfg;
    kfx_round_buffer& rb = p->rb;
    memset(&rb, 0, sizeof(rb));
    rb.voxels_size = cfg->voxels;
    rb.volume_size = cfg->volume_size;
    rb.buffer_length = cfg->voxels.x * cfg->voxels.y * cfg->voxels.z;
    for (int k = 0; k < 3; ++k) p->origin_world[k] = volume_origin_world[k];
    rb.origin_metric = kfx_point_3f{(float)p->origin_world[0], (float)p->origin_world[1], (float)p->origin_world[2]};
    p->cell = kfx_point_3f{cfg->volume_size.x / cfg->voxels.x, cfg->volume_size.y / cfg->voxels.y,
                           cfg->volume_size.z / cfg->voxels.z};
    for (int l = 0; l < KFX_ICP_LEVELS; ++l) {
        const float d = (float)(1 << l);
        p->lvl[l] = kfx_intrinsics{cfg->intr.fx / d, cfg->intr.fy / d, cfg->intr.cx / d, cfg->intr.cy / d,
                                   cfg->intr.cols >> l, cfg->intr.rows >> l};
    }
    const size_t npix = (size_t)cfg->intr.cols * cfg->intr.rows;
    hipError_t e = hipMalloc((void**)&p->vol, sizeof(kfx_tsdf) * (size_t)rb.buffer_length);
    if (e == hipSuccess) e = hipMalloc((void**)&p->raw, 2 * npix);
    if (e == hipSuccess) e = hipMalloc((void**)&p->scaled, 4 * npix);
    if (e == hipSuccess && cfg->colors) e = hipMalloc((void**)&p->cvol, sizeof(kfx_color) * (size_t)rb.buffer_length);
    if (e == hipSuccess && cfg->colors) e = hipMalloc((void**)&p->rgb, sizeof(kfx_color) * npix);
    for (int l = 0; l < KFX_ICP_LEVELS && e == hipSuccess; ++l) {
        const size_t n = (size_t)p->lvl[l].cols * p->lvl[l].rows;
        e = hipMalloc((void**)&p->dep[l], 2 * n);
        if (e == hipSuccess) e = hipMalloc((void**)&p->vc[l], sizeof(kfx_xyz) * n);
        if (e == hipSuccess) e = hipMalloc((void**)&p->nc[l], sizeof(kfx_xyz) * n);
        if (e == hipSuccess) e = hipMalloc((void**)&p->vp[l], sizeof(kfx_xyz) * n);
        if (e == hipSuccess) e = hipMalloc((void**)&p->np[l], sizeof(kfx_xyz) * n);
    }
    p->gbuf_stride = (int)(divup(cfg->intr.cols, 32) * divup(cfg->intr.rows, 8));
    if (e == hipSuccess) e = hipMalloc((void**)&p->gbuf, sizeof(float) * 27 * (size_t)p->gbuf_stride);
    if (e == hipSuccess) e = hipMalloc((void**)&p->mbuf, sizeof(float) * 27);
    p->ext_cap = 1u << 20;
    if (e == hipSuccess) e = hipMalloc((void**)&p->ext_pts, sizeof(kfx_xyzi) * p->ext_cap);
    if (e == hipSuccess) e = hipMalloc((void**)&p->ext_col, sizeof(kfx_color) * p->ext_cap);
    if (e == hipSuccess) e = hipHostMalloc((void**)&p->hstats, sizeof(double) * 4, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&p->dgate, sizeof(int32_t));
    if (e == hipSuccess) e = hipHostMalloc((void**)&p->hgate, sizeof(int32_t), hipHostMallocDefault);
    // Brick flags for the skipping raycast in fused mode (plain mode marches
    // every sample, as the reference does); the volume starts empty, so no
    // brick holds a negative value.
    const size_t nbricks = brick_flag_bytes(rb.voxels_size);
    if (e == hipSuccess && cfg->mode == KFX_PIPELINE_FUSED && nbricks > 0) e = hipMalloc((void**)&p->bricks, nbricks);
    if (e == hipSuccess && p->bricks) e = hipMemset(p->bricks, 0, nbricks);
    if (e != hipSuccess) {
        kfx_pipeline_destroy(p);
        return set_error(KFX_ERR_HIP, "kfx_pipeline_create: %s", hipGetErrorString(e));
    }
    int rc = kfx_icp_create((int)npix, &p->icp);
    if (!rc) rc = kfx_icp_set_sum_mode(p->icp, p->cfg.icp_sum_mode);
    if (!rc) rc = kfx_icp_set_launch_mode(p->icp, p->cfg.icp_launch);
    if (!rc) rc = kfx_slicer_create(&rb, &p->slicer);
    if (!rc) rc = workspace_reserve(p->ws, p->cfg.intr.cols, p->cfg.intr.rows, p->cvol != nullptr);
    if (!rc) rc = workspace_reserve_volume(p->ws, rb.voxels_size);
    if (!rc) rc = kfx_init_tsdf(p->vol, &rb);
    if (!rc && p->cvol) rc = kfx_init_colors(p->cvol, &rb);
    if (!rc) rc = kfx_synchronize();
    if (rc) {
        kfx_pipeline_destroy(p);
        return rc;
    }
    for (int i = 0; i < 9; ++i) p->R[i] = initial_pose_world->R[i];
    for (int k = 0; k < 3; ++k) p->t[k] = initial_pose_world->t[k] - p->origin_world[k];
    memcpy(p->Rprev, p->R, sizeof(p->R));
    memcpy(p->tprev, p->t, sizeof(p->t));
    *out = p;
    return KFX_OK;
}

static int process_frame(kfx_pipeline* p, const uint16_t* depth, const kfx_color* rgb, kfx_frame_info* info) {
    if (info) memset(info, 0, sizeof(*info));
    const size_t npix = (size_t)p->cfg.intr.cols * p->cfg.intr.rows;
    // The frame's kernels read the caller's depth buffer where they are
    // launched afresh each frame; a captured graph reads the pipeline's own
    // copy (its kernels' arguments are fixed).
    if (p->cfg.use_graph && depth != p->raw) {
        KFX_TRY(kfx_memcpy_dtod(p->raw, depth, 2 * npix));
        p->in = p->raw;
    } else {
        p->in = depth;
    }
    if (rgb != nullptr && rgb != p->rgb) KFX_TRY(kfx_memcpy_dtod(p->rgb, rgb, sizeof(kfx_color) * npix));
    kfx_icp_result r;
    bool have_pose = false;  // r holds the fused estimate of this frame

    if (p->frame == 0) {
        KFX_TRY(preprocess(p, use_denoiser(p)));
        KFX_TRY(integrate_host_pose(p));
        KFX_TRY(raycast_host_pose(p));
        KFX_TRY(colors_host_pose(p));
        if (p->cfg.mode == KFX_PIPELINE_FUSED) {
            kfx_pose pose;
            memcpy(pose.R, p->R, sizeof(p->R));
            memcpy(pose.t, p->t, sizeof(p->t));
            KFX_TRY(kfx_icp_reset(p->icp, &pose));
        }
        if (info) info->integrated = 1;
    } else if (p->cfg.mode == KFX_PIPELINE_PLAIN) {
        KFX_TRY(preprocess(p, false));
        KFX_TRY(icp_plain(p, info));
        const bool integ = motion(p->R, p->t, p->Rprev, p->tprev) >= p->cfg.integrate_threshold;
        if (integ) KFX_TRY(integrate_host_pose(p));
        KFX_TRY(raycast_host_pose(p));
        KFX_TRY(colors_host_pose(p));
        if (info) info->integrated = integ ? 1 : 0;
    } else if (!use_denoiser(p)) {
        KFX_TRY(arm_icp_read(p));
        KFX_TRY(run_part(p, bilateral_frame, &p->graph, &p->graph_exec));
    } else {
        // The learned frame as one graph; the gate's statistics and decision
        // are in pinned host memory when the pose read below returns (one
        // round trip per frame).
        // When the gate fell back, the device skipped integration and
        // raycast: the frame is redone with the bilateral filter, ICP from
        // the previous pose, then integration and raycast.
        KFX_TRY(arm_icp_read(p));
        KFX_TRY(run_part(p, learned_frame, &p->graph, &p->graph_exec));
        KFX_TRY(icp_read_frame(p, &r));  // waits for the ICP launch only
        have_pose = true;
        const double cnt = p->hstats[1];
        const double rms_mm = cnt > 0 ? 1000.0 * sqrt(p->hstats[0] / cnt) : 0.0;
        const bool fallback = *p->hgate != 0;
        if (info) {
            info->denoise_rms_mm = (float)rms_mm;
            info->fallback = fallback ? 1 : 0;
        }
        if (fallback) {
            kfx_pose prev;
            memcpy(prev.R, p->Rprev, sizeof(prev.R));
            memcpy(prev.t, p->tprev, sizeof(prev.t));
            KFX_TRY(kfx_icp_reset(p->icp, &prev));
            KFX_TRY(preprocess(p, false));
            KFX_TRY(arm_icp_read(p));
            KFX_TRY(icp_part(p, false));
            KFX_TRY(map_part(p));
            p->fallbacks += 1;
            have_pose = false;
        }
    }
    if (p->frame > 0 && p->cfg.mode == KFX_PIPELINE_FUSED) {
        if (!have_pose) KFX_TRY(icp_read_frame(p, &r));
        memcpy(p->R, r.pose.R, sizeof(p->R));
        memcpy(p->t, r.pose.t, sizeof(p->t));
        if (info) {
            info->icp_iterations = r.iterations;
            info->icp_singular = r.singular_count;
            info->integrated = motion(p->R, p->t, p->Rprev, p->tprev) >= p->cfg.integrate_threshold;
        }
    }
    if (info) {
        memcpy(info->pose_volume.R, p->R, sizeof(p->R));
        memcpy(info->pose_volume.t, p->t, sizeof(p->t));
    }
    KFX_TRY(maybe_shift(p, info));
    memcpy(p->Rprev, p->R, sizeof(p->R));
    memcpy(p->tprev, p->t, sizeof(p->t));
    if (info) {
        memcpy(info->pose_world.R, p->R, sizeof(p->R));
        for (int k = 0; k < 3; ++k) info->pose_world.t[k] = p->t[k] + p->origin_world[k];
    }
    p->frame += 1;
    return KFX_OK;
}

int kfx_pipeline_process(kfx_pipeline* p, const uint16_t* depth, kfx_frame_info* info) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(p && depth, "kfx_pipeline_process: null pointer");
    KFX_REQUIRE(!p->cvol, "kfx_pipeline_process: the pipeline keeps a colour volume; use kfx_pipeline_process_rgbd");
    return process_frame(p, depth, nullptr, info);
}

int kfx_pipeline_process_rgbd(kfx_pipeline* p, const uint16_t* depth, const kfx_color* rgb, kfx_frame_info* info) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(p && depth && rgb, "kfx_pipeline_process_rgbd: null pointer");
    KFX_REQUIRE(p->cvol, "kfx_pipeline_process_rgbd: the pipeline was created without colours (cfg.colors = 0)");
    return process_frame(p, depth, rgb, info);
}

int kfx_pipeline_set_denoiser(kfx_pipeline* p, kfx_denoiser* dn) {
    KFX_REQUIRE(p != nullptr, "kfx_pipeline_set_denoiser: null pipeline");
    KFX_REQUIRE(dn == nullptr || p->cfg.mode == KFX_PIPELINE_FUSED, "kfx_pipeline_set_denoiser: needs fused mode");
    p->denoiser = dn;
    drop_graph(p);
    return KFX_OK;
}

int kfx_pipeline_state(kfx_pipeline* p, kfx_pipeline_view* v) {
    KFX_REQUIRE(p && v, "kfx_pipeline_state: null pointer");
    KFX_REQUIRE_INIT();
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    v->volume = p->vol;
    v->round_buffer = p->rb;
    v->cell_size = p->cell;
    for (int k = 0; k < 3; ++k) v->origin_world[k] = p->origin_world[k];
    v->depth_filtered = p->dep[0];
    v->depth_raw = p->in != nullptr ? p->in : p->raw;
    v->vmap_model = p->vp[0];
    v->nmap_model = p->np[0];
    v->vmap_current = p->vc[0];
    v->nmap_current = p->nc[0];
    memcpy(v->pose_volume.R, p->R, sizeof(p->R));
    memcpy(v->pose_volume.t, p->t, sizeof(p->t));
    v->frames = p->frame;
    v->shifts = p->shifts;
    v->fallbacks = p->fallbacks;
    v->extracted_points = p->total_extracted;
    v->brick_flags = p->bricks;
    for (int l = 0; l < KFX_ICP_LEVELS; ++l) {
        v->vmap_model_levels[l] = p->vp[l];
        v->nmap_model_levels[l] = p->np[l];
    }
    v->depth_scaled = p->scaled;
    v->depth_bound = p->ws.bound;
    v->last_points = p->cfg.extract_records ? nullptr : p->ext_pts;
    v->last_records = p->cfg.extract_records ? reinterpret_cast<kfx_slice_record*>(p->ext_pts) : nullptr;
    v->last_colors = p->ext_col;
    v->color_volume = p->cvol;
    v->rgb = p->rgb;
    return KFX_OK;
}

}  // extern "C"
