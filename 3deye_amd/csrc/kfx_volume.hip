// kfx_volume.hip — the cyclic TSDF volume: address evaluation, reset and
// projective integration of one depth frame.
//
// Volume layout in HBM: kfx_tsdf {int16 value, int16 weight}, 4 B per voxel,
// x fastest then y then z after the cyclic offset (raycast.cl:29-47).  A wave
// of 64 lanes walks 64 neighbouring x columns, so every z step of a wave is
// one 256-byte contiguous segment; each lane sweeps a run of z for its
// column ("z-major sweep"), the reference's (Nx, Ny) launch with the z loop
// split over several workgroups so the chip has enough waves in flight.
#include <vector>

#include "kfx_common.h"

namespace kfx {

// Voxels of one column that a thread handles (grid.z = Nz / run): 32 up to
// 512 slices, 64 beyond.
constexpr int kIntegrateZRun = 32;
constexpr int kIntegrateZRunLarge = 64;
__host__ __device__ inline int integrate_zrun(int nz) { return nz > 512 ? kIntegrateZRunLarge : kIntegrateZRun; }
// Frustum side planes are pushed out by this many pixels and the clipped z
// range by one voxel on each side, so the clip never removes a voxel the
// per-voxel projection test would accept.
constexpr float kClipMarginPx = 2.0f;
constexpr int kIntegrateBlockX = 64;
constexpr int kIntegrateBlockY = 4;
// Voxels of a run whose loads are issued before any of them is updated.
constexpr int kIntegrateBatch = 8;

// The sweep's workgroups: (bx, by, bz) covers the columns x in [64 bx,
// 64 bx + 63], y in [4 by, 4 by + 3] over the z run bz.  The hardware deals
// consecutive workgroups round-robin to the XCDs, so the launch grid takes
// the y blocks fastest: each XCD then gets every eighth row band of every
// region of the volume.  With x fastest, XCD k would get the x slabs
// 64 (8 j + k) .. 64 (8 j + k) + 63 -- for a camera looking along x, a few
// XCDs would hold the whole frustum.
__host__ __device__ inline dim3 sweep_grid(int nx, int ny, int nz, int zrun) {
    return dim3(divup((unsigned)ny, kIntegrateBlockY), divup((unsigned)nx, kIntegrateBlockX), divup((unsigned)nz, (unsigned)zrun));
}
struct SweepBlock {
    int bx, by, bz;
    uint32_t linear;  // bx + nbx (by + nby bz): the index of the pre-pass's workgroup flags
};
__device__ __forceinline__ SweepBlock sweep_block() {
    SweepBlock b;
    b.by = blockIdx.x;
    b.bx = blockIdx.y;
    b.bz = blockIdx.z;
    b.linear = (uint32_t)b.bx + gridDim.y * ((uint32_t)b.by + gridDim.x * (uint32_t)b.bz);
    return b;
}

struct IntegrateParams {
    float r[9];          // rotation_inverse, row-major
    float tx, ty, tz;    // camera position in volume coordinates
    float csx, csy, csz; // cell size
    float fx, fy, cx, cy;
    int cols, rows;
    float trunc, inv_trunc;
    int nx, ny, nz;
    int ox, oy, oz;      // cyclic origin, each in [0, n)
    int zrun;            // voxels of a column per thread (integrate_zrun)
    uint32_t slice;      // nx * ny
    // Frustum planes through the camera centre in volume coordinates:
    // a voxel centre g (relative to the camera) can project into the image
    // only if dot(plane[k], g) >= 0 for every k (near, left, right, top, bottom).
    float plane[5][3];
    // Depth-bound pyramid (null: no culling): level l in 1..kBoundLevels
    // holds, per texel, the largest scaled depth over its 2^l x 2^l pixels,
    // at bound + bound_off[l]; level 0 is the scaled depth itself.
    const float* bound;
    int bound_off[kBoundLevels + 1];
    // Brick flags the raycast's skipping march reads (null: none kept): one
    // byte per 16^3 brick of local voxel coordinates (nbx x nby x ...),
    // set when the rule writes a negative value into the brick.
    uint8_t* bricks;
    int nbx, nby;
};

// The pose-dependent terms of IntegrateParams, kept apart so that a kernel
// taking its pose from device memory fills a small register-resident struct
// instead of rewriting its (large) parameter block, which the compiler would
// then place in scratch memory.
struct PoseTerms {
    float r[9];
    float tx, ty, tz;
    float plane[5][3];
};

__device__ __forceinline__ PoseTerms host_terms(const IntegrateParams& p) {
    PoseTerms q;
#pragma unroll
    for (int i = 0; i < 9; ++i) q.r[i] = p.r[i];
    q.tx = p.tx; q.ty = p.ty; q.tz = p.tz;
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) q.plane[k][j] = p.plane[k][j];
    return q;
}

// Range [lo, hi] of z in [z0, z1) whose voxel centres of column (gx, gy) lie
// inside every frustum plane, widened by one voxel; lo > hi when empty.
__device__ __forceinline__ void clip_column(const IntegrateParams& p, const PoseTerms& q, float gx, float gy,
                                            int z0, int z1, int& lo, int& hi) {
    float zl = (float)z0, zh = (float)(z1 - 1);
    bool empty = false;
    const float base_z = 0.5f * p.csz - q.tz;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const float a = q.plane[k][0] * gx + q.plane[k][1] * gy + q.plane[k][2] * base_z;
        const float b = q.plane[k][2] * p.csz;
        if (b > 0.0f) zl = fmaxf(zl, -a * __builtin_amdgcn_rcpf(b));
        else if (b < 0.0f) zh = fminf(zh, -a * __builtin_amdgcn_rcpf(b));
        else if (a < 0.0f) empty = true;
    }
    zl = fmaxf(zl, (float)z0 - 2.0f);
    zh = fminf(zh, (float)z1 + 2.0f);
    lo = max(z0, (int)floorf(zl) - 1);
    hi = min(z1 - 1, (int)ceilf(zh) + 1);
    if (empty || !(zl <= zh + 2.0f)) { lo = z1; hi = z0 - 1; }
}

__device__ __forceinline__ int wave_min(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// scale_depth (opencl_wrappers.cpp:37-43; kernel absent from the reference,
// KinectFusion convention): metres along the pixel's ray,
//   lambda = sqrt(((u-cx)/fx)^2 + ((v-cy)/fy)^2 + 1),  D = d_mm * lambda / 1000.
__global__ void __launch_bounds__(256) scale_depth_kernel(const uint16_t* __restrict__ depth,
                                                          float* __restrict__ scaled,
                                                          kfx_intrinsics intr) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y * blockDim.y + threadIdx.y;
    if (u >= intr.cols || v >= intr.rows) return;
    const float xl = ((float)u - intr.cx) / intr.fx;
    const float yl = ((float)v - intr.cy) / intr.fy;
    const float lambda = sqrtf(xl * xl + yl * yl + 1.0f);
    const int idx = v * intr.cols + u;
    scaled[idx] = (float)depth[idx] * lambda / 1000.0f;
}

int scale_depth_launch(const uint16_t* depth, float* depth_scaled, const kfx_intrinsics* intr);

// One voxel of the update rule.  Returns true when the voxel's stored value
// changes and has to be written back.
__device__ __forceinline__ bool integrate_voxel(const IntegrateParams& p, const PoseTerms& q, float gx, float gy,
                                                float part, int z, const float* __restrict__ scaled,
                                                float& out_dp, bool& in_band) {
    const float gz = ((float)z + 0.5f) * p.csz - q.tz;
    const float camx = q.r[0] * gx + q.r[1] * gy + q.r[2] * gz;
    const float camy = q.r[3] * gx + q.r[4] * gy + q.r[5] * gz;
    const float camz = q.r[6] * gx + q.r[7] * gy + q.r[8] * gz;
    in_band = false;
    if (!(camz >= kRcpMin && camz < kRcpMax)) return false;
    const float inv_z = rcp_rn(camz);
    const float fu = rintf(camx * p.fx * inv_z + p.cx);
    const float fv = rintf(camy * p.fy * inv_z + p.cy);
    if (!(fu >= 0.0f && fu < (float)p.cols && fv >= 0.0f && fv < (float)p.rows)) return false;
    const float dp = scaled[(int)fv * p.cols + (int)fu];
    if (dp == 0.0f) return false;
    const float sdf = dp - sqrtf(gz * gz + part);
    if (!(sdf >= -p.trunc)) return false;
    out_dp = fminf(1.0f, sdf * p.inv_trunc);
    in_band = true;
    return true;
}

// Fills the pose-dependent terms q from a device-resident pose (R
// camera -> volume row-major, then t: 12 floats); false when the camera moved
// less than motion_thr since the previous pose (R^T, t), the KinectFusion
// integration metric (|rotation angle| + |translation|) / 2.
__device__ __forceinline__ bool load_device_pose(const IntegrateParams& p, PoseTerms& q, const float* __restrict__ dpose,
                                                 const float* __restrict__ dprev, float motion_thr) {
    float R[9], t[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = dpose[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) t[i] = dpose[9 + i];
    if (dprev != nullptr && motion_thr > 0.0f) {
        float tr = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) tr += R[3 * i + j] * dprev[3 * j + i];
        const float ang = acosf(fminf(1.0f, fmaxf(-1.0f, 0.5f * (tr - 1.0f))));
        const float dx = t[0] - dprev[9], dy = t[1] - dprev[10], dz = t[2] - dprev[11];
        if (0.5f * (ang + sqrtf(dx * dx + dy * dy + dz * dz)) < motion_thr) return false;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) q.r[3 * r + c] = R[3 * c + r];
    q.tx = t[0]; q.ty = t[1]; q.tz = t[2];
    const float m = kClipMarginPx;
    const float cl = p.cx + 0.5f + m, cr = (float)p.cols - 0.5f + m - p.cx;
    const float ct = p.cy + 0.5f + m, cb = (float)p.rows - 0.5f + m - p.cy;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        q.plane[0][k] = q.r[6 + k];
        q.plane[1][k] = p.fx * q.r[k] + cl * q.r[6 + k];
        q.plane[2][k] = -p.fx * q.r[k] + cr * q.r[6 + k];
        q.plane[3][k] = p.fy * q.r[3 + k] + ct * q.r[6 + k];
        q.plane[4][k] = -p.fy * q.r[3 + k] + cb * q.r[6 + k];
    }
    return true;
}

// Frustum test of the integration workgroup (bx, by, bz) -- voxel columns x
// in [64 bx, 64 bx + 63], y in [4 by, 4 by + 3], z in one run of p.zrun:
// 0 when for some plane every voxel centre of the box lies outside it (the
// largest value over the box's corners, taken per axis, stays below the
// plane), 2 when every centre lies inside every plane (the smallest value
// reaches it), 1 otherwise.  Both decisions carry a tolerance far above the
// rounding of these sums, so 0 is conservative (a voxel the update rule
// accepts projects into the image, so it lies inside every plane, which
// carry a margin of kClipMarginPx pixels) and 2 means that the per-column
// clip would keep the whole run.
__device__ __forceinline__ int block_frustum_class(const IntegrateParams& p, const PoseTerms& q, int bx, int by,
                                                   int bz) {
    const int x0 = bx * kIntegrateBlockX, x1 = min(x0 + kIntegrateBlockX, p.nx) - 1;
    const int y0 = by * kIntegrateBlockY, y1 = min(y0 + kIntegrateBlockY, p.ny) - 1;
    const int z0 = bz * p.zrun, z1 = min(z0 + p.zrun, p.nz) - 1;
    const float gx0 = ((float)x0 + 0.5f) * p.csx - q.tx, gx1 = ((float)x1 + 0.5f) * p.csx - q.tx;
    const float gy0 = ((float)y0 + 0.5f) * p.csy - q.ty, gy1 = ((float)y1 + 0.5f) * p.csy - q.ty;
    const float gz0 = ((float)z0 + 0.5f) * p.csz - q.tz, gz1 = ((float)z1 + 0.5f) * p.csz - q.tz;
    const float ex = fmaxf(fabsf(gx0), fabsf(gx1)), ey = fmaxf(fabsf(gy0), fabsf(gy1)), ez = fmaxf(fabsf(gz0), fabsf(gz1));
    bool all_inside = true;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const float n0 = q.plane[k][0], n1 = q.plane[k][1], n2 = q.plane[k][2];
        const float hi = fmaxf(n0 * gx0, n0 * gx1) + fmaxf(n1 * gy0, n1 * gy1) + fmaxf(n2 * gz0, n2 * gz1);
        const float lo = fminf(n0 * gx0, n0 * gx1) + fminf(n1 * gy0, n1 * gy1) + fminf(n2 * gz0, n2 * gz1);
        const float tol = 1e-4f * (fabsf(n0) * ex + fabsf(n1) * ey + fabsf(n2) * ez) +
                          1e-3f * (fabsf(n0) + fabsf(n1) + fabsf(n2));
        if (hi < -tol) return 0;
        all_inside = all_inside && lo > tol;
    }
    return all_inside ? 2 : 1;
}

// Per-frame pre-pass of the integration, one launch:
//  - scaled depth and the depth-bound pyramid: one 32 x 32 pixel tile per
//    workgroup writes its pixels' scaled depth (every pixel belongs to one
//    tile) and builds levels 1..5 by repeated 2 x 2 maxima in LDS; pixels
//    outside the image count as 0 (no depth);
//  - with a device pose, the pose terms of the frame and its skip decision
//    (the gate's flag, or too little motion), written once by thread 0;
//  - one flag per integration workgroup: 0 when the frame is skipped or the
//    workgroup's box lies outside the frustum, so those workgroups exit at
//    once; 2 when it lies inside, so its columns need no clip; 1 otherwise
//    (threads take the workgroups in turn over the whole grid).
template <bool kDevicePose>
__global__ void __launch_bounds__(kBoundTile * kBoundTile)
integrate_prepass_kernel(const uint16_t* __restrict__ depth, float* __restrict__ scaled, kfx_intrinsics intr,
                         float* __restrict__ bound, IntegrateParams p, const float* __restrict__ dpose,
                         const float* __restrict__ dprev, float motion_thr, const int32_t* __restrict__ dskip,
                         PoseTerms* __restrict__ terms, int32_t* __restrict__ frame_skip,
                         uint8_t* __restrict__ flags, int nbx, int nby, int nbz, int pixels) {
    __shared__ float t[kBoundTile * kBoundTile];
    const int cols = intr.cols, rows = intr.rows;
    const int tx = threadIdx.x % kBoundTile, ty = threadIdx.x / kBoundTile;
    const int u = blockIdx.x * kBoundTile + tx, v = blockIdx.y * kBoundTile + ty;
    float d = 0.0f;
    if (pixels && u < cols && v < rows) {
        d = scaled_depth(depth, intr, u, v);
        scaled[v * cols + u] = d;
    }
    if (pixels && bound != nullptr) {
        t[threadIdx.x] = d;
        __syncthreads();
        for (int l = 1; l <= kBoundLevels; ++l) {
            const int n = kBoundTile >> l;  // texels per tile side at level l
            float m = 0.0f;
            if (tx < n && ty < n) {
                const int sd = kBoundTile >> (l - 1);  // side of the previous level's tile in t[]
                m = fmaxf(fmaxf(t[(2 * ty) * sd + 2 * tx], t[(2 * ty) * sd + 2 * tx + 1]),
                          fmaxf(t[(2 * ty + 1) * sd + 2 * tx], t[(2 * ty + 1) * sd + 2 * tx + 1]));
            }
            __syncthreads();
            if (tx < n && ty < n) {
                t[ty * n + tx] = m;
                const int lw = (cols + (1 << l) - 1) >> l, lh = (rows + (1 << l) - 1) >> l;
                const int a = blockIdx.x * n + tx, b = blockIdx.y * n + ty;
                if (a < lw && b < lh) bound[p.bound_off[l] + b * lw + a] = m;
            }
            __syncthreads();
        }
    }
    if (flags == nullptr) return;
    const int nthreads = gridDim.x * gridDim.y * blockDim.x;
    const int gid = (blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
    const int nblk = nbx * nby * nbz;
    if (gid >= nblk) return;
    PoseTerms q;
    bool skip = false;
    if (kDevicePose) {
        skip = (dskip != nullptr && *dskip != 0) || !load_device_pose(p, q, dpose, dprev, motion_thr);
        if (gid == 0) {
            if (!skip) *terms = q;
            *frame_skip = skip ? 1 : 0;
        }
    } else {
        q = host_terms(p);
    }
    for (int g = gid; g < nblk; g += nthreads) {
        const int bx = g % nbx, by = (g / nbx) % nby, bz = g / (nbx * nby);
        flags[g] = skip ? 0 : (uint8_t)block_frustum_class(p, q, bx, by, bz);
    }
}

// Pixel footprint of the column segment z in [za, zb]: the box spanned by
// the projections of its end voxels' centres, widened by one pixel, and the
// smallest pyramid level l (texels of 2^l x 2^l pixels) whose 2 x 2 texels at
// a0..a1, b0..b1 cover it.
struct Footprint {
    int l;
    int u0, u1, v0, v1;  // pixels (level 0)
    float gza, gzb;      // z of the end voxels' centres relative to the camera
};
enum { kSegmentOutside = 0, kSegmentUndecided = 1, kSegmentFootprint = 2 };

// kSegmentOutside when no voxel centre of the segment can project into the
// image (all behind the camera, or the footprint misses the image);
// kSegmentUndecided when the test cannot tell (a centre too close to the
// camera plane, or a footprint wider than the coarsest level).
__device__ __forceinline__ int segment_footprint(const IntegrateParams& p, const PoseTerms& q, float cxy0,
                                                 float cxy1, float cxy2, int za, int zb, Footprint& f) {
    f.gza = ((float)za + 0.5f) * p.csz - q.tz;
    f.gzb = ((float)zb + 0.5f) * p.csz - q.tz;
    const float ax = cxy0 + q.r[2] * f.gza, ay = cxy1 + q.r[5] * f.gza, az = cxy2 + q.r[8] * f.gza;
    const float bx = cxy0 + q.r[2] * f.gzb, by = cxy1 + q.r[5] * f.gzb, bz = cxy2 + q.r[8] * f.gzb;
    if (az <= 0.0f && bz <= 0.0f) return kSegmentOutside;
    if (!(az > 1e-4f && bz > 1e-4f)) return kSegmentUndecided;
    const float ia = __builtin_amdgcn_rcpf(az), ib = __builtin_amdgcn_rcpf(bz);
    const float ua = ax * p.fx * ia + p.cx, ub = bx * p.fx * ib + p.cx;
    const float va = ay * p.fy * ia + p.cy, vb = by * p.fy * ib + p.cy;
    const float fu0 = floorf(fminf(ua, ub)) - 1.0f, fu1 = ceilf(fmaxf(ua, ub)) + 1.0f;
    const float fv0 = floorf(fminf(va, vb)) - 1.0f, fv1 = ceilf(fmaxf(va, vb)) + 1.0f;
    if (fu1 < 0.0f || fv1 < 0.0f || fu0 > (float)(p.cols - 1) || fv0 > (float)(p.rows - 1)) return kSegmentOutside;
    f.u0 = (int)fmaxf(fu0, 0.0f);
    f.u1 = (int)fminf(fu1, (float)(p.cols - 1));
    f.v0 = (int)fmaxf(fv0, 0.0f);
    f.v1 = (int)fminf(fv1, (float)(p.rows - 1));
    const int span = max(f.u1 - f.u0, f.v1 - f.v0) + 1;
    f.l = span <= 1 ? 0 : 32 - __clz(span - 1);  // 2^l >= span: at most 2 x 2 texels
    return f.l > kBoundLevels ? kSegmentUndecided : kSegmentFootprint;
}

// Offset of pyramid level l (1..kBoundLevels) without dynamic indexing of
// the parameter block.
__device__ __forceinline__ int level_offset(const IntegrateParams& p, int l) {
    int off = p.bound_off[1];
    off = l == 2 ? p.bound_off[2] : off;
    off = l == 3 ? p.bound_off[3] : off;
    off = l == 4 ? p.bound_off[4] : off;
    off = l == 5 ? p.bound_off[5] : off;
    return off;
}

// Per z of a workgroup's run: the z terms of the camera coordinates,
// r[2] gz, r[5] gz, r[8] gz, and gz^2, so that a voxel's camera coordinates
// (r0 gx + r1 gy) + r2 gz and |g|^2 = gz^2 + (gx^2 + gy^2) are one add each
// with the per-voxel evaluation's rounding; rows past the run's end keep the
// batch loop's overhang in bounds.
constexpr int kZTerms = kIntegrateZRunLarge + kIntegrateBatch;

// Largest scaled depth under the pixel footprint of the lane's column
// segment from camera point a to camera point b (its end voxels' centres):
// the box spanned by their projections widened by one pixel, looked up as
// 2 x 2 texels of the depth-bound pyramid level whose texels cover it.
// Returns 0 when no voxel centre of the segment can project into the image
// or no pixel under it has depth, a negative value when the test cannot
// decide (a centre near the camera plane, a footprint wider than the
// coarsest level).
__device__ __forceinline__ float segment_max_depth(const IntegrateParams& p, const float* __restrict__ scaled,
                                                   float ax, float ay, float az, float bx, float by, float bz) {
    if (az <= 0.0f && bz <= 0.0f) return 0.0f;
    if (!(az > 1e-4f && bz > 1e-4f)) return -1.0f;
    const float ia = __builtin_amdgcn_rcpf(az), ib = __builtin_amdgcn_rcpf(bz);
    const float ua = ax * p.fx * ia + p.cx, ub = bx * p.fx * ib + p.cx;
    const float va = ay * p.fy * ia + p.cy, vb = by * p.fy * ib + p.cy;
    const float fu0 = floorf(fminf(ua, ub)) - 1.0f, fu1 = ceilf(fmaxf(ua, ub)) + 1.0f;
    const float fv0 = floorf(fminf(va, vb)) - 1.0f, fv1 = ceilf(fmaxf(va, vb)) + 1.0f;
    if (fu1 < 0.0f || fv1 < 0.0f || fu0 > (float)(p.cols - 1) || fv0 > (float)(p.rows - 1)) return 0.0f;
    const int u0 = (int)fmaxf(fu0, 0.0f), u1 = (int)fminf(fu1, (float)(p.cols - 1));
    const int v0 = (int)fmaxf(fv0, 0.0f), v1 = (int)fminf(fv1, (float)(p.rows - 1));
    const int span = max(u1 - u0, v1 - v0) + 1;
    const int l = span <= 1 ? 0 : 32 - __clz(span - 1);  // 2^l >= span: at most 2 x 2 texels
    if (l > kBoundLevels) return -1.0f;
    if (l == 0) return scaled[v0 * p.cols + u0];
    const int lw = (p.cols + (1 << l) - 1) >> l;
    const float* b = p.bound + level_offset(p, l);
    const int a0 = u0 >> l, a1 = u1 >> l, b0 = v0 >> l, b1 = v1 >> l;
    return fmaxf(fmaxf(b[b0 * lw + a0], b[b0 * lw + a1]), fmaxf(b[b1 * lw + a0], b[b1 * lw + a1]));
}

// Distance bound of the update rule (dist <= depth + trunc), with a 1e-4
// relative margin that dwarfs the rounding of the squares compared with it.
__device__ __forceinline__ float reach(const IntegrateParams& p, float maxd) {
    return (maxd + p.trunc) * 1.0001f + 1e-4f;
}

// Buffer resource over one slice of the volume (bytes [0, slice_bytes) from
// base): the voxel loads and stores of a wave address one slice at a time
// with a 32-bit offset inside it (raw buffer, no swizzle).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slice_rsrc(const char* base, uint32_t slice_bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(base), (short)0, (int)slice_bytes, 0x00020000);
}

// (w, w + 1, RN(1 / (w + 1)), min(w + 1, KFX_MAX_TSDF_WEIGHT) << 16) for the
// running average from a voxel of weight w: the divisor's correctly rounded
// reciprocal comes from v_rcp_f32 and one Newton step, exact for every
// divisor up to 129.
struct WeightTerms {
    float fw, fb, y;
    uint32_t w_new_hi;  // min(w + 1, KFX_MAX_TSDF_WEIGHT) << 16, the stored word's weight half
};

// True when no voxel of the workgroup's box can be updated because every
// voxel centre lies farther from the camera than the largest depth under the
// box's footprint plus the truncation: the footprint is the pixel box of
// the projected corners (all in front of the camera), widened by one pixel,
// looked up as level-5 texels of the depth-bound pyramid (one per lane, at
// most 64), and the nearest point of the box comes from clamping the camera
// position per axis.  The inputs are uniform across the workgroup, so its
// waves agree.  Conservative: false whenever the test cannot decide.
__device__ __forceinline__ bool block_beyond_surface(const IntegrateParams& p, const PoseTerms& q, int bx, int by,
                                                     int bz) {
    if (p.bound == nullptr) return false;
    const int x0 = bx * kIntegrateBlockX, x1 = min(x0 + kIntegrateBlockX, p.nx) - 1;
    const int y0 = by * kIntegrateBlockY, y1 = min(y0 + kIntegrateBlockY, p.ny) - 1;
    const int z0 = bz * p.zrun, z1 = min(z0 + p.zrun, p.nz) - 1;
    const float g[2][3] = {{((float)x0 + 0.5f) * p.csx - q.tx, ((float)y0 + 0.5f) * p.csy - q.ty,
                            ((float)z0 + 0.5f) * p.csz - q.tz},
                           {((float)x1 + 0.5f) * p.csx - q.tx, ((float)y1 + 0.5f) * p.csy - q.ty,
                            ((float)z1 + 0.5f) * p.csz - q.tz}};
    float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float gx = g[c & 1][0], gy = g[(c >> 1) & 1][1], gz = g[c >> 2][2];
        const float cx = q.r[0] * gx + q.r[1] * gy + q.r[2] * gz;
        const float cy = q.r[3] * gx + q.r[4] * gy + q.r[5] * gz;
        const float cz = q.r[6] * gx + q.r[7] * gy + q.r[8] * gz;
        if (!(cz > 1e-3f)) return false;  // near or behind the camera plane: undecided
        const float iz = __builtin_amdgcn_rcpf(cz);
        const float u = cx * p.fx * iz + p.cx, v = cy * p.fy * iz + p.cy;
        umin = fminf(umin, u); umax = fmaxf(umax, u);
        vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
    }
    // The projection of the box lies in the hull of its corners' projections.
    const float fu0 = floorf(umin) - 1.0f, fu1 = ceilf(umax) + 1.0f;
    const float fv0 = floorf(vmin) - 1.0f, fv1 = ceilf(vmax) + 1.0f;
    if (fu1 < 0.0f || fv1 < 0.0f || fu0 > (float)(p.cols - 1) || fv0 > (float)(p.rows - 1)) return true;
    const int u0 = (int)fmaxf(fu0, 0.0f) >> kBoundLevels, u1 = (int)fminf(fu1, (float)(p.cols - 1)) >> kBoundLevels;
    const int v0 = (int)fmaxf(fv0, 0.0f) >> kBoundLevels, v1 = (int)fminf(fv1, (float)(p.rows - 1)) >> kBoundLevels;
    const int tw = u1 - u0 + 1, th = v1 - v0 + 1;
    if (tw * th > 64) return false;
    const int lane = threadIdx.x & 63;
    const int lw = (p.cols + kBoundTile - 1) >> kBoundLevels;
    float m = 0.0f;
    if (lane < tw * th) m = p.bound[p.bound_off[kBoundLevels] + (v0 + lane / tw) * lw + (u0 + lane % tw)];
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (m == 0.0f) return true;  // no depth under the footprint
    float d2 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float d = fmaxf(0.0f, fmaxf(g[0][a], -g[1][a]));
        d2 += d * d;
    }
    const float r = reach(p, m);
    return d2 > r * r;
}

// The TSDF sweep.  A workgroup is 64 x 4 columns over a run of p.zrun
// slices; a wave holds 64 neighbouring x columns, so each slice it visits
// is one 256-byte segment of the volume.
//
// Per lane: the column's run is clipped to the frustum (workgroups the
// pre-pass found inside the frustum keep the whole run), then bounded by
// distance: no voxel farther from the camera than the largest depth under
// the run's footprint plus the truncation can be updated, and |g| is convex
// along the column, so the remaining voxels form one z interval.  The wave
// walks the union of its lanes' ranges in batches of kIntegrateBatch
// slices.  Per batch: each lane bounds its part of the batch against the
// pyramid in the same way and the wave skips the batch when every lane is
// culled; otherwise every voxel of the batch is projected (the exact rule's
// arithmetic: correctly rounded 1/z, rint) and its scaled depth gathered,
// while the lanes that are not culled load their voxels of the batch; then
// the rule is applied.
//
// The distance |g| = sqrt(gz^2 + gx^2 + gy^2) is needed only near the
// surface: when |g|^2 < f |f| with f = 0.999999 dp - 1.01 trunc, then
// sdf = dp - |g| exceeds 1.01 trunc - 2e-7 dp (the rounding of the square,
// the root and the difference are each below 1.2e-7 of dp), so
// min(1, sdf / trunc) = 1 without computing it; otherwise the correctly
// rounded root decides, as in the oracle.
//
// The running average divides by w + 1 with its correctly rounded
// reciprocal from an LDS table and Markstein's correction: the IEEE
// quotient.
template <bool kDevicePose>
__global__ void __launch_bounds__(kIntegrateBlockX * kIntegrateBlockY)
integrate_kernel(const float* __restrict__ scaled, kfx_tsdf* __restrict__ volume, const IntegrateParams p,
                 const PoseTerms* __restrict__ dterms, const int32_t* __restrict__ frame_skip,
                 const uint8_t* __restrict__ flags) {
    __shared__ float4 zt[kZTerms];
    __shared__ float gzt[kZTerms];
    __shared__ WeightTerms qt[KFX_MAX_TSDF_WEIGHT + 1];
    const SweepBlock blk = sweep_block();
    int cls = 1;
    if (flags != nullptr) {
        cls = flags[blk.linear];
        if (cls == 0) return;  // uniform: the frame is skipped, or the workgroup lies outside the frustum
    }
    if (kDevicePose && *frame_skip != 0) return;
    PoseTerms q;
    if (kDevicePose) q = *dterms;
    else q = host_terms(p);
    if (block_beyond_surface(p, q, blk.bx, blk.by, blk.bz)) return;  // uniform across the workgroup
    const int tid = threadIdx.y * kIntegrateBlockX + threadIdx.x;
    const int z_begin = blk.bz * p.zrun;
    const int z_end = min(z_begin + p.zrun, p.nz);
    if (tid < kZTerms) {
        const float gz = ((float)(z_begin + tid) + 0.5f) * p.csz - q.tz;
        zt[tid] = make_float4(q.r[2] * gz, q.r[5] * gz, q.r[8] * gz, gz * gz);
        gzt[tid] = gz;
    }
    if (tid <= KFX_MAX_TSDF_WEIGHT) {
        const float fb = (float)(tid + 1);
        const float y0 = __builtin_amdgcn_rcpf(fb);
        WeightTerms t;
        t.fw = (float)tid;
        t.fb = fb;
        t.y = fmaf(fmaf(-fb, y0, 1.0f), y0, y0);
        t.w_new_hi = (uint32_t)min(tid + 1, KFX_MAX_TSDF_WEIGHT) << 16;
        qt[tid] = t;
    }
    __syncthreads();

    // Wave w of the workgroup takes its 16 x 4 columns x0 + 16 w .. + 15,
    // y0 .. y0 + 3 (not one row of 64): neighbouring columns' z ranges
    // differ less, so the range the wave walks (the union of its lanes')
    // is shorter.  A 16-lane group is one brick row.
    const int lane = tid & 63, wave = tid >> 6;
    const int x = blk.bx * kIntegrateBlockX + 16 * wave + (lane & 15);
    const int y = blk.by * kIntegrateBlockY + (lane >> 4);
    const bool inside = x < p.nx && y < p.ny;
    const float gx = ((float)x + 0.5f) * p.csx - q.tx;
    const float gy = ((float)y + 0.5f) * p.csy - q.ty;
    const float part = gx * gx + gy * gy;
    // Column terms of the camera coordinates: (r0 gx + r1 gy) + r2 gz.
    const float cxy0 = q.r[0] * gx + q.r[1] * gy;
    const float cxy1 = q.r[3] * gx + q.r[4] * gy;
    const float cxy2 = q.r[6] * gx + q.r[7] * gy;
    int lo = z_begin, hi = z_end - 1;
    if (cls != 2) clip_column(p, q, gx, gy, z_begin, z_end, lo, hi);
    if (inside && lo <= hi && p.bound != nullptr) {
        // Distance bound of the whole run.
        const float4 ta = zt[lo - z_begin], tb = zt[hi - z_begin];
        const float maxd = segment_max_depth(p, scaled, cxy0 + ta.x, cxy1 + ta.y, cxy2 + ta.z, cxy0 + tb.x,
                                             cxy1 + tb.y, cxy2 + tb.z);
        if (maxd == 0.0f) {
            lo = z_end;
            hi = z_begin - 1;
        } else if (maxd > 0.0f) {
            const float r = reach(p, maxd);
            const float room = r * r - part;
            if (!(room > 0.0f)) {
                lo = z_end;
                hi = z_begin - 1;
            } else {
                // |gz| <= sqrt(room): z in [(tz - s) / csz - 1/2, (tz + s) / csz - 1/2],
                // widened by a voxel for the rounding of gz.
                const float sq = sqrtf(room) * 1.0001f;
                const float zl = (q.tz - sq) / p.csz - 0.5f, zh = (q.tz + sq) / p.csz - 0.5f;
                lo = max(lo, (int)fmaxf(floorf(zl) - 1.0f, (float)z_begin - 1.0f));
                hi = min(hi, (int)fminf(ceilf(zh) + 1.0f, (float)z_end));
            }
        }
    }
    if (!inside) { lo = z_end; hi = z_begin - 1; }
    const int own_lo = lo, own_hi = hi;
    // The wave's range, in scalar registers: the batch loop, the slice
    // addresses and the z tests are then scalar work.
    lo = __builtin_amdgcn_readfirstlane(wave_min(lo));
    hi = __builtin_amdgcn_readfirstlane(wave_max(hi));
    if (!inside) return;

    const float fcols = (float)p.cols, frows = (float)p.rows;
    const float free_c = 1.01f * p.trunc;
    const uint32_t col4 = 4u * ((uint32_t)wrap_once(x + p.ox, p.nx) + (uint32_t)p.nx * (uint32_t)wrap_once(y + p.oy, p.ny));
    const uint64_t slice4 = 4ull * p.slice;
    const uint32_t slice_bytes = 4u * p.slice;
    const uint32_t nz = (uint32_t)p.nz;
    const char* __restrict__ sbase = reinterpret_cast<const char*>(scaled);
    char* __restrict__ vbase = reinterpret_cast<char*>(volume);
    for (int z0 = lo; z0 <= hi; z0 += kIntegrateBatch) {
        const int za = max(z0, own_lo), zb = min(z0 + kIntegrateBatch - 1, own_hi);
        bool culled = za > zb;
        if (!culled && p.bound != nullptr) {
            const float4 ta = zt[za - z_begin], tb = zt[zb - z_begin];
            const float maxd = segment_max_depth(p, scaled, cxy0 + ta.x, cxy1 + ta.y, cxy2 + ta.z, cxy0 + tb.x,
                                                 cxy1 + tb.y, cxy2 + tb.z);
            if (maxd == 0.0f) {
                culled = true;
            } else if (maxd > 0.0f) {
                // Nearest point of the segment to the camera, squared.
                const float gza = gzt[za - z_begin], gzb = gzt[zb - z_begin];
                const float dz = (gza <= 0.0f && gzb >= 0.0f) ? 0.0f : fminf(fabsf(gza), fabsf(gzb));
                const float r = reach(p, maxd);
                culled = part + dz * dz > r * r;
            }
        }
        if (__all(culled)) continue;
        float dp[kIntegrateBatch], d2[kIntegrateBatch];
        bool inimg[kIntegrateBatch], act[kIntegrateBatch];
        uint32_t word[kIntegrateBatch];
        const char* zbase[kIntegrateBatch];
        // Phase 1: projection and depth gather of every voxel of the batch
        // (outside the image: pixel 0, masked afterwards); the voxel loads
        // of the lanes not culled.  The slice of each voxel is the same for
        // the whole wave.
        float4 tn = zt[z0 - z_begin];
#pragma unroll
        for (int k = 0; k < kIntegrateBatch; ++k) {
            const int z = z0 + k;
            const float4 t = tn;
            if (k + 1 < kIntegrateBatch) tn = zt[z + 1 - z_begin];
            const float camx = cxy0 + t.x;
            const float camy = cxy1 + t.y;
            const float camz = cxy2 + t.z;
            // In front of the camera: camz in [2^-126, 2^126), where rcp_rn is
            // the correctly rounded reciprocal.
            const float inv_z = rcp_rn(camz);
            const float fu = rintf(camx * p.fx * inv_z + p.cx);
            const float fv = rintf(camy * p.fy * inv_z + p.cy);
            const bool in = (z <= hi) & (camz >= kRcpMin) & (camz < kRcpMax) & (fu >= 0.0f) & (fu < fcols) &
                            (fv >= 0.0f) & (fv < frows);
            // v cols + u is an integer below 2^24 inside the image, exact in
            // one fma.
            const float pixf = fmaf(fv, fcols, fu);
            const uint32_t pix = in ? (uint32_t)pixf : 0u;
            dp[k] = *reinterpret_cast<const float*>(sbase + 4u * pix);
            d2[k] = t.w + part;
            inimg[k] = in;
            uint32_t rz = (uint32_t)z + (uint32_t)p.oz;
            rz = min(rz, rz - nz);  // (z + oz) mod nz for z + oz < 2 nz (uniform)
            zbase[k] = vbase + slice4 * rz;
            act[k] = !culled & (z >= own_lo) & (z <= own_hi);
            word[k] = 0u;
            if (act[k]) word[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(slice_rsrc(zbase[k], slice_bytes), col4, 0, 0);
        }
        // Phase 2: the update test, the running average, and the write of
        // what changed.
#pragma unroll
        for (int k = 0; k < kIntegrateBatch; ++k) {
            const bool valid = act[k] & inimg[k] & (dp[k] != 0.0f);
            const float f = dp[k] * 0.999999f - free_c;
            const bool free_space = d2[k] < f * fabsf(f);
            float tsdf_new = 1.0f;
            bool upd = valid & free_space;
            const bool near = valid & !free_space;
            if (__builtin_amdgcn_ballot_w64(near) != 0) {
                // Near the surface: the rule with the correctly rounded root.
                const float sdf = dp[k] - sqrt_rn_in_range(d2[k]);
                if (near && sdf >= -p.trunc) {
                    upd = true;
                    tsdf_new = fminf(1.0f, sdf * p.inv_trunc);
                }
            }
            if (!upd) continue;
            const int16_t v_old = (int16_t)(word[k] & 0xffffu);
            // Weights outside [0, KFX_MAX_TSDF_WEIGHT] (a volume not written
            // by this rule) count as the nearest bound.
            const int w = max(0, min((int)(int16_t)(word[k] >> 16), KFX_MAX_TSDF_WEIGHT));
            const WeightTerms bq = qt[w];
            const float a = tsdf_decode(v_old) * bq.fw + tsdf_new;
            const float q0 = a * bq.y;
            const float avg = fmaf(fmaf(-q0, bq.fb, a), bq.y, q0);
            const int16_t v_new = tsdf_encode_unit(avg);
            const uint32_t out = bq.w_new_hi | (uint32_t)(uint16_t)v_new;
            if (out != word[k]) __builtin_amdgcn_raw_buffer_store_b32(out, slice_rsrc(zbase[k], slice_bytes), col4, 0, 0);
            if (p.bricks != nullptr) {
                // The first lane of each 16-column group (one brick) that
                // writes a negative value sets the brick's flag.
                const bool neg = v_new < 0;
                const uint64_t m = __builtin_amdgcn_ballot_w64(neg);
                const int lane = threadIdx.x & 63;
                const uint32_t grp = (uint32_t)(m >> (lane & ~15)) & 0xffffu;
                if (neg && (lane & 15) == __builtin_ctz(grp)) {
                    const uint32_t z = (uint32_t)(z0 + k);
                    p.bricks[((z >> 4) * (uint32_t)p.nby + ((uint32_t)y >> 4)) * (uint32_t)p.nbx + ((uint32_t)x >> 4)] = 1;
                }
            }
        }
    }
}

// Read-only pass of the update rule: counts the voxels integrate_kernel
// updates (one 64-bit atomic per wave).
__global__ void __launch_bounds__(kIntegrateBlockX * kIntegrateBlockY)
integrate_count_kernel(const float* __restrict__ scaled, IntegrateParams p, unsigned long long* __restrict__ count) {
    const SweepBlock blk = sweep_block();
    const int x = blk.bx * kIntegrateBlockX + threadIdx.x;
    const int y = blk.by * kIntegrateBlockY + threadIdx.y;
    unsigned n = 0;
    const PoseTerms q = host_terms(p);
    if (x < p.nx && y < p.ny) {
        const int z_begin = blk.bz * p.zrun;
        const int z_end = min(z_begin + p.zrun, p.nz);
        const float gx = ((float)x + 0.5f) * p.csx - q.tx;
        const float gy = ((float)y + 0.5f) * p.csy - q.ty;
        const float part = gx * gx + gy * gy;
        int lo, hi;
        clip_column(p, q, gx, gy, z_begin, z_end, lo, hi);
        for (int z = lo; z <= hi; ++z) {
            float tn;
            bool band;
            n += integrate_voxel(p, q, gx, gy, part, z, scaled, tn, band) ? 1u : 0u;
        }
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(count, (unsigned long long)n);
}

// ---------------------------------------------------------------------------
// Colour volume: Integrate_Colors (common/src/opencl_wrappers.cpp:181-214);
// its kernel integrate_color_volume is absent from the reference, whose
// arguments (RGB frame, raycast vertex map, colour volume, truncation,
// maximum weight, R^-1, t, intrinsics, cell size, round buffer) are those of
// the KinectFusion colour volume (PCL gpu/kinfu updateColorVolume).  Rule,
// per voxel of the cyclic volume:
//   voxel centre g = (i + 1/2) cell, camera point c = R^-1 (g - t), c.z > 0;
//   pixel (u, v) = rint(f c.xy / c.z + c0) inside the image (the projection
//   of the TSDF integration, so both pick the same pixel);
//   raycast vertex p at that pixel valid (x not NaN) and |p - g| < trunc;
//   then per channel rgb = rint_even((rgb * w + rgb_frame) / (w + 1)) in
//   FP32 clamped to [0, 255], w = min(w + 1, max_weight).
// The sweep is the TSDF integration's (z-major, frustum-clipped, batches of
// kIntegrateBatch); batches are culled with a pyramid of the raycast
// vertices' distances from the camera: an update needs
// | |g - t| - |p - t| | <= |p - g| < trunc, so a segment whose distance
// range misses [min |p - t| - trunc, max |p - t| + trunc] over its
// footprint has nothing to update.

// Colour-bound pyramid: one 32 x 32 pixel tile per workgroup; per texel of
// levels 0..kBoundLevels the smallest and largest distance from the camera
// to a valid raycast vertex under it (+inf and 0 when there is none).
__global__ void __launch_bounds__(kBoundTile * kBoundTile)
color_bound_kernel(const kfx_xyz* __restrict__ vmap, float2* __restrict__ bound, IntegrateParams p,
                   const float* __restrict__ dpose, const int32_t* __restrict__ dskip) {
    if (dskip != nullptr && *dskip != 0) return;  // uniform across the grid
    float tx = p.tx, ty = p.ty, tz = p.tz;
    if (dpose != nullptr) { tx = dpose[9]; ty = dpose[10]; tz = dpose[11]; }
    __shared__ float2 t[kBoundTile * kBoundTile];
    const int lx = threadIdx.x % kBoundTile, ly = threadIdx.x / kBoundTile;
    const int u = blockIdx.x * kBoundTile + lx, v = blockIdx.y * kBoundTile + ly;
    float2 m = make_float2(INFINITY, 0.0f);
    if (u < p.cols && v < p.rows) {
        const kfx_xyz q = vmap[v * p.cols + u];
        if (!isnan(q.x)) {
            const float dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
            const float r = sqrtf(dx * dx + dy * dy + dz * dz);
            m = make_float2(r, r);  // NaN when y or z is NaN: such a vertex never updates
        }
        bound[v * p.cols + u] = m;
    }
    t[threadIdx.x] = m;
    __syncthreads();
    for (int l = 1; l <= kBoundLevels; ++l) {
        const int n = kBoundTile >> l;
        float2 r = make_float2(INFINITY, 0.0f);
        if (lx < n && ly < n) {
            const int s = kBoundTile >> (l - 1);
            const float2 a = t[(2 * ly) * s + 2 * lx], b = t[(2 * ly) * s + 2 * lx + 1];
            const float2 c = t[(2 * ly + 1) * s + 2 * lx], d = t[(2 * ly + 1) * s + 2 * lx + 1];
            r = make_float2(fminf(fminf(a.x, b.x), fminf(c.x, d.x)), fmaxf(fmaxf(a.y, b.y), fmaxf(c.y, d.y)));
        }
        __syncthreads();
        if (lx < n && ly < n) {
            t[ly * n + lx] = r;
            const int lw = (p.cols + (1 << l) - 1) >> l, lh = (p.rows + (1 << l) - 1) >> l;
            const int a = blockIdx.x * n + lx, b = blockIdx.y * n + ly;
            if (a < lw && b < lh) bound[p.bound_off[l] + b * lw + a] = r;
        }
        __syncthreads();
    }
}

// True when no voxel of the column segment z in [za, zb] can take a colour
// (see the bound above).  Conservative: false whenever the test cannot
// decide; a NaN bound never culls.
__device__ __forceinline__ bool color_segment_culled(const IntegrateParams& p, const PoseTerms& q,
                                                     const float2* __restrict__ cb, float gx, float gy, float cxy0,
                                                     float cxy1, float cxy2, int za, int zb) {
    Footprint f;
    const int k = segment_footprint(p, q, cxy0, cxy1, cxy2, za, zb, f);
    if (k != kSegmentFootprint) return k == kSegmentOutside;
    const int lw = (p.cols + (1 << f.l) - 1) >> f.l;
    const float2* b = cb + (f.l == 0 ? 0 : level_offset(p, f.l));
    const int a0 = f.u0 >> f.l, a1 = f.u1 >> f.l, b0 = f.v0 >> f.l, b1 = f.v1 >> f.l;
    const float2 e00 = b[b0 * lw + a0], e01 = b[b0 * lw + a1], e10 = b[b1 * lw + a0], e11 = b[b1 * lw + a1];
    const float rmin = fminf(fminf(e00.x, e01.x), fminf(e10.x, e11.x));
    const float rmax = fmaxf(fmaxf(e00.y, e01.y), fmaxf(e10.y, e11.y));
    if (rmax == 0.0f) return true;  // no valid vertex under the footprint
    const float part = gx * gx + gy * gy;
    const float dz = (f.gza <= 0.0f && f.gzb >= 0.0f) ? 0.0f : fminf(fabsf(f.gza), fabsf(f.gzb));
    const float dmin = sqrt_rn(part + dz * dz);
    const float dmax = sqrt_rn(part + fmaxf(f.gza * f.gza, f.gzb * f.gzb));
    const float slack = p.trunc * 1.0001f + 1e-4f;
    return dmin > rmax * 1.0001f + slack || dmax < rmin * 0.9999f - slack;
}

template <bool kDevicePose>
__global__ void __launch_bounds__(kIntegrateBlockX * kIntegrateBlockY)
integrate_colors_kernel(const kfx_color* __restrict__ rgb, const kfx_xyz* __restrict__ vmap,
                        kfx_color* __restrict__ volume, const IntegrateParams p, const float2* __restrict__ cbound,
                        int max_weight, const float* __restrict__ dpose, const int32_t* __restrict__ dskip) {
    if (kDevicePose && dskip != nullptr && *dskip != 0) return;  // uniform across the grid
    PoseTerms q;
    if (kDevicePose) (void)load_device_pose(p, q, dpose, nullptr, 0.0f);
    else q = host_terms(p);
    const SweepBlock blk = sweep_block();
    const int x = blk.bx * kIntegrateBlockX + threadIdx.x;
    const int y = blk.by * kIntegrateBlockY + threadIdx.y;
    const bool inside = x < p.nx && y < p.ny;
    const int z_begin = blk.bz * p.zrun;
    const int z_end = min(z_begin + p.zrun, p.nz);

    const float vgx = ((float)x + 0.5f) * p.csx, vgy = ((float)y + 0.5f) * p.csy;
    const float gx = vgx - q.tx, gy = vgy - q.ty;
    const uint32_t col = inside ? (uint32_t)wrap_once(x + p.ox, p.nx) + (uint32_t)p.nx * (uint32_t)wrap_once(y + p.oy, p.ny) : 0u;
    int lo, hi;
    clip_column(p, q, gx, gy, z_begin, z_end, lo, hi);
    if (!inside) { lo = z_end; hi = z_begin - 1; }
    const int own_lo = lo, own_hi = hi;
    lo = wave_min(lo);
    hi = wave_max(hi);
    if (!inside) return;

    const float cxy0 = q.r[0] * gx + q.r[1] * gy;
    const float cxy1 = q.r[3] * gx + q.r[4] * gy;
    const float cxy2 = q.r[6] * gx + q.r[7] * gy;
    const float fcols = (float)p.cols, frows = (float)p.rows;
    const uint32_t nz = (uint32_t)p.nz;
    for (int z0 = lo; z0 <= hi; z0 += kIntegrateBatch) {
        if (cbound != nullptr) {
            const int za = max(z0, own_lo), zb = min(z0 + kIntegrateBatch - 1, own_hi);
            const bool culled = za > zb || color_segment_culled(p, q, cbound, gx, gy, cxy0, cxy1, cxy2, za, zb);
            if (__all(culled)) continue;
        }
        float px[kIntegrateBatch], py[kIntegrateBatch], pz[kIntegrateBatch];
        int pix[kIntegrateBatch];
        bool inimg[kIntegrateBatch];
        // Phase 1 (no branches): projection and vertex gather.  Pixels
        // outside the image read pixel 0 and are masked afterwards.
#pragma unroll
        for (int k = 0; k < kIntegrateBatch; ++k) {
            const int z = z0 + k;
            const float gz = ((float)z + 0.5f) * p.csz - q.tz;
            const float camx = cxy0 + q.r[2] * gz;
            const float camy = cxy1 + q.r[5] * gz;
            const float camz = cxy2 + q.r[8] * gz;
            // In front of the camera: camz in [2^-126, 2^126), where rcp_rn is
            // the correctly rounded reciprocal.
            const float inv_z = rcp_rn(camz);
            const float fu = rintf(camx * p.fx * inv_z + p.cx);
            const float fv = rintf(camy * p.fy * inv_z + p.cy);
            const bool in = (z <= hi) & (camz >= kRcpMin) & (camz < kRcpMax) & (fu >= 0.0f) & (fu < fcols) &
                            (fv >= 0.0f) & (fv < frows);
            pix[k] = (int)(in ? fv : 0.0f) * p.cols + (int)(in ? fu : 0.0f);
            const kfx_xyz v = vmap[pix[k]];
            px[k] = v.x; py[k] = v.y; pz[k] = v.z;
            inimg[k] = in;
        }
        // Phase 2: the distance test, and the loads of what it selects.
        bool upd[kIntegrateBatch];
        kfx_color prev[kIntegrateBatch], cur[kIntegrateBatch];
        uint32_t addr[kIntegrateBatch];
#pragma unroll
        for (int k = 0; k < kIntegrateBatch; ++k) {
            const float vgz = ((float)(z0 + k) + 0.5f) * p.csz;
            const float dx = px[k] - vgx, dy = py[k] - vgy, dz = pz[k] - vgz;
            const float dist = sqrt_rn(dx * dx + dy * dy + dz * dz);
            upd[k] = inimg[k] & !isnan(px[k]) & (dist < p.trunc);
            uint32_t rz = (uint32_t)(z0 + k) + (uint32_t)p.oz;
            rz = min(rz, rz - nz);
            addr[k] = col + p.slice * rz;
            if (upd[k]) {
                prev[k] = volume[addr[k]];
                cur[k] = rgb[pix[k]];
            }
        }
        // Phase 3: running average per channel, write what changed.
#pragma unroll
        for (int k = 0; k < kIntegrateBatch; ++k) {
            if (!upd[k]) continue;
            const int w = prev[k].w;
            const float den = (float)(w + 1);
            kfx_color out;
            out.r = (uint8_t)min(255, max(0, (int)rintf(((float)(prev[k].r * w) + (float)cur[k].r) / den)));
            out.g = (uint8_t)min(255, max(0, (int)rintf(((float)(prev[k].g * w) + (float)cur[k].g) / den)));
            out.b = (uint8_t)min(255, max(0, (int)rintf(((float)(prev[k].b * w) + (float)cur[k].b) / den)));
            out.w = (uint8_t)min(max_weight, w + 1);
            if (out.r != prev[k].r || out.g != prev[k].g || out.b != prev[k].b || out.w != prev[k].w)
                volume[addr[k]] = out;
        }
    }
}

__global__ void cyclic_address_kernel(const int32_t* __restrict__ xyz, uint32_t* __restrict__ out,
                                      size_t n, kfx_round_buffer rb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = cyclic_address(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], rb);
}

// Checks shared by every kernel that sweeps the volume with wrap_once().
int check_round_buffer(const kfx_round_buffer* rb, const char* who) {
    KFX_REQUIRE(rb != nullptr, "%s: null round_buffer", who);
    const kfx_point_3i n = rb->voxels_size;
    KFX_REQUIRE(n.x > 0 && n.y > 0 && n.z > 0, "%s: voxels_size must be positive", who);
    KFX_REQUIRE((int64_t)n.x * n.y * n.z == (int64_t)rb->buffer_length,
                "%s: buffer_length %d != Nx*Ny*Nz", who, rb->buffer_length);
    KFX_REQUIRE(rb->start_index == 0,
                "%s: start_index must be 0 (the reference always sets 0; its '>' fold would "
                "address one voxel past the end otherwise)", who);
    KFX_REQUIRE(rb->origin_GRID.x >= 0 && rb->origin_GRID.x < n.x && rb->origin_GRID.y >= 0 &&
                    rb->origin_GRID.y < n.y && rb->origin_GRID.z >= 0 && rb->origin_GRID.z < n.z,
                "%s: origin_GRID (%d,%d,%d) must lie in [0, N) per axis", who, rb->origin_GRID.x,
                rb->origin_GRID.y, rb->origin_GRID.z);
    return KFX_OK;
}

}  // namespace kfx

using namespace kfx;

extern "C" {

uint32_t kfx_cyclic_address(int x, int y, int z, const kfx_round_buffer* rb) {
    return cyclic_address(x, y, z, *rb);
}

int kfx_cyclic_address_batch(const int32_t* xyz, uint32_t* out, size_t n, const kfx_round_buffer* rb) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(rb != nullptr, "kfx_cyclic_address_batch: null round_buffer");
    KFX_REQUIRE(rb->voxels_size.x > 0 && rb->voxels_size.y > 0 && rb->voxels_size.z > 0 && rb->buffer_length > 0,
                "kfx_cyclic_address_batch: sizes must be positive");
    if (n == 0) return KFX_OK;
    KFX_REQUIRE(xyz && out, "kfx_cyclic_address_batch: null pointer");
    const unsigned blocks = (unsigned)((n + 255) / 256);
    cyclic_address_kernel<<<blocks, 256, 0, stream()>>>(xyz, out, n, *rb);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_init_tsdf(kfx_tsdf* volume, const kfx_round_buffer* rb) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(volume && rb, "kfx_init_tsdf: null pointer");
    KFX_REQUIRE(rb->buffer_length > 0, "kfx_init_tsdf: empty volume");
    return kfx_fill_u8(volume, 0, (size_t)rb->buffer_length * sizeof(kfx_tsdf));
}

}  // extern "C"

namespace kfx {

static size_t bound_layout(int cols, int rows, int off[kBoundLevels + 1]) {
    size_t total = 0;
    off[0] = 0;
    for (int l = 1; l <= kBoundLevels; ++l) {
        off[l] = (int)total;
        total += (size_t)((cols + (1 << l) - 1) >> l) * (size_t)((rows + (1 << l) - 1) >> l);
    }
    return total;
}

size_t bound_floats(int cols, int rows) {
    int off[kBoundLevels + 1];
    return bound_layout(cols, rows, off);
}

// Frame terms buffer of a workspace: the PoseTerms of the frame, then its
// skip flag.
constexpr size_t kTermsSkipOffset = (sizeof(PoseTerms) + 15) / 16 * 16;
constexpr size_t kTermsBytes = kTermsSkipOffset + 16;

size_t integrate_blocks(const kfx_point_3i& n) {
    return (size_t)divup((unsigned)n.x, kIntegrateBlockX) * divup((unsigned)n.y, kIntegrateBlockY) *
           divup((unsigned)n.z, (unsigned)integrate_zrun(n.z));
}

// The integration pre-pass (one launch): the scaled depth into depth_scaled
// and, as far as ws holds them, the depth-bound pyramid (p.bound points at
// it) and the workgroup flags of the sweep grid igrid (*flags_out), with a
// device pose also the frame's pose terms and skip flag.  Without pyramid or
// flags (the library workspace inside a stream capture before any
// reservation) the host-pose sweep runs unculled: same result, more work.
//
// bound_ready: depth_scaled and the pyramid were built by the maps launch
// (depth_bound_job); the launch then covers the flags only, one thread per
// sweep workgroup.
template <bool kDevicePose>
static int integrate_prepass(const Workspace& ws, bool profile, const uint16_t* depth, float* depth_scaled,
                             const kfx_intrinsics* intr, IntegrateParams& p, dim3 igrid, const float* dpose,
                             const float* dprev, float motion_thr, const int32_t* dskip, const uint8_t** flags_out,
                             bool bound_ready = false) {
    p.bound = nullptr;
    *flags_out = nullptr;
    const size_t need = bound_layout(intr->cols, intr->rows, p.bound_off);
    const bool have_bound = ws.bound != nullptr && need <= ws.bound_cap;
    const size_t nblk = (size_t)igrid.x * igrid.y * igrid.z;
    const bool have_flags = ws.flags != nullptr && nblk <= ws.flags_cap && (!kDevicePose || ws.terms != nullptr);
    KFX_REQUIRE(!kDevicePose || have_flags,
                "kfx_integrate_tsdf_posed: no workspace for a %zu-workgroup sweep (reserve it outside a capture)",
                nblk);
    if (!have_bound && !have_flags) return scale_depth_launch(depth, depth_scaled, intr);
    PoseTerms* terms = kDevicePose ? static_cast<PoseTerms*>(ws.terms) : nullptr;
    int32_t* skip = kDevicePose ? reinterpret_cast<int32_t*>(static_cast<char*>(ws.terms) + kTermsSkipOffset) : nullptr;
    ProfScope prof(profile ? "depth_bound" : nullptr);
    const bool pixels = !(bound_ready && have_bound && have_flags);
    // The flags alone: one wave per workgroup, so that each wave's serial
    // chain (pose terms, frustum test) has a SIMD to itself.
    constexpr unsigned kFlagsBlock = 64;
    const dim3 grid = pixels ? dim3(divup(intr->cols, kBoundTile), divup(intr->rows, kBoundTile))
                             : dim3((unsigned)divup(nblk, (size_t)kFlagsBlock));
    const unsigned block = pixels ? kBoundTile * kBoundTile : kFlagsBlock;
    // igrid is sweep_grid's (y blocks, x blocks, z runs).
    integrate_prepass_kernel<kDevicePose><<<grid, block, 0, stream()>>>(
        depth, depth_scaled, *intr, have_bound ? ws.bound : nullptr, p, dpose, dprev, motion_thr, dskip, terms, skip,
        have_flags ? ws.flags : nullptr, (int)igrid.y, (int)igrid.x, (int)igrid.z, pixels ? 1 : 0);
    KFX_CHECK_LAUNCH();
    if (have_bound) p.bound = ws.bound;
    if (have_flags) *flags_out = ws.flags;
    return KFX_OK;
}

bool depth_bound_job(const Workspace& ws, const uint16_t* depth, float* depth_scaled, const kfx_intrinsics* intr,
                     DepthBoundJob* job) {
    const size_t need = bound_layout(intr->cols, intr->rows, job->bound_off);
    if (ws.bound == nullptr || need > ws.bound_cap || ws.flags == nullptr || ws.terms == nullptr) return false;
    job->depth = depth;
    job->scaled = depth_scaled;
    job->bound = ws.bound;
    job->intr = *intr;
    job->tiles_x = (int)divup(intr->cols, kBoundTile);
    job->tiles = job->tiles_x * (int)divup(intr->rows, kBoundTile);
    return true;
}

// Camera, volume and frustum terms of a column sweep (integration of depth or
// colour): everything but the depth-bound pyramid.
static int sweep_setup(const char* who, const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                       const kfx_point_3f* translation, float truncation_distance, const kfx_point_3f* cell_size,
                       const kfx_round_buffer* rb, IntegrateParams& p) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(intr && rotation_inverse && translation && cell_size, "%s: null pointer", who);
    int rc = check_round_buffer(rb, who);
    if (rc) return rc;
    KFX_REQUIRE(intr->cols > 0 && intr->rows > 0, "%s: empty image", who);
    KFX_REQUIRE(truncation_distance > 0.0f, "%s: truncation distance must be > 0", who);
    for (int r = 0; r < 3; ++r) {
        p.r[3 * r + 0] = rotation_inverse->data[r].x;
        p.r[3 * r + 1] = rotation_inverse->data[r].y;
        p.r[3 * r + 2] = rotation_inverse->data[r].z;
    }
    p.tx = translation->x; p.ty = translation->y; p.tz = translation->z;
    p.csx = cell_size->x; p.csy = cell_size->y; p.csz = cell_size->z;
    p.fx = intr->fx; p.fy = intr->fy; p.cx = intr->cx; p.cy = intr->cy;
    p.cols = intr->cols; p.rows = intr->rows;
    p.trunc = truncation_distance;
    p.inv_trunc = 1.0f / truncation_distance;
    p.nx = rb->voxels_size.x; p.ny = rb->voxels_size.y; p.nz = rb->voxels_size.z;
    p.ox = rb->origin_GRID.x; p.oy = rb->origin_GRID.y; p.oz = rb->origin_GRID.z;
    p.slice = (uint32_t)p.nx * (uint32_t)p.ny;
    p.zrun = integrate_zrun(p.nz);
    // Frustum planes: camera coordinates c = R^-1 g; a voxel can project
    // inside the image only if c.z > 0 and -0.5 - m <= f c.xy / c.z + c0 <
    // size - 0.5 + m.  In volume coordinates each plane normal is the
    // corresponding combination of rows of R^-1.
    {
        double ri[3][3];
        for (int r = 0; r < 3; ++r) {
            ri[r][0] = rotation_inverse->data[r].x;
            ri[r][1] = rotation_inverse->data[r].y;
            ri[r][2] = rotation_inverse->data[r].z;
        }
        const double m = kClipMarginPx;
        const double cl = intr->cx + 0.5 + m, cr = intr->cols - 0.5 + m - intr->cx;
        const double ct = intr->cy + 0.5 + m, cb = intr->rows - 0.5 + m - intr->cy;
        for (int k = 0; k < 3; ++k) {
            p.plane[0][k] = (float)ri[2][k];
            p.plane[1][k] = (float)(intr->fx * ri[0][k] + cl * ri[2][k]);
            p.plane[2][k] = (float)(-intr->fx * ri[0][k] + cr * ri[2][k]);
            p.plane[3][k] = (float)(intr->fy * ri[1][k] + ct * ri[2][k]);
            p.plane[4][k] = (float)(-intr->fy * ri[1][k] + cb * ri[2][k]);
        }
    }
    p.bound = nullptr;
    p.bricks = nullptr;
    p.nbx = p.nx >> 4;
    p.nby = p.ny >> 4;
    return KFX_OK;
}

int scale_depth_launch(const uint16_t* depth, float* depth_scaled, const kfx_intrinsics* intr) {
    ProfScope prof("scale_depth");
    dim3 sblock(64, 4);
    dim3 sgrid(divup(intr->cols, 64), divup(intr->rows, 4));
    scale_depth_kernel<<<sgrid, sblock, 0, stream()>>>(depth, depth_scaled, *intr);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

// Sweep terms; with scale, also the scaled depth (the integrations instead
// scale it in the integration pre-pass).
static int integrate_setup(const char* who, const uint16_t* depth, float* depth_scaled,
                           const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                           const kfx_point_3f* translation, float truncation_distance,
                           const kfx_point_3f* cell_size, const kfx_round_buffer* rb, IntegrateParams& p,
                           bool scale = true) {
    KFX_REQUIRE(depth && depth_scaled, "%s: null pointer", who);
    KFX_TRY(sweep_setup(who, intr, rotation_inverse, translation, truncation_distance, cell_size, rb, p));
    return scale ? scale_depth_launch(depth, depth_scaled, intr) : KFX_OK;
}

static size_t color_bound_layout(int cols, int rows, int off[kBoundLevels + 1]) {
    size_t total = (size_t)cols * (size_t)rows;
    off[0] = 0;
    for (int l = 1; l <= kBoundLevels; ++l) {
        off[l] = (int)total;
        total += (size_t)((cols + (1 << l) - 1) >> l) * (size_t)((rows + (1 << l) - 1) >> l);
    }
    return total;
}

size_t cbound_texels(int cols, int rows) {
    int off[kBoundLevels + 1];
    return color_bound_layout(cols, rows, off);
}

size_t cons_blocks(int cols, int rows) { return divup((unsigned)(cols * rows), 256u); }

// Builds the colour-bound pyramid of vertex_map into ws.cbound and returns it
// (null: the workspace is too small, run unculled).
static int colors_bound(const Workspace& ws, bool profile, const kfx_xyz* vertex_map, IntegrateParams& p,
                        const float* dpose, const int32_t* dskip, const float2** out) {
    *out = nullptr;
    const size_t need = color_bound_layout(p.cols, p.rows, p.bound_off);
    if (ws.cbound == nullptr || need > ws.cbound_cap) return KFX_OK;
    ProfScope prof(profile ? "color_bound" : nullptr);
    dim3 grid(divup(p.cols, kBoundTile), divup(p.rows, kBoundTile));
    color_bound_kernel<<<grid, kBoundTile * kBoundTile, 0, stream()>>>(vertex_map, ws.cbound, p, dpose, dskip);
    KFX_CHECK_LAUNCH();
    *out = ws.cbound;
    return KFX_OK;
}

int workspace_reserve(Workspace& ws, int cols, int rows, bool colors) {
    const size_t nb = bound_floats(cols, rows), nc = cbound_texels(cols, rows), ng = cons_blocks(cols, rows);
    if (ws.bound_cap < nb) {
        if (ws.bound) KFX_CHECK_HIP(hipFree(ws.bound));
        ws.bound = nullptr;
        ws.bound_cap = 0;
        KFX_CHECK_HIP(hipMalloc((void**)&ws.bound, nb * sizeof(float)));
        ws.bound_cap = nb;
    }
    if (colors && ws.cbound_cap < nc) {
        if (ws.cbound) KFX_CHECK_HIP(hipFree(ws.cbound));
        ws.cbound = nullptr;
        ws.cbound_cap = 0;
        KFX_CHECK_HIP(hipMalloc((void**)&ws.cbound, nc * sizeof(float2)));
        ws.cbound_cap = nc;
    }
    if (ws.cons_cap < ng) {
        if (ws.cons) KFX_CHECK_HIP(hipFree(ws.cons));
        ws.cons = nullptr;
        ws.cons_cap = 0;
        KFX_CHECK_HIP(hipMalloc((void**)&ws.cons, ng * 4 * sizeof(double)));
        ws.cons_cap = ng;
    }
    return KFX_OK;
}

int workspace_reserve_volume(Workspace& ws, const kfx_point_3i& voxels) {
    const size_t nb = integrate_blocks(voxels);
    if (ws.terms == nullptr) KFX_CHECK_HIP(hipMalloc(&ws.terms, kTermsBytes));
    if (ws.flags_cap < nb) {
        if (ws.flags) KFX_CHECK_HIP(hipFree(ws.flags));
        ws.flags = nullptr;
        ws.flags_cap = 0;
        KFX_CHECK_HIP(hipMalloc((void**)&ws.flags, nb));
        ws.flags_cap = nb;
    }
    return KFX_OK;
}

void workspace_free(Workspace& ws) {
    if (ws.bound) (void)hipFree(ws.bound);
    if (ws.cbound) (void)hipFree(ws.cbound);
    if (ws.cons) (void)hipFree(ws.cons);
    if (ws.terms) (void)hipFree(ws.terms);
    if (ws.flags) (void)hipFree(ws.flags);
    ws = Workspace();
}

namespace {
Workspace g_lib_ws;
std::vector<void*> g_outgrown;  // library buffers replaced by larger ones, kept until exit
}  // namespace

Workspace& library_workspace() { return g_lib_ws; }

int library_workspace_reserve(int cols, int rows, bool colors, bool gate, const kfx_point_3i* voxels) {
    const bool need_b = g_lib_ws.bound_cap < bound_floats(cols, rows);
    const bool need_c = colors && g_lib_ws.cbound_cap < cbound_texels(cols, rows);
    const bool need_g = gate && g_lib_ws.cons_cap < cons_blocks(cols, rows);
    const size_t nflags = voxels ? integrate_blocks(*voxels) : 0;
    const bool need_f = voxels != nullptr && (g_lib_ws.flags_cap < nflags || g_lib_ws.terms == nullptr);
    if (!need_b && !need_c && !need_g && !need_f) return KFX_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    KFX_CHECK_HIP(hipStreamIsCapturing(stream(), &cs));
    if (cs != hipStreamCaptureStatusNone)
        return set_error(KFX_ERR_INVALID_ARG, "no library workspace reserved for %dx%d images inside a stream capture",
                         cols, rows);
    Workspace grown;
    if (need_b) KFX_CHECK_HIP(hipMalloc((void**)&grown.bound, bound_floats(cols, rows) * sizeof(float)));
    if (need_c) {
        hipError_t e = hipMalloc((void**)&grown.cbound, cbound_texels(cols, rows) * sizeof(float2));
        if (e != hipSuccess) { workspace_free(grown); KFX_CHECK_HIP(e); }
    }
    if (need_g) {
        hipError_t e = hipMalloc((void**)&grown.cons, cons_blocks(cols, rows) * 4 * sizeof(double));
        if (e != hipSuccess) { workspace_free(grown); KFX_CHECK_HIP(e); }
    }
    if (need_f && g_lib_ws.flags_cap < nflags) {
        hipError_t e = hipMalloc((void**)&grown.flags, nflags);
        if (e != hipSuccess) { workspace_free(grown); KFX_CHECK_HIP(e); }
    }
    if (need_f && g_lib_ws.terms == nullptr) {
        hipError_t e = hipMalloc(&grown.terms, kTermsBytes);
        if (e != hipSuccess) { workspace_free(grown); KFX_CHECK_HIP(e); }
    }
    if (need_b) {
        if (g_lib_ws.bound) g_outgrown.push_back(g_lib_ws.bound);
        g_lib_ws.bound = grown.bound;
        g_lib_ws.bound_cap = bound_floats(cols, rows);
    }
    if (need_c) {
        if (g_lib_ws.cbound) g_outgrown.push_back(g_lib_ws.cbound);
        g_lib_ws.cbound = grown.cbound;
        g_lib_ws.cbound_cap = cbound_texels(cols, rows);
    }
    if (need_g) {
        if (g_lib_ws.cons) g_outgrown.push_back(g_lib_ws.cons);
        g_lib_ws.cons = grown.cons;
        g_lib_ws.cons_cap = cons_blocks(cols, rows);
    }
    if (grown.flags) {
        if (g_lib_ws.flags) g_outgrown.push_back(g_lib_ws.flags);
        g_lib_ws.flags = grown.flags;
        g_lib_ws.flags_cap = nflags;
    }
    if (grown.terms) g_lib_ws.terms = grown.terms;
    return KFX_OK;
}

// Library workspace for a standalone call: reserved outside a capture;
// inside one, whatever was reserved before (too small: unculled).
static int library_ws_for(int cols, int rows, bool colors, const Workspace** out,
                          const kfx_point_3i* voxels = nullptr) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    KFX_CHECK_HIP(hipStreamIsCapturing(stream(), &cs));
    if (cs == hipStreamCaptureStatusNone) KFX_TRY(library_workspace_reserve(cols, rows, colors, false, voxels));
    *out = &g_lib_ws;
    return KFX_OK;
}

static int colors_check(const char* who, kfx_color* color_volume, const kfx_color* rgb_frame,
                        const kfx_xyz* vertex_map, int max_weight) {
    KFX_REQUIRE(color_volume && rgb_frame && vertex_map, "%s: null pointer", who);
    KFX_REQUIRE(max_weight >= 1 && max_weight <= 255, "%s: max_weight %d outside [1, 255] (8-bit weight)", who,
                max_weight);
    return KFX_OK;
}


int integrate_tsdf_ws(const Workspace& ws, bool profile, const uint16_t* depth, float* depth_scaled,
                      kfx_tsdf* volume, const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                      const kfx_point_3f* translation, float truncation_distance, const kfx_point_3f* cell_size,
                      const kfx_round_buffer* rb, uint8_t* bricks) {
    KFX_REQUIRE(volume != nullptr, "kfx_integrate_tsdf: null volume");
    IntegrateParams p;
    int rc = integrate_setup("kfx_integrate_tsdf", depth, depth_scaled, intr, rotation_inverse, translation,
                             truncation_distance, cell_size, rb, p, false);
    if (rc) return rc;
    p.bricks = bricks;
    dim3 block(kIntegrateBlockX, kIntegrateBlockY);
    const dim3 grid = sweep_grid(p.nx, p.ny, p.nz, p.zrun);
    const uint8_t* flags = nullptr;
    KFX_TRY(integrate_prepass<false>(ws, profile, depth, depth_scaled, intr, p, grid, nullptr, nullptr, 0.0f, nullptr,
                                     &flags));
    ProfScope prof(profile ? "integrate" : nullptr);
    integrate_kernel<false><<<grid, block, 0, stream()>>>(depth_scaled, volume, p, nullptr, nullptr, flags);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int integrate_tsdf_posed_ws(const Workspace& ws, const uint16_t* depth, float* depth_scaled, kfx_tsdf* volume,
                            const kfx_intrinsics* intr, const float* dev_pose, const float* dev_prev_pose_inv,
                            float motion_threshold, float truncation_distance, const kfx_point_3f* cell_size,
                            const kfx_round_buffer* rb, const int32_t* dev_skip, uint8_t* bricks, bool bound_ready) {
    KFX_REQUIRE(volume != nullptr && dev_pose != nullptr, "kfx_integrate_tsdf_posed: null pointer");
    const kfx_matrix_3f ident = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    const kfx_point_3f zero = {0, 0, 0};
    IntegrateParams p;
    int rc = integrate_setup("kfx_integrate_tsdf_posed", depth, depth_scaled, intr, &ident, &zero,
                             truncation_distance, cell_size, rb, p, false);
    if (rc) return rc;
    p.bricks = bricks;
    dim3 block(kIntegrateBlockX, kIntegrateBlockY);
    const dim3 grid = sweep_grid(p.nx, p.ny, p.nz, p.zrun);
    const uint8_t* flags = nullptr;
    KFX_TRY(integrate_prepass<true>(ws, true, depth, depth_scaled, intr, p, grid, dev_pose, dev_prev_pose_inv,
                                    motion_threshold, dev_skip, &flags, bound_ready));
    ProfScope prof("integrate");
    const PoseTerms* terms = static_cast<const PoseTerms*>(ws.terms);
    const int32_t* skip = reinterpret_cast<const int32_t*>(static_cast<const char*>(ws.terms) + kTermsSkipOffset);
    integrate_kernel<true><<<grid, block, 0, stream()>>>(depth_scaled, volume, p, terms, skip, flags);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int integrate_colors_ws(const Workspace& ws, kfx_color* color_volume, const kfx_color* rgb_frame,
                        const kfx_xyz* vertex_map, const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                        const kfx_point_3f* translation, float truncation_distance, int max_weight,
                        const kfx_point_3f* cell_size, const kfx_round_buffer* rb) {
    KFX_TRY(colors_check("kfx_integrate_colors", color_volume, rgb_frame, vertex_map, max_weight));
    IntegrateParams p;
    KFX_TRY(sweep_setup("kfx_integrate_colors", intr, rotation_inverse, translation, truncation_distance, cell_size,
                        rb, p));
    const float2* cb = nullptr;
    KFX_TRY(colors_bound(ws, true, vertex_map, p, nullptr, nullptr, &cb));
    ProfScope prof("integrate_colors");
    dim3 block(kIntegrateBlockX, kIntegrateBlockY);
    const dim3 grid = sweep_grid(p.nx, p.ny, p.nz, p.zrun);
    integrate_colors_kernel<false><<<grid, block, 0, stream()>>>(rgb_frame, vertex_map, color_volume, p, cb,
                                                                 max_weight, nullptr, nullptr);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int integrate_colors_posed_ws(const Workspace& ws, kfx_color* color_volume, const kfx_color* rgb_frame,
                              const kfx_xyz* vertex_map, const kfx_intrinsics* intr, const float* dev_pose,
                              float truncation_distance, int max_weight, const kfx_point_3f* cell_size,
                              const kfx_round_buffer* rb, const int32_t* dev_skip) {
    KFX_TRY(colors_check("kfx_integrate_colors_posed", color_volume, rgb_frame, vertex_map, max_weight));
    KFX_REQUIRE(dev_pose != nullptr, "kfx_integrate_colors_posed: null pose");
    const kfx_matrix_3f ident = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    const kfx_point_3f zero = {0, 0, 0};
    IntegrateParams p;
    KFX_TRY(sweep_setup("kfx_integrate_colors_posed", intr, &ident, &zero, truncation_distance, cell_size, rb, p));
    const float2* cb = nullptr;
    KFX_TRY(colors_bound(ws, true, vertex_map, p, dev_pose, dev_skip, &cb));
    ProfScope prof("integrate_colors");
    dim3 block(kIntegrateBlockX, kIntegrateBlockY);
    const dim3 grid = sweep_grid(p.nx, p.ny, p.nz, p.zrun);
    integrate_colors_kernel<true><<<grid, block, 0, stream()>>>(rgb_frame, vertex_map, color_volume, p, cb,
                                                                max_weight, dev_pose, dev_skip);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

// Depth-bound pyramid of kfx_integrate_tsdf_noalloc: a fixed array in device
// memory, one copy per device that loads the library, held until the
// process ends.  Host code never passes the array's name to a kernel: it is
// a host-side stand-in there.  noalloc_setup resolves each device's address
// once (kfx_init, with the device current) and the entry point uses that.
constexpr size_t kNoAllocBoundFloats = size_t(1) << 20;  // 4 MiB: levels 1..5 of images up to 2048 x 1536
static __device__ __align__(256) float scratch[kNoAllocBoundFloats];
constexpr int kNoAllocMaxDevices = 64;
static float* g_scratch_ptr[kNoAllocMaxDevices] = {};

int noalloc_setup(int device) {
    KFX_REQUIRE(device >= 0 && device < kNoAllocMaxDevices, "kfx_init: device %d beyond the %d the library supports",
                device, kNoAllocMaxDevices);
    if (g_scratch_ptr[device] != nullptr) return KFX_OK;
    float* ptr = nullptr;
    KFX_CHECK_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    KFX_REQUIRE(ptr != nullptr, "kfx_init: no device address for the integration scratch array");
    g_scratch_ptr[device] = ptr;
    return KFX_OK;
}

}  // namespace kfx

extern "C" {

int kfx_integrate_tsdf(const uint16_t* depth, float* depth_scaled, kfx_tsdf* volume,
                       const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                       const kfx_point_3f* translation, float truncation_distance,
                       const kfx_point_3f* cell_size, const kfx_round_buffer* rb) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(intr != nullptr && rb != nullptr, "kfx_integrate_tsdf: null pointer");
    const Workspace* ws = nullptr;
    KFX_TRY(library_ws_for(intr->cols, intr->rows, false, &ws, &rb->voxels_size));
    return integrate_tsdf_ws(*ws, true, depth, depth_scaled, volume, intr, rotation_inverse, translation,
                             truncation_distance, cell_size, rb);
}

int kfx_integrate_tsdf_posed(const uint16_t* depth, float* depth_scaled, kfx_tsdf* volume,
                             const kfx_intrinsics* intr, const float* dev_pose, const float* dev_prev_pose_inv,
                             float motion_threshold, float truncation_distance, const kfx_point_3f* cell_size,
                             const kfx_round_buffer* rb, const int32_t* dev_skip) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(intr != nullptr && rb != nullptr, "kfx_integrate_tsdf_posed: null pointer");
    const Workspace* ws = nullptr;
    KFX_TRY(library_ws_for(intr->cols, intr->rows, false, &ws, &rb->voxels_size));
    return integrate_tsdf_posed_ws(*ws, depth, depth_scaled, volume, intr, dev_pose, dev_prev_pose_inv,
                                   motion_threshold, truncation_distance, cell_size, rb, dev_skip);
}

int kfx_integrate_tsdf_noalloc(const uint16_t* depth, float* depth_scaled, kfx_tsdf* volume,
                               const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                               const kfx_point_3f* translation, float truncation_distance,
                               const kfx_point_3f* cell_size, const kfx_round_buffer* rb) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(intr != nullptr, "kfx_integrate_tsdf_noalloc: null pointer");
    KFX_REQUIRE(intr->cols > 0 && intr->rows > 0, "kfx_integrate_tsdf_noalloc: empty image");
    const size_t need = bound_floats(intr->cols, intr->rows);
    KFX_REQUIRE(need <= kNoAllocBoundFloats,
                "kfx_integrate_tsdf_noalloc: a %dx%d image needs %zu floats of scratch, more than the %zu reserved",
                intr->cols, intr->rows, need, kNoAllocBoundFloats);
    const int dev = current_device();
    KFX_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices && g_scratch_ptr[dev] != nullptr,
                "kfx_integrate_tsdf_noalloc: scratch not set up for device %d", dev);
    Workspace ws;
    ws.bound = g_scratch_ptr[dev];
    ws.bound_cap = kNoAllocBoundFloats;
    return integrate_tsdf_ws(ws, false, depth, depth_scaled, volume, intr, rotation_inverse, translation,
                             truncation_distance, cell_size, rb);
}

int kfx_integrate_count_updates(const uint16_t* depth, float* depth_scaled, const kfx_intrinsics* intr,
                                const kfx_matrix_3f* rotation_inverse, const kfx_point_3f* translation,
                                float truncation_distance, const kfx_point_3f* cell_size,
                                const kfx_round_buffer* rb, uint64_t* count) {
    KFX_REQUIRE(count != nullptr, "kfx_integrate_count_updates: null output");
    IntegrateParams p;
    int rc = integrate_setup("kfx_integrate_count_updates", depth, depth_scaled, intr, rotation_inverse,
                             translation, truncation_distance, cell_size, rb, p);
    if (rc) return rc;
    unsigned long long* dcount = nullptr;
    KFX_CHECK_HIP(hipMallocAsync((void**)&dcount, sizeof(unsigned long long), stream()));
    hipError_t e = hipMemsetAsync(dcount, 0, sizeof(unsigned long long), stream());
    if (e == hipSuccess) {
        dim3 block(kIntegrateBlockX, kIntegrateBlockY);
        const dim3 grid = sweep_grid(p.nx, p.ny, p.nz, p.zrun);
        integrate_count_kernel<<<grid, block, 0, stream()>>>(depth_scaled, p, dcount);
        e = hipGetLastError();
    }
    unsigned long long host = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&host, dcount, sizeof(host), hipMemcpyDeviceToHost, stream());
    hipError_t e2 = hipFreeAsync(dcount, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    KFX_CHECK_HIP(e);
    KFX_CHECK_HIP(e2);
    *count = host;
    return KFX_OK;
}

int kfx_integrate_colors(kfx_color* color_volume, const kfx_color* rgb_frame, const kfx_xyz* vertex_map,
                         const kfx_intrinsics* intr, const kfx_matrix_3f* rotation_inverse,
                         const kfx_point_3f* translation, float truncation_distance, int max_weight,
                         const kfx_point_3f* cell_size, const kfx_round_buffer* rb) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(intr != nullptr, "kfx_integrate_colors: null pointer");
    const Workspace* ws = nullptr;
    KFX_TRY(library_ws_for(intr->cols, intr->rows, true, &ws));
    return integrate_colors_ws(*ws, color_volume, rgb_frame, vertex_map, intr, rotation_inverse, translation,
                               truncation_distance, max_weight, cell_size, rb);
}

int kfx_integrate_colors_posed(kfx_color* color_volume, const kfx_color* rgb_frame, const kfx_xyz* vertex_map,
                               const kfx_intrinsics* intr, const float* dev_pose, float truncation_distance,
                               int max_weight, const kfx_point_3f* cell_size, const kfx_round_buffer* rb,
                               const int32_t* dev_skip) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(intr != nullptr, "kfx_integrate_colors_posed: null pointer");
    const Workspace* ws = nullptr;
    KFX_TRY(library_ws_for(intr->cols, intr->rows, true, &ws));
    return integrate_colors_posed_ws(*ws, color_volume, rgb_frame, vertex_map, intr, dev_pose, truncation_distance,
                                     max_weight, cell_size, rb, dev_skip);
}

int kfx_init_colors(kfx_color* volume, const kfx_round_buffer* rb) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(volume && rb, "kfx_init_colors: null pointer");
    KFX_REQUIRE(rb->buffer_length > 0, "kfx_init_colors: empty volume");
    return kfx_fill_u8(volume, 0, (size_t)rb->buffer_length * sizeof(kfx_color));
}

}  // extern "C"
