// kfx_rgbd.hip — RGBD photometric odometry (Estimate_Isometry_By_RGBD,
// common/src/opencl_wrappers.cpp:1328-1466, with the kernels
// scan_for_interest_point.cl, scan_list_of_valid_correspondence.cl and
// compute_least_squares_equation.cl).  The reference's path does not work
// (uninitialised local memory, an inverted depth test, a least-squares
// matrix that only accumulates its first row, swapped memsets, float data
// wrapped as double); the algorithm it follows is the dense RGBD odometry of
// the OpenCV rgbd module (Steinbruecker, Sturm, Cremers 2011), whose weights
// (1 / (sigma + |diff|)), Sobel scale (1/8), Jacobian rows and twist
// exponential the reference's code reproduces
// (compute_least_squares_equation.cl:37-71, opencl_wrappers.cpp:1282-1326).
// This file implements that algorithm as the reference intends it:
//
//   frame preparation: intensity 0.299 r + 0.587 g + 0.114 b and metric depth
//     (d_mm / 1000, 0 = invalid), pyramids (intensity: 2x2 mean; depth: the
//     top-left sample of each 2x2 block), 3x3 Sobel derivatives (borders
//     reflected without repeating the edge pixel).
//   per iteration of a level (coarsest first), on the device:
//     1. correspondences: every textured current-frame pixel (Sobel magnitude
//        >= min_gradient, valid depth) is back-projected, moved into the
//        previous frame by the inverse of the estimate and projected; it
//        pairs with the previous-frame pixel it lands on when that pixel has
//        depth within max_depth_diff of the moved point.  Several pixels
//        landing on one previous-frame pixel: the nearest moved point wins
//        (ties: the later pixel in row-major order), a 64-bit atomic minimum
//        of (depth bits, reversed index).
//     2. sigma = sqrt(mean diff^2), diff = I_prev(u0) - I_curr(u1).  Fewer
//        than 120 correspondences end the level.
//     3. normal equations A += C C^T, b += C w diff with
//        w = 1 / (sigma + |diff|) and C the Jacobian row of the photometric
//        residual (derivatives of the current intensity scaled by w / 8), in
//        double precision, per workgroup then in a fixed order.
//     4. one workgroup: |det A| < 1e-6 (or not finite) ends the level;
//        otherwise the Cholesky solution ksi, the twist exponential
//        exp([ksi0..2]x, ksi3..5) and estimate <- exp * estimate.
//   The estimate maps previous-camera points to current-camera points.
#include <utility>

#include "kfx_common.h"

namespace kfx {

constexpr int kRgbdMaxLevels = 4;
constexpr int kRgbdThreads = 256;
constexpr int kRgbdMinCorresp = 120;  // 20 per unknown
constexpr double kRgbdSobelScale = 1.0 / 8.0;
constexpr double kRgbdDetThreshold = 1e-6;

struct RgbdState {
    double R[9], t[3];     // estimate: previous camera -> current camera
    float Rif[9], tif[3];  // its inverse in single precision (correspondences)
    double sigma;
    double A[36], b[6];    // last normal equations
    int count;             // correspondences of the last iteration
    int stop;              // the current level has ended
    int ok;                // at least one solve succeeded
    int solves;
};

struct RgbdLevel {
    const float *i_prev, *d_prev, *i_curr, *d_curr, *dx_curr, *dy_curr;
    int cols, rows;
    float fx, fy, cx, cy;
};

__device__ __forceinline__ int reflect101(int i, int n) {
    if (i < 0) return n > 1 ? -i : 0;
    if (i >= n) return n > 1 ? 2 * n - 2 - i : 0;
    return i;
}

// Level 0 of a frame: intensity and metric depth.
__global__ void __launch_bounds__(256) rgbd_frame_kernel(const kfx_color* __restrict__ rgb,
                                                         const uint16_t* __restrict__ depth, int n,
                                                         float* __restrict__ intensity, float* __restrict__ metres) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const kfx_color c = rgb[i];
    intensity[i] = (0.299f * (float)c.r + 0.587f * (float)c.g) + 0.114f * (float)c.b;
    metres[i] = (float)depth[i] * 0.001f;
}

// Level l from level l - 1 (cols x rows is level l's size).
__global__ void __launch_bounds__(256) rgbd_down_kernel(const float* __restrict__ i_in, const float* __restrict__ d_in,
                                                        int in_cols, float* __restrict__ i_out,
                                                        float* __restrict__ d_out, int cols, int rows) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols || y >= rows) return;
    const float* r0 = i_in + (2 * y) * in_cols + 2 * x;
    const float* r1 = r0 + in_cols;
    i_out[y * cols + x] = 0.25f * ((r0[0] + r0[1]) + (r1[0] + r1[1]));
    d_out[y * cols + x] = d_in[(2 * y) * in_cols + 2 * x];
}

__global__ void __launch_bounds__(256) rgbd_sobel_kernel(const float* __restrict__ in, int cols, int rows,
                                                         float* __restrict__ dx, float* __restrict__ dy) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cols || y >= rows) return;
    const int xm = reflect101(x - 1, cols), xp = reflect101(x + 1, cols);
    const int ym = reflect101(y - 1, rows), yp = reflect101(y + 1, rows);
    const float* rm = in + ym * cols;
    const float* r0 = in + y * cols;
    const float* rp = in + yp * cols;
    dx[y * cols + x] = ((rm[xp] + 2.0f * r0[xp]) + rp[xp]) - ((rm[xm] + 2.0f * r0[xm]) + rp[xm]);
    dy[y * cols + x] = ((rp[xm] + 2.0f * rp[x]) + rp[xp]) - ((rm[xm] + 2.0f * rm[x]) + rm[xp]);
}

// Start of a level: clear the correspondence map and the level's stop flag.
__global__ void __launch_bounds__(256) rgbd_clear_kernel(unsigned long long* __restrict__ corr, int n,
                                                         RgbdState* __restrict__ st, int new_level) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (new_level) {
        if (i == 0) st->stop = 0;
    } else if (st->stop) {
        return;
    }
    if (i < n) corr[i] = ~0ull;
}

// Step 1: correspondences (one thread per current-frame pixel).
__global__ void __launch_bounds__(256) rgbd_corr_kernel(const RgbdLevel L, const RgbdState* __restrict__ st,
                                                        float min_gradient_sq, float max_depth_diff,
                                                        unsigned long long* __restrict__ corr) {
    if (st->stop) return;
    const int u1 = blockIdx.x * blockDim.x + threadIdx.x, v1 = blockIdx.y;
    if (u1 >= L.cols || v1 >= L.rows) return;
    const int i1 = v1 * L.cols + u1;
    const float d1 = L.d_curr[i1];
    const float gx = L.dx_curr[i1], gy = L.dy_curr[i1];
    if (!(d1 > 0.0f) || gx * gx + gy * gy < min_gradient_sq) return;
    const float x1 = ((float)u1 - L.cx) / L.fx * d1, y1 = ((float)v1 - L.cy) / L.fy * d1;
    const float* R = st->Rif;
    const float px = (R[0] * x1 + R[1] * y1) + R[2] * d1 + st->tif[0];
    const float py = (R[3] * x1 + R[4] * y1) + R[5] * d1 + st->tif[1];
    const float pz = (R[6] * x1 + R[7] * y1) + R[8] * d1 + st->tif[2];
    if (!(pz > 0.0f)) return;
    const float inv_z = 1.0f / pz;
    const float fu = rintf(px * L.fx * inv_z + L.cx), fv = rintf(py * L.fy * inv_z + L.cy);
    if (!(fu >= 0.0f && fu < (float)L.cols && fv >= 0.0f && fv < (float)L.rows)) return;
    const int i0 = (int)fv * L.cols + (int)fu;
    const float d0 = L.d_prev[i0];
    if (!(d0 > 0.0f) || !(fabsf(pz - d0) <= max_depth_diff)) return;
    const unsigned long long key =
        ((unsigned long long)__float_as_uint(pz) << 32) | (unsigned long long)(0xFFFFFFFFu - (uint32_t)i1);
    atomicMin(&corr[i0], key);
}

// Fixed-order sum of N doubles per thread over the workgroup (an LDS tree);
// thread k < N stores entry k's total in out[k].
template <int N>
__device__ __forceinline__ void rgbd_block_sum(double (&v)[N], double (*sh)[kRgbdThreads], double* out) {
    for (int k = 0; k < N; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kRgbdThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < N; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < N) out[threadIdx.x] = sh[threadIdx.x][0];
}

__device__ __forceinline__ bool rgbd_pair(const RgbdLevel& L, const unsigned long long* corr, int i0, int& i1,
                                          double& diff) {
    const unsigned long long key = corr[i0];
    if (key == ~0ull) return false;
    i1 = (int)(0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFull));
    diff = (double)(L.i_prev[i0] - L.i_curr[i1]);
    return true;
}

// Step 2a: per-workgroup (count, sum diff^2); partials [2][nblk].
__global__ void __launch_bounds__(kRgbdThreads) rgbd_stats_kernel(const RgbdLevel L, const RgbdState* __restrict__ st,
                                                                 const unsigned long long* __restrict__ corr,
                                                                 double* __restrict__ part) {
    __shared__ double sh[2][kRgbdThreads];
    if (st->stop) return;
    const int i0 = blockIdx.x * kRgbdThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    int i1;
    double diff;
    if (i0 < L.cols * L.rows && rgbd_pair(L, corr, i0, i1, diff)) {
        v[0] = 1.0;
        v[1] = diff * diff;
    }
    double out[2];
    rgbd_block_sum<2>(v, sh, out);
    if (threadIdx.x < 2) part[threadIdx.x * gridDim.x + blockIdx.x] = out[threadIdx.x];
}

// Step 2b (one workgroup): sigma, or the end of the level.
__global__ void __launch_bounds__(kRgbdThreads) rgbd_sigma_kernel(RgbdState* __restrict__ st,
                                                                 const double* __restrict__ part, int nblk) {
    __shared__ double sh[2][kRgbdThreads];
    __shared__ double out[2];
    if (st->stop) return;
    double v[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += kRgbdThreads) {
        v[0] += part[b];
        v[1] += part[nblk + b];
    }
    rgbd_block_sum<2>(v, sh, out);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int count = (int)out[0];
        st->count = count;
        if (count < kRgbdMinCorresp) st->stop = 1;
        else st->sigma = sqrt(out[1] / (double)count);
    }
}

// Step 3: per-workgroup normal equations; partials [27][nblk] (21 upper
// entries of A row by row, then b).
__global__ void __launch_bounds__(kRgbdThreads) rgbd_sums_kernel(const RgbdLevel L, const RgbdState* __restrict__ st,
                                                                const unsigned long long* __restrict__ corr,
                                                                double* __restrict__ part) {
    __shared__ double sh[27][kRgbdThreads];
    if (st->stop) return;
    const int i0 = blockIdx.x * kRgbdThreads + threadIdx.x;
    double v[27];
    for (int k = 0; k < 27; ++k) v[k] = 0.0;
    int i1;
    double diff;
    if (i0 < L.cols * L.rows && rgbd_pair(L, corr, i0, i1, diff)) {
        double w = st->sigma + fabs(diff);
        w = w > 2.220446049250313e-16 ? 1.0 / w : 1.0;
        const double ws = w * kRgbdSobelScale;
        // The previous frame's point, moved into the current camera.
        const int u0 = i0 % L.cols, v0 = i0 / L.cols;
        const float d0 = L.d_prev[i0];
        const float x0 = ((float)u0 - L.cx) / L.fx * d0, y0 = ((float)v0 - L.cy) / L.fy * d0;
        const double* R = st->R;
        const float X = (float)(R[0] * x0 + R[1] * y0 + R[2] * d0 + st->t[0]);
        const float Y = (float)(R[3] * x0 + R[4] * y0 + R[5] * d0 + st->t[1]);
        const float Z = (float)(R[6] * x0 + R[7] * y0 + R[8] * d0 + st->t[2]);
        const double inv_z = 1.0 / (double)Z;
        const double c3 = ws * (double)L.dx_curr[i1] * (double)L.fx * inv_z;
        const double c4 = ws * (double)L.dy_curr[i1] * (double)L.fy * inv_z;
        const double c5 = -(c3 * (double)X + c4 * (double)Y) * inv_z;
        const double C[6] = {-(double)Z * c4 + (double)Y * c5, (double)Z * c3 - (double)X * c5,
                             -(double)Y * c3 + (double)X * c4, c3, c4, c5};
        int k = 0;
        for (int a = 0; a < 6; ++a)
            for (int c = a; c < 6; ++c) v[k++] = C[a] * C[c];
        for (int a = 0; a < 6; ++a) v[21 + a] = C[a] * w * diff;
    }
    double out[27];
    rgbd_block_sum<27>(v, sh, out);
    if (threadIdx.x < 27) part[threadIdx.x * gridDim.x + blockIdx.x] = out[threadIdx.x];
}

// R = exp([w]x), t = V v: the exponential of the twist (w, v) (Rodrigues;
// V = I + (1 - cos a)/a^2 [w]x + (a - sin a)/a^3 [w]x^2, series for small a).
__host__ __device__ inline void twist_exp(const double ksi[6], double R[9], double t[3]) {
    const double wx = ksi[0], wy = ksi[1], wz = ksi[2];
    const double a2 = wx * wx + wy * wy + wz * wz, a = sqrt(a2);
    double A, B, Cc;  // sin a / a, (1 - cos a) / a^2, (a - sin a) / a^3
    if (a < 1e-4) {
        A = 1.0 - a2 / 6.0 + a2 * a2 / 120.0;
        B = 0.5 - a2 / 24.0 + a2 * a2 / 720.0;
        Cc = 1.0 / 6.0 - a2 / 120.0 + a2 * a2 / 5040.0;
    } else {
        A = sin(a) / a;
        B = (1.0 - cos(a)) / a2;
        Cc = (a - sin(a)) / (a2 * a);
    }
    const double K[9] = {0.0, -wz, wy, wz, 0.0, -wx, -wy, wx, 0.0};
    double K2[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) K2[3 * r + c] = K[3 * r] * K[c] + K[3 * r + 1] * K[3 + c] + K[3 * r + 2] * K[6 + c];
    double V[9];
    for (int i = 0; i < 9; ++i) {
        const double I = (i % 4 == 0) ? 1.0 : 0.0;
        R[i] = I + A * K[i] + B * K2[i];
        V[i] = I + B * K[i] + Cc * K2[i];
    }
    for (int r = 0; r < 3; ++r) t[r] = V[3 * r] * ksi[3] + V[3 * r + 1] * ksi[4] + V[3 * r + 2] * ksi[5];
}

// 6x6 Cholesky solve of A x = b; false when |det A| < kRgbdDetThreshold or
// the factorisation fails (the reference's solvability test, inverted there,
// opencl_wrappers.cpp:1286-1293).
__host__ __device__ inline bool rgbd_solve6(const double A[36], const double b[6], double x[6]) {
    double L[36] = {0};
    double det = 1.0;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = A[i * 6 + j];
            for (int p = 0; p < j; ++p) s -= L[i * 6 + p] * L[j * 6 + p];
            if (i == j) {
                if (!(s > 0.0)) return false;
                L[i * 6 + i] = sqrt(s);
                det *= s;
            } else {
                L[i * 6 + j] = s / L[j * 6 + j];
            }
        }
    if (!(fabs(det) >= kRgbdDetThreshold) || !isfinite(det)) return false;
    double y[6];
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
    return true;
}

// Step 4 (one workgroup): the fixed-order sum of the partials, the solve and
// the update estimate <- exp(ksi) * estimate.
__global__ void __launch_bounds__(kRgbdThreads) rgbd_solve_kernel(RgbdState* __restrict__ st,
                                                                 const double* __restrict__ part, int nblk) {
    __shared__ double sh[27][kRgbdThreads];
    __shared__ double tot[27];
    if (st->stop) return;
    double v[27];
    for (int k = 0; k < 27; ++k) v[k] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kRgbdThreads)
        for (int k = 0; k < 27; ++k) v[k] += part[k * nblk + b];
    rgbd_block_sum<27>(v, sh, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    double A[36], b[6];
    int k = 0;
    for (int r = 0; r < 6; ++r)
        for (int c = r; c < 6; ++c, ++k) A[r * 6 + c] = A[c * 6 + r] = tot[k];
    for (int r = 0; r < 6; ++r) b[r] = tot[21 + r];
    for (int i = 0; i < 36; ++i) st->A[i] = A[i];
    for (int i = 0; i < 6; ++i) st->b[i] = b[i];
    double ksi[6];
    if (!rgbd_solve6(A, b, ksi)) {
        st->stop = 1;
        return;
    }
    double Rc[9], tc[3];
    twist_exp(ksi, Rc, tc);
    double R[9], t[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            R[3 * r + c] = Rc[3 * r] * st->R[c] + Rc[3 * r + 1] * st->R[3 + c] + Rc[3 * r + 2] * st->R[6 + c];
        t[r] = Rc[3 * r] * st->t[0] + Rc[3 * r + 1] * st->t[1] + Rc[3 * r + 2] * st->t[2] + tc[r];
    }
    for (int i = 0; i < 9; ++i) st->R[i] = R[i];
    for (int i = 0; i < 3; ++i) st->t[i] = t[i];
    // Inverse for the next correspondences: R^T, -R^T t.
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) st->Rif[3 * r + c] = (float)R[3 * c + r];
        st->tif[r] = (float)(-(R[r] * t[0] + R[3 + r] * t[1] + R[6 + r] * t[2]));
    }
    st->ok = 1;
    st->solves += 1;
}

__global__ void rgbd_init_kernel(RgbdState* st, kfx_pose init) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i < 9; ++i) st->R[i] = init.R[i];
    for (int i = 0; i < 3; ++i) st->t[i] = init.t[i];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) st->Rif[3 * r + c] = (float)init.R[3 * c + r];
        st->tif[r] = (float)(-(init.R[r] * init.t[0] + init.R[3 + r] * init.t[1] + init.R[6 + r] * init.t[2]));
    }
    st->sigma = 0.0;
    st->count = 0;
    st->stop = 0;
    st->ok = 0;
    st->solves = 0;
}

}  // namespace kfx

using namespace kfx;

struct kfx_rgbd {
    int cols, rows, levels;
    float* buf[2][kRgbdMaxLevels][4];  // [frame][level][intensity, depth, dx, dy]
    bool have[2];
    unsigned long long* corr;
    double* part;
    RgbdState* state;
    RgbdState* host_state;
};

static int rgbd_level_size(int n, int l) { return n >> l; }

extern "C" {

int kfx_rgbd_create(int cols, int rows, int levels, kfx_rgbd** out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(out != nullptr, "kfx_rgbd_create: null pointer");
    KFX_REQUIRE(levels >= 1 && levels <= kRgbdMaxLevels, "kfx_rgbd_create: %d levels (1..%d)", levels, kRgbdMaxLevels);
    KFX_REQUIRE(cols >= (8 << (levels - 1)) && rows >= (8 << (levels - 1)),
                "kfx_rgbd_create: %dx%d too small for %d levels", cols, rows, levels);
    kfx_rgbd* h = new kfx_rgbd();
    h->cols = cols;
    h->rows = rows;
    h->levels = levels;
    const size_t n0 = (size_t)cols * rows;
    const int nblk = (int)divup((unsigned)n0, (unsigned)kRgbdThreads);
    hipError_t e = hipSuccess;
    for (int f = 0; f < 2; ++f)
        for (int l = 0; l < levels; ++l)
            for (int k = 0; k < 4; ++k) {
                const size_t n = (size_t)rgbd_level_size(cols, l) * rgbd_level_size(rows, l);
                if (e == hipSuccess) e = hipMalloc((void**)&h->buf[f][l][k], sizeof(float) * n);
            }
    if (e == hipSuccess) e = hipMalloc((void**)&h->corr, sizeof(unsigned long long) * n0);
    if (e == hipSuccess) e = hipMalloc((void**)&h->part, sizeof(double) * 27 * (size_t)nblk);
    if (e == hipSuccess) e = hipMalloc((void**)&h->state, sizeof(RgbdState));
    if (e == hipSuccess) e = hipHostMalloc((void**)&h->host_state, sizeof(RgbdState), hipHostMallocDefault);
    if (e != hipSuccess) {
        kfx_rgbd_destroy(h);
        return set_error(KFX_ERR_HIP, "kfx_rgbd_create: %s", hipGetErrorString(e));
    }
    *out = h;
    return KFX_OK;
}

int kfx_rgbd_destroy(kfx_rgbd* h) {
    if (h == nullptr) return KFX_OK;
    (void)hipStreamSynchronize(stream());
    for (int f = 0; f < 2; ++f)
        for (int l = 0; l < kRgbdMaxLevels; ++l)
            for (int k = 0; k < 4; ++k) (void)hipFree(h->buf[f][l][k]);
    (void)hipFree(h->corr);
    (void)hipFree(h->part);
    (void)hipFree(h->state);
    if (h->host_state) (void)hipHostFree(h->host_state);
    delete h;
    return KFX_OK;
}

int kfx_rgbd_set_frame(kfx_rgbd* h, int slot, const kfx_color* rgb, const uint16_t* depth) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && rgb && depth, "kfx_rgbd_set_frame: null pointer");
    KFX_REQUIRE(slot == 0 || slot == 1, "kfx_rgbd_set_frame: slot %d (0 previous, 1 current)", slot);
    const int n = h->cols * h->rows;
    float* const* b0 = h->buf[slot][0];
    rgbd_frame_kernel<<<divup(n, 256), 256, 0, stream()>>>(rgb, depth, n, b0[0], b0[1]);
    for (int l = 1; l < h->levels; ++l) {
        const int c = rgbd_level_size(h->cols, l), r = rgbd_level_size(h->rows, l);
        float* const* bp = h->buf[slot][l - 1];
        float* const* bl = h->buf[slot][l];
        rgbd_down_kernel<<<dim3(divup(c, 256), r), 256, 0, stream()>>>(bp[0], bp[1], rgbd_level_size(h->cols, l - 1),
                                                                    bl[0], bl[1], c, r);
    }
    for (int l = 0; l < h->levels; ++l) {
        const int c = rgbd_level_size(h->cols, l), r = rgbd_level_size(h->rows, l);
        float* const* bl = h->buf[slot][l];
        rgbd_sobel_kernel<<<dim3(divup(c, 256), r), 256, 0, stream()>>>(bl[0], c, r, bl[2], bl[3]);
    }
    KFX_CHECK_LAUNCH();
    h->have[slot] = true;
    return KFX_OK;
}

int kfx_rgbd_swap(kfx_rgbd* h) {
    KFX_REQUIRE(h != nullptr, "kfx_rgbd_swap: null handle");
    for (int l = 0; l < kRgbdMaxLevels; ++l)
        for (int k = 0; k < 4; ++k) std::swap(h->buf[0][l][k], h->buf[1][l][k]);
    std::swap(h->have[0], h->have[1]);
    return KFX_OK;
}

int kfx_rgbd_download_level(kfx_rgbd* h, int slot, int level, float* intensity, float* depth, float* dx, float* dy) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && (slot == 0 || slot == 1) && level >= 0 && level < h->levels,
                "kfx_rgbd_download_level: bad arguments");
    const size_t n = (size_t)rgbd_level_size(h->cols, level) * rgbd_level_size(h->rows, level);
    float* dst[4] = {intensity, depth, dx, dy};
    for (int k = 0; k < 4; ++k)
        if (dst[k])
            KFX_CHECK_HIP(hipMemcpyAsync(dst[k], h->buf[slot][level][k], sizeof(float) * n, hipMemcpyDeviceToHost,
                                         stream()));
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    return KFX_OK;
}

int kfx_rgbd_estimate(kfx_rgbd* h, const kfx_intrinsics* intr, const kfx_rgbd_params* prm, const kfx_pose* init,
                      kfx_rgbd_result* out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && intr && prm && init && out, "kfx_rgbd_estimate: null pointer");
    KFX_REQUIRE(h->have[0] && h->have[1], "kfx_rgbd_estimate: set both frames first");
    KFX_REQUIRE(intr->cols == h->cols && intr->rows == h->rows, "kfx_rgbd_estimate: intrinsics of %dx%d, handle %dx%d",
                intr->cols, intr->rows, h->cols, h->rows);
    KFX_REQUIRE(prm->levels >= 1 && prm->levels <= h->levels, "kfx_rgbd_estimate: %d levels (handle has %d)",
                prm->levels, h->levels);
    KFX_REQUIRE(prm->max_depth_diff > 0.0f && prm->min_gradient >= 0.0f, "kfx_rgbd_estimate: bad thresholds");
    ProfScope prof("rgbd");
    rgbd_init_kernel<<<1, 64, 0, stream()>>>(h->state, *init);
    const float g2 = prm->min_gradient * prm->min_gradient;
    for (int l = prm->levels - 1; l >= 0; --l) {
        KFX_REQUIRE(prm->iterations[l] >= 0, "kfx_rgbd_estimate: negative iteration count");
        const float d = (float)(1 << l);
        RgbdLevel L;
        L.i_prev = h->buf[0][l][0];
        L.d_prev = h->buf[0][l][1];
        L.i_curr = h->buf[1][l][0];
        L.d_curr = h->buf[1][l][1];
        L.dx_curr = h->buf[1][l][2];
        L.dy_curr = h->buf[1][l][3];
        L.cols = rgbd_level_size(h->cols, l);
        L.rows = rgbd_level_size(h->rows, l);
        L.fx = intr->fx / d;
        L.fy = intr->fy / d;
        L.cx = intr->cx / d;
        L.cy = intr->cy / d;
        const int n = L.cols * L.rows, nblk = divup(n, kRgbdThreads);
        for (int it = 0; it < prm->iterations[l]; ++it) {
            rgbd_clear_kernel<<<divup(n, 256), 256, 0, stream()>>>(h->corr, n, h->state, it == 0);
            rgbd_corr_kernel<<<dim3(divup(L.cols, 256), L.rows), 256, 0, stream()>>>(L, h->state, g2,
                                                                                     prm->max_depth_diff, h->corr);
            rgbd_stats_kernel<<<nblk, kRgbdThreads, 0, stream()>>>(L, h->state, h->corr, h->part);
            rgbd_sigma_kernel<<<1, kRgbdThreads, 0, stream()>>>(h->state, h->part, nblk);
            rgbd_sums_kernel<<<nblk, kRgbdThreads, 0, stream()>>>(L, h->state, h->corr, h->part);
            rgbd_solve_kernel<<<1, kRgbdThreads, 0, stream()>>>(h->state, h->part, nblk);
        }
    }
    KFX_CHECK_LAUNCH();
    KFX_CHECK_HIP(hipMemcpyAsync(h->host_state, h->state, sizeof(RgbdState), hipMemcpyDeviceToHost, stream()));
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    const RgbdState& s = *h->host_state;
    for (int i = 0; i < 9; ++i) out->pose.R[i] = s.R[i];
    for (int i = 0; i < 3; ++i) out->pose.t[i] = s.t[i];
    for (int i = 0; i < 36; ++i) out->A[i] = s.A[i];
    for (int i = 0; i < 6; ++i) out->b[i] = s.b[i];
    out->sigma = s.sigma;
    out->correspondences = s.count;
    out->solves = s.solves;
    out->ok = s.ok;
    return KFX_OK;
}

}  // extern "C"
