// kfx_preprocess.hip — per-frame depth preprocessing: bilateral filter,
// depth pyramid, vertex and normal maps, rigid transform of maps and the 2x
// downsampling of raycast maps for the coarser ICP levels.
//
// Images are row-major.  Depth is uint16 millimetres; vertex and normal maps
// are arrays of 12-byte XYZ structures (the reference's Point_XYZ layout),
// NaN in x marking an invalid pixel.  The stencil kernels stage a tile plus
// halo in LDS so each depth/vertex sample is read from HBM once per tile.
#include <math.h>
#include <string.h>

#include "kfx_common.h"

namespace kfx {

// --------------------------------------------------------------------------
// Bilateral filter (GPU/opencl/compute_bilateral_filtering.cl:8-57).
// Window [c-6, min(c+7, dim-1)) per axis: the far-edge row and column of the
// image never enter a window, as in the reference.  Weight
//   exp(-(space2 * 0.5/sigma_s^2 + color2 * 0.5/sigma_r^2)),
// output clamp(rint(sum(w d) / sum(w)), 0, 32767), zeroed above
// max_distance_mm.  (The reference kernel stores rint(sum(w d)) instead of
// the normalised value it computes; this is its intended output.)
// --------------------------------------------------------------------------
constexpr int kBilR = 6;
constexpr int kBilTX = 64, kBilTY = 4;
constexpr int kBilLX = kBilTX + 2 * kBilR, kBilLY = kBilTY + 2 * kBilR;

__global__ void __launch_bounds__(kBilTX * kBilTY)
bilateral_kernel(const uint16_t* __restrict__ src, uint16_t* __restrict__ dst, int cols, int rows,
                 float space_coef, float color_coef, uint32_t max_distance) {
    __shared__ uint16_t tile[kBilLY][kBilLX];
    const int bx = blockIdx.x * kBilTX, by = blockIdx.y * kBilTY;
    const int tid = threadIdx.y * kBilTX + threadIdx.x;
    for (int i = tid; i < kBilLX * kBilLY; i += kBilTX * kBilTY) {
        const int ly = i / kBilLX, lx = i - ly * kBilLX;
        const int gx = bx + lx - kBilR, gy = by + ly - kBilR;
        tile[ly][lx] = (gx >= 0 && gx < cols && gy >= 0 && gy < rows) ? src[gy * cols + gx] : (uint16_t)0;
    }
    __syncthreads();
    const int c = bx + threadIdx.x, r = by + threadIdx.y;
    if (c >= cols || r >= rows) return;
    const int value = tile[threadIdx.y + kBilR][threadIdx.x + kBilR];
    const int max_x = min(c + kBilR + 1, cols - 1), max_y = min(r + kBilR + 1, rows - 1);
    const int min_x = max(c - kBilR, 0), min_y = max(r - kBilR, 0);
    float sum1 = 0.0f, sum2 = 0.0f;
    for (int cy = min_y; cy < max_y; ++cy) {
        const uint16_t* row = &tile[cy - by + kBilR][0];
        const float dy2 = (float)((r - cy) * (r - cy));
        for (int cx = min_x; cx < max_x; ++cx) {
            const int tmp = row[cx - bx + kBilR];
            const float space2 = (float)((c - cx) * (c - cx)) + dy2;
            const float diff = (float)(value - tmp);
            const float w = __expf(-(space2 * space_coef + diff * diff * color_coef));
            sum1 += (float)tmp * w;
            sum2 += w;
        }
    }
    int res = (int)rintf(sum1 / sum2);
    res = max(0, min(res, KFX_MAX_DEPTH_VALUE));
    if ((uint32_t)res > max_distance) res = 0;
    dst[r * cols + c] = (uint16_t)res;
}

// --------------------------------------------------------------------------
// Depth pyramid (compute_depth_pyramid; call site opencl_wrappers.cpp:244-271,
// kernel absent — KinectFusion pyrDown): dst(x, y) averages the 5x5
// neighbourhood of src(2x, 2y) with separable weights {1,4,6,4,1}/16 over the
// samples within 3 sigma_r (90 mm) of the centre, truncated to integer mm.
// --------------------------------------------------------------------------
// The range test |d - centre| < 3 sigma_r on integer differences d (|d| <
// 2^16, exact in float) as |d| < thr with thr = ceil(3 sigma_r) clamped to
// [0, 65536]; 0 (no tap passes) for a NaN or non-positive sigma_r.
__device__ __forceinline__ int pyr_threshold(float sigma_color) {
    const float c = ceilf(3.0f * sigma_color);
    return !(c > 0.0f) ? 0 : (c >= 65536.0f ? 65536 : (int)c);
}

// One output pixel (x, y) of the 2x downsample from a source of scols x srows
// read through at(xx, yy), thr from pyr_threshold.  The weights are the
// integers 16 wt[|xi|] x 16 wt[|yi|] of {6, 4, 1} x {6, 4, 1}: every term of
// the float sums sum(d w) and sum(w) is a multiple of 1/256 and every partial
// sum is below 2^16 (the weights add up to 1), so those float sums are exact
// and equal the integer sums / 256, and their quotient is the correctly
// rounded s / wsum of the integers.
template <typename At>
__device__ __forceinline__ uint16_t pyr_down_px(At at, int scols, int srows, int x, int y, int thr) {
    constexpr int kW[3] = {6, 4, 1};
    const int center = at(2 * x, 2 * y);
    const int x_lo = max(0, 2 * x - 2) - 2 * x, y_lo = max(0, 2 * y - 2) - 2 * y;
    const int x_hi = min(scols, 2 * x + 3) - 2 * x, y_hi = min(srows, 2 * y + 3) - 2 * y;
    // The 25 taps' loads are issued together (taps outside the image read
    // the centre and are masked).
    int val[5][5];
#pragma unroll
    for (int yi = -2; yi <= 2; ++yi)
#pragma unroll
        for (int xi = -2; xi <= 2; ++xi) {
            const bool in = yi >= y_lo && yi < y_hi && xi >= x_lo && xi < x_hi;
            val[yi + 2][xi + 2] = at(in ? 2 * x + xi : 2 * x, in ? 2 * y + yi : 2 * y);
        }
    int s = 0, wsum = 0;
#pragma unroll
    for (int yi = -2; yi <= 2; ++yi)
#pragma unroll
        for (int xi = -2; xi <= 2; ++xi) {
            const int v = val[yi + 2][xi + 2];
            const bool in = yi >= y_lo && yi < y_hi && xi >= x_lo && xi < x_hi;
            const int w = kW[xi < 0 ? -xi : xi] * kW[yi < 0 ? -yi : yi];
            const int wt = in && abs(v - center) < thr ? w : 0;
            s += v * wt;
            wsum += wt;
        }
    return (uint16_t)(int)((float)s / (float)wsum);
}

__global__ void __launch_bounds__(256) pyr_down_kernel(const uint16_t* __restrict__ src, int scols, int srows,
                                                       uint16_t* __restrict__ dst, int dcols, int drows,
                                                       float sigma_color) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= dcols || y >= drows) return;
    dst[y * dcols + x] = pyr_down_px([&](int xx, int yy) { return (int)src[yy * scols + xx]; }, scols, srows, x, y,
                                     pyr_threshold(sigma_color));
}

// --------------------------------------------------------------------------
// Vertex map (compute_vertex_map; call site opencl_wrappers.cpp:281-305,
// kernel absent): z = d / 1000 m; v = (z (u - cx) / fx, z (v - cy) / fy, z);
// NaN where d == 0.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) vertex_map_kernel(const uint16_t* __restrict__ depth,
                                                         kfx_xyz* __restrict__ vmap, kfx_intrinsics in,
                                                         float fx_inv, float fy_inv) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y * blockDim.y + threadIdx.y;
    if (u >= in.cols || v >= in.rows) return;
    const int i = v * in.cols + u;
    const float z = (float)depth[i] / 1000.0f;
    kfx_xyz out{NAN, NAN, NAN};
    if (z != 0.0f) out = kfx_xyz{z * ((float)u - in.cx) * fx_inv, z * ((float)v - in.cy) * fy_inv, z};
    vmap[i] = out;
}

// --------------------------------------------------------------------------
// Normal map (compute_normal_map_eigen; call site opencl_wrappers.cpp:315-343,
// kernel absent — KinectFusion "eigen" normals): over the window
// [u-3, min(u+4, cols-1)) x [v-3, min(v+4, rows-1)) of valid vertices (at
// least 24 of them), the normal is the eigenvector of the smallest eigenvalue
// of the covariance about the centroid, oriented toward the camera (n . v < 0).
// --------------------------------------------------------------------------
constexpr int kNmR = 3;
constexpr int kNmTX = 32, kNmTY = 8;
constexpr int kNmLX = kNmTX + 2 * kNmR + 1, kNmLY = kNmTY + 2 * kNmR + 1;

// Eigenvector of the smallest eigenvalue of a symmetric 3x3 matrix
// {a00, a01, a02, a11, a12, a22}: eigenvalues from the trigonometric form of
// the characteristic cubic (O. K. Smith, CACM 4(4), 1961), eigenvector as the
// largest cross product of two rows of (A - lambda I).
__device__ void smallest_eigvec(const float* cov, float& nx, float& ny, float& nz) {
    float scale = fmaxf(fmaxf(fabsf(cov[0]), fabsf(cov[1])),
                        fmaxf(fmaxf(fabsf(cov[2]), fabsf(cov[3])), fmaxf(fabsf(cov[4]), fabsf(cov[5]))));
    if (!(scale > 0.0f)) { nx = ny = nz = NAN; return; }
    const float inv = 1.0f / scale;
    const float a00 = cov[0] * inv, a01 = cov[1] * inv, a02 = cov[2] * inv;
    const float a11 = cov[3] * inv, a12 = cov[4] * inv, a22 = cov[5] * inv;
    const float q = (a00 + a11 + a22) * (1.0f / 3.0f);
    const float p1 = a01 * a01 + a02 * a02 + a12 * a12;
    const float b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
    const float p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0f * p1;
    float lambda;
    if (p2 <= 1e-30f) {
        lambda = q;
    } else {
        const float p = sqrtf(p2 * (1.0f / 6.0f));
        const float ip = 1.0f / p;
        const float c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip;
        const float c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
        const float det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) + c02 * (c01 * c12 - c11 * c02);
        const float r = fminf(1.0f, fmaxf(-1.0f, 0.5f * det));
        const float phi = acosf(r) * (1.0f / 3.0f);
        lambda = q + 2.0f * p * cosf(phi + 2.0943951023931957f);
    }
    const float r0x = a00 - lambda, r0y = a01, r0z = a02;
    const float r1x = a01, r1y = a11 - lambda, r1z = a12;
    const float r2x = a02, r2y = a12, r2z = a22 - lambda;
    const float c0x = r0y * r1z - r0z * r1y, c0y = r0z * r1x - r0x * r1z, c0z = r0x * r1y - r0y * r1x;
    const float c1x = r0y * r2z - r0z * r2y, c1y = r0z * r2x - r0x * r2z, c1z = r0x * r2y - r0y * r2x;
    const float c2x = r1y * r2z - r1z * r2y, c2y = r1z * r2x - r1x * r2z, c2z = r1x * r2y - r1y * r2x;
    const float d0 = c0x * c0x + c0y * c0y + c0z * c0z;
    const float d1 = c1x * c1x + c1y * c1y + c1z * c1z;
    const float d2 = c2x * c2x + c2y * c2y + c2z * c2z;
    float vx = c0x, vy = c0y, vz = c0z, dm = d0;
    if (d1 > dm) { vx = c1x; vy = c1y; vz = c1z; dm = d1; }
    if (d2 > dm) { vx = c2x; vy = c2y; vz = c2z; dm = d2; }
    if (!(dm > 0.0f)) { nx = ny = nz = NAN; return; }
    const float s = 1.0f / sqrtf(dm);
    nx = vx * s; ny = vy * s; nz = vz * s;
}

typedef float NmTile[kNmLY][kNmLX];

// Normal of pixel (u, v) from the vertex tile in LDS (tile position lu, lv).
__device__ __forceinline__ kfx_xyz tile_normal(const NmTile& sx, const NmTile& sy, const NmTile& sz, int lu, int lv,
                                               int u, int v, int cols, int rows) {
    kfx_xyz out{NAN, NAN, NAN};
    const float pcx = sx[lv][lu], pcy = sy[lv][lu], pcz = sz[lv][lu];
    if (!isnan(pcx)) {
        // Window rows [max(v-3, 0), min(v+4, rows-1)), columns likewise:
        // out-of-image taps are NaN in the tile; the far-edge row and column
        // are excluded by these limits.  One pass over the 49 taps, fully
        // unrolled so the LDS reads are issued together, accumulating the
        // count, the sum and the second moments of the taps relative to the
        // centre vertex (small numbers, so the covariance
        // S2 - S1 S1^T / n keeps its precision).
        const int dymax = min(kNmR, rows - 2 - v), dxmax = min(kNmR, cols - 2 - u);
        float n = 0.0f, s1x = 0.0f, s1y = 0.0f, s1z = 0.0f;
        float sxx = 0.0f, sxy = 0.0f, sxz = 0.0f, syy = 0.0f, syz = 0.0f, szz = 0.0f;
#pragma unroll 1
        for (int dy = -kNmR; dy <= kNmR; ++dy)
#pragma unroll
            for (int dx = -kNmR; dx <= kNmR; ++dx) {
                const float px = sx[lv + dy][lu + dx], py = sy[lv + dy][lu + dx], pz = sz[lv + dy][lu + dx];
                const bool ok = (px == px) & (dy <= dymax) & (dx <= dxmax);
                const float qx = ok ? px - pcx : 0.0f;
                const float qy = ok ? py - pcy : 0.0f;
                const float qz = ok ? pz - pcz : 0.0f;
                n += ok ? 1.0f : 0.0f;
                s1x += qx; s1y += qy; s1z += qz;
                // Second moments with fused multiply-adds (the normals are
                // compared with the double-precision oracle within a
                // tolerance, not bit for bit).
                sxx = fmaf(qx, qx, sxx); sxy = fmaf(qx, qy, sxy); sxz = fmaf(qx, qz, sxz);
                syy = fmaf(qy, qy, syy); syz = fmaf(qy, qz, syz); szz = fmaf(qz, qz, szz);
            }
        if (n >= (float)((2 * kNmR + 1) * (2 * kNmR + 1) / 2)) {
            const float in = 1.0f / n;
            const float mx = s1x * in, my = s1y * in, mz = s1z * in;
            float cov[6] = {sxx - s1x * mx, sxy - s1x * my, sxz - s1x * mz,
                            syy - s1y * my, syz - s1y * mz, szz - s1z * mz};
            float nx, ny, nz;
            smallest_eigvec(cov, nx, ny, nz);
            if (!isnan(nx)) {
                if (nx * pcx + ny * pcy + nz * pcz > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
                out = kfx_xyz{nx, ny, nz};
            }
        }
    }
    return out;
}

__global__ void __launch_bounds__(kNmTX * kNmTY)
normal_map_kernel(const kfx_xyz* __restrict__ vmap, kfx_xyz* __restrict__ nmap, int cols, int rows) {
    __shared__ NmTile sx, sy, sz;
    const int bx = blockIdx.x * kNmTX, by = blockIdx.y * kNmTY;
    const int tid = threadIdx.y * kNmTX + threadIdx.x;
    for (int i = tid; i < kNmLX * kNmLY; i += kNmTX * kNmTY) {
        const int ly = i / kNmLX, lx = i - ly * kNmLX;
        const int gx = bx + lx - kNmR, gy = by + ly - kNmR;
        kfx_xyz v{NAN, NAN, NAN};
        if (gx >= 0 && gx < cols && gy >= 0 && gy < rows) v = vmap[gy * cols + gx];
        sx[ly][lx] = v.x; sy[ly][lx] = v.y; sz[ly][lx] = v.z;
    }
    __syncthreads();
    const int u = bx + threadIdx.x, v = by + threadIdx.y;
    if (u >= cols || v >= rows) return;
    nmap[v * cols + u] = tile_normal(sx, sy, sz, threadIdx.x + kNmR, threadIdx.y + kNmR, u, v, cols, rows);
}

// Vertex and normal maps of every pyramid level in one launch: each
// workgroup owns one 32 x 8 tile of one level, back-projects its depth tile
// and halo into LDS (vertex_map_kernel's arithmetic), writes the tile's
// vertices and computes its normals from the same LDS tile (normal_map_kernel's
// arithmetic): the per-level kernels' results, bit for bit.
struct MapsLevel {
    const uint16_t* depth;
    kfx_xyz* vmap;
    kfx_xyz* nmap;
    kfx_intrinsics in;
    float fx_inv, fy_inv;
    int tiles_x, first_tile;
};
struct MapsParams {
    MapsLevel lv[KFX_ICP_LEVELS];
    int levels;
};

__global__ void __launch_bounds__(kNmTX * kNmTY) maps_kernel(const MapsParams P) {
    __shared__ NmTile sx, sy, sz;
    const int t = blockIdx.x;
    int l = 0;
    for (int k = 1; k < P.levels; ++k) l = t >= P.lv[k].first_tile ? k : l;
    const MapsLevel& L = P.lv[l];
    const int tl = t - L.first_tile;
    const int bx = (tl % L.tiles_x) * kNmTX, by = (tl / L.tiles_x) * kNmTY;
    const int cols = L.in.cols, rows = L.in.rows;
    const int tid = threadIdx.y * kNmTX + threadIdx.x;
    for (int i = tid; i < kNmLX * kNmLY; i += kNmTX * kNmTY) {
        const int ly = i / kNmLX, lx = i - ly * kNmLX;
        const int gx = bx + lx - kNmR, gy = by + ly - kNmR;
        kfx_xyz v{NAN, NAN, NAN};
        if (gx >= 0 && gx < cols && gy >= 0 && gy < rows) {
            const float z = (float)L.depth[gy * cols + gx] / 1000.0f;
            if (z != 0.0f) v = kfx_xyz{z * ((float)gx - L.in.cx) * L.fx_inv, z * ((float)gy - L.in.cy) * L.fy_inv, z};
        }
        sx[ly][lx] = v.x; sy[ly][lx] = v.y; sz[ly][lx] = v.z;
    }
    __syncthreads();
    const int u = bx + threadIdx.x, v = by + threadIdx.y;
    if (u >= cols || v >= rows) return;
    const int lu = threadIdx.x + kNmR, lv = threadIdx.y + kNmR;
    L.vmap[v * cols + u] = kfx_xyz{sx[lv][lu], sy[lv][lu], sz[lv][lu]};
    L.nmap[v * cols + u] = tile_normal(sx, sy, sz, lu, lv, u, v, cols, rows);
}

// The same maps from level-0 depth alone: a tile of level l >= 1 derives its
// depth halo with pyr_down_px from an LDS region of level l - 1, and level
// 2's region of level 1 is itself derived from a staged region of level 0,
// so no depth pyramid is written beforehand.  A halo of hx x kNmLY pixels
// (hx = tile width + 7) reads (2 hx + 3) x (2 kNmLY + 3) pixels of the level
// above (5x5 taps about every second pixel), and that region reads twice as
// much again of the level above it.  Level-2 tiles, whose two derivations
// make them the costliest, are kL2TX = 16 pixels wide (half the other
// levels' 32) so that their work spreads over more workgroups (8 wide
// repeats more of the level-1 derivation, 32 wide leaves the launch waiting
// on 75 long tiles), and they come first in the grid.
constexpr int kL2TX = 16;
template <int TW>
struct DeriveGeom {
    static constexpr int kHX = TW + 2 * kNmR + 1;                 // halo width
    static constexpr int k1X = 2 * kHX + 3, k1Y = 2 * kNmLY + 3;  // region of level l - 1
    static constexpr int k0X = 2 * k1X + 3, k0Y = 2 * k1Y + 3;    // its region of level l - 2
};
static_assert(KFX_ICP_LEVELS <= 3, "maps_derive_kernel derives at most two levels");
static_assert(DeriveGeom<kL2TX>::kHX <= kNmLX, "a level-2 halo fits the vertex tile");

struct MapsDeriveParams {
    MapsLevel lv[KFX_ICP_LEVELS];  // lv[l].depth unused for l >= 1
    uint16_t* depth_out[KFX_ICP_LEVELS];  // derived depth of level l >= 1 (tile cores), or null
    int levels;
    float sigma_color;
    DepthBoundJob job;  // the integration's depth-bound tiles, the grid's first job.tiles workgroups
};

// One 32 x 32 tile of the integration pre-pass's pose-independent half
// (integrate_prepass_kernel's arithmetic, by 256 threads): the tile's scaled
// depth (0 outside the image), then levels 1..5 of the depth-bound pyramid
// by repeated 2 x 2 maxima in LDS (t, 1024 floats).
__device__ __forceinline__ void depth_bound_tile(const DepthBoundJob& J, int tb, int tid, float* t) {
    constexpr int kThreads = kNmTX * kNmTY;
    const int cols = J.intr.cols, rows = J.intr.rows;
    const int bx = tb % J.tiles_x, by = tb / J.tiles_x;
#pragma unroll
    for (int k = 0; k < kBoundTile * kBoundTile / kThreads; ++k) {
        const int i = tid + k * kThreads;
        const int u = bx * kBoundTile + i % kBoundTile, v = by * kBoundTile + i / kBoundTile;
        float d = 0.0f;
        if (u < cols && v < rows) {
            d = scaled_depth(J.depth, J.intr, u, v);
            J.scaled[v * cols + u] = d;
        }
        t[i] = d;
    }
    __syncthreads();
    for (int l = 1; l <= kBoundLevels; ++l) {
        const int n = kBoundTile >> l;        // texels per tile side at level l (n * n <= 256)
        const int sd = kBoundTile >> (l - 1);  // side of the previous level's tile in t[]
        const int tx = tid % n, ty = tid / n;
        const bool own = tid < n * n;
        float m = 0.0f;
        if (own)
            m = fmaxf(fmaxf(t[(2 * ty) * sd + 2 * tx], t[(2 * ty) * sd + 2 * tx + 1]),
                      fmaxf(t[(2 * ty + 1) * sd + 2 * tx], t[(2 * ty + 1) * sd + 2 * tx + 1]));
        __syncthreads();
        if (own) {
            t[ty * n + tx] = m;
            const int lw = (cols + (1 << l) - 1) >> l, lh = (rows + (1 << l) - 1) >> l;
            const int a = bx * n + tx, b = by * n + ty;
            if (a < lw && b < lh) J.bound[J.bound_off[l] + b * lw + a] = m;
        }
        __syncthreads();
    }
}

// Copies the W x H region of src (scols x srows) whose corner is (ox, oy)
// into dst (row stride W).  Out-of-image positions hold a clamped in-image
// pixel and are never read, since pyr_down_px reads only in-image taps.  The
// loop has a fixed trip count and no branch (the last pass's spare lanes
// repeat the final element), so its loads are issued together rather than
// one latency per pass.
template <int W, int H>
__device__ __forceinline__ void stage_depth(const uint16_t* __restrict__ src, int scols, int srows, int ox, int oy,
                                            uint16_t* dst, int tid) {
    constexpr int kN = W * H, kThreads = kNmTX * kNmTY, kPasses = (kN + kThreads - 1) / kThreads;
    uint16_t val[kPasses];
#pragma unroll
    for (int k = 0; k < kPasses; ++k) {
        const int i = min(tid + k * kThreads, kN - 1);
        const int ry = i / W, rx = i - ry * W;
        const int x = min(max(ox + rx, 0), scols - 1), y = min(max(oy + ry, 0), srows - 1);
        val[k] = src[y * scols + x];
    }
#pragma unroll
    for (int k = 0; k < kPasses; ++k) dst[min(tid + k * kThreads, kN - 1)] = val[k];
}

// Downsampled pixel (x, y) of a scols x srows level held in LDS as the
// region s (row stride w, corner (ox, oy)) that covers its 5x5 window.
__device__ __forceinline__ int pyr_down_lds(const uint16_t* s, int w, int ox, int oy, int scols, int srows, int x,
                                            int y, int thr) {
    return pyr_down_px([&](int xx, int yy) { return (int)s[(yy - oy) * w + (xx - ox)]; }, scols, srows, x, y, thr);
}

// One tile (tile index tl within level l) of tile width TW.
template <int TW>
__device__ __forceinline__ void maps_derive_tile(const MapsDeriveParams& P, int l, int tl, int tid, uint16_t* d0,
                                                 NmTile& vx, NmTile& vy, NmTile& vz, uint16_t* d1) {
    using G = DeriveGeom<TW>;
    const MapsLevel& L = P.lv[l];
    const int bx = (tl % L.tiles_x) * TW, by = (tl / L.tiles_x) * kNmTY;
    const int hx = bx - kNmR, hy = by - kNmR;  // halo corner
    const int cols = L.in.cols, rows = L.in.rows;
    const int thr = pyr_threshold(P.sigma_color);
    const int sox = 2 * hx - 2, soy = 2 * hy - 2;  // corner of the level l - 1 region
    if (l == 2) {
        const kfx_intrinsics& i0 = P.lv[0].in;
        const kfx_intrinsics& i1 = P.lv[1].in;
        const int o0x = 2 * sox - 2, o0y = 2 * soy - 2;
        stage_depth<G::k0X, G::k0Y>(P.lv[0].depth, i0.cols, i0.rows, o0x, o0y, d0, tid);
        __syncthreads();
        for (int i = tid; i < G::k1X * G::k1Y; i += kNmTX * kNmTY) {
            const int ry = i / G::k1X, rx = i - ry * G::k1X;
            const int x = sox + rx, y = soy + ry;
            if (x >= 0 && x < i1.cols && y >= 0 && y < i1.rows)
                d1[i] = (uint16_t)pyr_down_lds(d0, G::k0X, o0x, o0y, i0.cols, i0.rows, x, y, thr);
        }
        __syncthreads();
    } else if (l == 1) {
        stage_depth<G::k1X, G::k1Y>(P.lv[0].depth, P.lv[0].in.cols, P.lv[0].in.rows, sox, soy, d1, tid);
        __syncthreads();
    }
    const kfx_intrinsics& is = P.lv[l > 0 ? l - 1 : 0].in;  // source level of a derived tile
    uint16_t* dout = l > 0 ? P.depth_out[l] : nullptr;
    for (int i = tid; i < G::kHX * kNmLY; i += kNmTX * kNmTY) {
        const int ly = i / G::kHX, lx = i - ly * G::kHX;
        const int gx = hx + lx, gy = hy + ly;
        kfx_xyz v{NAN, NAN, NAN};
        if (gx >= 0 && gx < cols && gy >= 0 && gy < rows) {
            int d;
            if (l == 0) {
                d = L.depth[gy * cols + gx];
            } else {
                d = pyr_down_lds(d1, G::k1X, sox, soy, is.cols, is.rows, gx, gy, thr);
                if (dout && lx >= kNmR && lx < kNmR + TW && ly >= kNmR && ly < kNmR + kNmTY)
                    dout[gy * cols + gx] = (uint16_t)d;
            }
            const float z = (float)(uint16_t)d / 1000.0f;
            if (z != 0.0f) v = kfx_xyz{z * ((float)gx - L.in.cx) * L.fx_inv, z * ((float)gy - L.in.cy) * L.fy_inv, z};
        }
        vx[ly][lx] = v.x; vy[ly][lx] = v.y; vz[ly][lx] = v.z;
    }
    __syncthreads();
    const int u = bx + threadIdx.x, v = by + threadIdx.y;
    if (threadIdx.x >= TW || u >= cols || v >= rows) return;
    const int lu = threadIdx.x + kNmR, lv = threadIdx.y + kNmR;
    L.vmap[v * cols + u] = kfx_xyz{vx[lv][lu], vy[lv][lu], vz[lv][lu]};
    L.nmap[v * cols + u] = tile_normal(vx, vy, vz, lu, lv, u, v, cols, rows);
}

__global__ void __launch_bounds__(kNmTX * kNmTY) maps_derive_kernel(const MapsDeriveParams P) {
    using G2 = DeriveGeom<kL2TX>;
    using G1 = DeriveGeom<kNmTX>;
    __shared__ union {
        uint16_t d0[G2::k0X * G2::k0Y];  // level 0 under a level-2 tile's level-1 region
        struct {
            NmTile x, y, z;
        } v;  // the tile's vertices (written once d0 is consumed)
    } lds;
    __shared__ uint16_t d1[G1::k1X * G1::k1Y > G2::k1X * G2::k1Y ? G1::k1X * G1::k1Y : G2::k1X * G2::k1Y];
    const int tid = threadIdx.y * kNmTX + threadIdx.x;
    if ((int)blockIdx.x < P.job.tiles) {
        static_assert(sizeof(lds.d0) >= kBoundTile * kBoundTile * sizeof(float), "a bound tile fits in d0");
        depth_bound_tile(P.job, blockIdx.x, tid, reinterpret_cast<float*>(lds.d0));
        return;
    }
    // Tiles run from the deepest level to level 0.
    const int t = blockIdx.x - P.job.tiles;
    int l = P.levels - 1;
    for (int k = P.levels - 2; k >= 0; --k) l = t >= P.lv[k].first_tile ? k : l;
    const int tl = t - P.lv[l].first_tile;
    if (l == 2)
        maps_derive_tile<kL2TX>(P, 2, tl, tid, lds.d0, lds.v.x, lds.v.y, lds.v.z, d1);
    else
        maps_derive_tile<kNmTX>(P, l, tl, tid, lds.d0, lds.v.x, lds.v.y, lds.v.z, d1);
}

// --------------------------------------------------------------------------
// transform_map (GPU/opencl/transform_map.cl:6-65): v' = R v + t, n' = R n,
// NaN passthrough keyed on the x component.
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(256) transform_maps_kernel(const kfx_xyz* __restrict__ vs,
                                                             const kfx_xyz* __restrict__ ns, kfx_matrix_3f R,
                                                             kfx_point_3f t, int cols, int rows,
                                                             kfx_xyz* __restrict__ vd, kfx_xyz* __restrict__ nd) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= cols || y >= rows) return;
    const int i = x + cols * y;
    auto mul = [&](kfx_xyz a) {
        kfx_xyz o;
        float s;
        s = 0.0f; s = s + R.data[0].x * a.x; s = s + R.data[0].y * a.y; s = s + R.data[0].z * a.z; o.x = s;
        s = 0.0f; s = s + R.data[1].x * a.x; s = s + R.data[1].y * a.y; s = s + R.data[1].z * a.z; o.y = s;
        s = 0.0f; s = s + R.data[2].x * a.x; s = s + R.data[2].y * a.y; s = s + R.data[2].z * a.z; o.z = s;
        return o;
    };
    const kfx_xyz v = vs[i];
    kfx_xyz vo{NAN, NAN, NAN};
    if (!isnan(v.x)) {
        vo = mul(v);
        vo.x = vo.x + t.x; vo.y = vo.y + t.y; vo.z = vo.z + t.z;
    }
    vd[i] = vo;
    const kfx_xyz n = ns[i];
    kfx_xyz no{NAN, NAN, NAN};
    if (!isnan(n.x)) no = mul(n);
    nd[i] = no;
}

// --------------------------------------------------------------------------
// resize_vertex_map / resize_normal_map (call sites opencl_wrappers.cpp:97-146,
// kernels absent — KinectFusion resizeMap): out(x, y) is the mean of the 2x2
// block at (2x, 2y) when all four are valid, NaN otherwise; normals are
// renormalised.
// --------------------------------------------------------------------------
// (resize4, the mean of a 2x2 block, is in kfx_common.h: the raycast's
// fused pyramid uses it too.)

// Output pixel (x, y) of one 2x downsampling of `in` (icols wide).
__device__ __forceinline__ kfx_xyz resize_at(const kfx_xyz* __restrict__ in, int icols, int x, int y, bool normalize) {
    return resize4(in[(2 * y) * icols + 2 * x], in[(2 * y) * icols + 2 * x + 1], in[(2 * y + 1) * icols + 2 * x],
                   in[(2 * y + 1) * icols + 2 * x + 1], normalize);
}

__global__ void __launch_bounds__(256) resize_map_kernel(const kfx_xyz* __restrict__ in, int icols,
                                                         kfx_xyz* __restrict__ out, int ocols, int orows,
                                                         int normalize) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= ocols || y >= orows) return;
    out[y * ocols + x] = resize_at(in, icols, x, y, normalize != 0);
}

// Levels 1 and 2 of the vertex and normal maps from level 0 in one launch:
// a level-2 pixel recomputes its four level-1 inputs from level 0 with the
// same arithmetic, so the result is two resize_map_kernel passes bit for bit.
__global__ void __launch_bounds__(256) resize_pyramid_kernel(const kfx_xyz* __restrict__ v0,
                                                             const kfx_xyz* __restrict__ n0, int c0,
                                                             kfx_xyz* __restrict__ v1, kfx_xyz* __restrict__ n1,
                                                             int c1, int r1, kfx_xyz* __restrict__ v2,
                                                             kfx_xyz* __restrict__ n2, int c2, int r2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int p1 = c1 * r1;
    if (i < p1) {
        const int x = i % c1, y = i / c1;
        v1[i] = resize_at(v0, c0, x, y, false);
        n1[i] = resize_at(n0, c0, x, y, true);
        return;
    }
    const int j = i - p1;
    if (j >= c2 * r2) return;
    const int x = j % c2, y = j / c2;
    v2[j] = resize4(resize_at(v0, c0, 2 * x, 2 * y, false), resize_at(v0, c0, 2 * x + 1, 2 * y, false),
                    resize_at(v0, c0, 2 * x, 2 * y + 1, false), resize_at(v0, c0, 2 * x + 1, 2 * y + 1, false), false);
    n2[j] = resize4(resize_at(n0, c0, 2 * x, 2 * y, true), resize_at(n0, c0, 2 * x + 1, 2 * y, true),
                    resize_at(n0, c0, 2 * x, 2 * y + 1, true), resize_at(n0, c0, 2 * x + 1, 2 * y + 1, true), true);
}

}  // namespace kfx

using namespace kfx;

int kfx::maps_pyramid_launch(const uint16_t* depth, uint16_t* const* depth_out, kfx_xyz* const* vmap,
                             kfx_xyz* const* nmap, const kfx_intrinsics* intr, int levels, float sigma_color,
                             const DepthBoundJob* job) {
    KFX_REQUIRE(depth && vmap && nmap && intr, "kfx_maps_pyramid_from_depth: null pointer");
    KFX_REQUIRE(levels >= 1 && levels <= KFX_ICP_LEVELS, "kfx_maps_pyramid_from_depth: levels %d outside [1, %d]",
                levels, KFX_ICP_LEVELS);
    MapsDeriveParams P;
    memset(&P, 0, sizeof(P));
    P.levels = levels;
    P.sigma_color = sigma_color;
    P.lv[0].depth = depth;
    for (int l = 0; l < levels; ++l) {
        KFX_REQUIRE(vmap[l] && nmap[l], "kfx_maps_pyramid_from_depth: null map at level %d", l);
        KFX_REQUIRE(vmap[l] != nmap[l], "kfx_maps_pyramid_from_depth: vertex and normal maps alias at level %d", l);
        KFX_REQUIRE(intr[l].cols > 0 && intr[l].rows > 0, "kfx_maps_pyramid_from_depth: empty level %d", l);
        if (l > 0) {
            KFX_REQUIRE(intr[l].cols == intr[l - 1].cols / 2 && intr[l].rows == intr[l - 1].rows / 2,
                        "kfx_maps_pyramid_from_depth: level %d is %dx%d, not half of level %d (%dx%d)", l,
                        intr[l].cols, intr[l].rows, l - 1, intr[l - 1].cols, intr[l - 1].rows);
            P.depth_out[l] = depth_out ? depth_out[l] : nullptr;
            KFX_REQUIRE(P.depth_out[l] != depth, "kfx_maps_pyramid_from_depth: level %d depth overwrites level 0", l);
        }
        MapsLevel& L = P.lv[l];
        L.vmap = vmap[l];
        L.nmap = nmap[l];
        L.in = intr[l];
        L.fx_inv = 1.0f / intr[l].fx;
        L.fy_inv = 1.0f / intr[l].fy;
        L.tiles_x = (int)divup(intr[l].cols, l == 2 ? kL2TX : kNmTX);
    }
    int tiles = 0;
    for (int l = levels - 1; l >= 0; --l) {
        P.lv[l].first_tile = tiles;
        tiles += P.lv[l].tiles_x * (int)divup(intr[l].rows, kNmTY);
    }
    if (job != nullptr) P.job = *job;
    ProfScope prof("maps_pyramid");
    maps_derive_kernel<<<tiles + P.job.tiles, dim3(kNmTX, kNmTY), 0, stream()>>>(P);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

extern "C" {

int kfx_bilateral_filter(const uint16_t* src, uint16_t* dst, int cols, int rows, float sigma_space,
                         float sigma_color, uint32_t max_distance_mm) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(src && dst, "kfx_bilateral_filter: null pointer");
    KFX_REQUIRE(src != dst, "kfx_bilateral_filter: in-place filtering is not supported");
    KFX_REQUIRE(cols > 0 && rows > 0, "kfx_bilateral_filter: empty image");
    KFX_REQUIRE(sigma_space > 0.0f && sigma_color > 0.0f, "kfx_bilateral_filter: sigmas must be > 0");
    ProfScope prof("bilateral");
    const float space_coef = 0.5f / (sigma_space * sigma_space);
    const float color_coef = 0.5f / (sigma_color * sigma_color);
    dim3 block(kBilTX, kBilTY), grid(divup(cols, kBilTX), divup(rows, kBilTY));
    bilateral_kernel<<<grid, block, 0, stream()>>>(src, dst, cols, rows, space_coef, color_coef, max_distance_mm);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_pyr_down(const uint16_t* src, int src_cols, int src_rows, uint16_t* dst, float sigma_color) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(src && dst, "kfx_pyr_down: null pointer");
    KFX_REQUIRE(src_cols >= 2 && src_rows >= 2, "kfx_pyr_down: source smaller than 2x2");
    ProfScope prof("pyr_down");
    const int dc = src_cols / 2, dr = src_rows / 2;
    dim3 block(64, 4), grid(divup(dc, 64), divup(dr, 4));
    pyr_down_kernel<<<grid, block, 0, stream()>>>(src, src_cols, src_rows, dst, dc, dr, sigma_color);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_vertex_map(const uint16_t* depth, kfx_xyz* vmap, const kfx_intrinsics* intr) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(depth && vmap && intr, "kfx_vertex_map: null pointer");
    KFX_REQUIRE(intr->cols > 0 && intr->rows > 0, "kfx_vertex_map: empty image");
    ProfScope prof("vertex_map");
    dim3 block(64, 4), grid(divup(intr->cols, 64), divup(intr->rows, 4));
    vertex_map_kernel<<<grid, block, 0, stream()>>>(depth, vmap, *intr, 1.0f / intr->fx, 1.0f / intr->fy);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_normal_map(const kfx_xyz* vmap, kfx_xyz* nmap, int cols, int rows) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(vmap && nmap, "kfx_normal_map: null pointer");
    KFX_REQUIRE(vmap != nmap, "kfx_normal_map: in-place is not supported");
    KFX_REQUIRE(cols > 0 && rows > 0, "kfx_normal_map: empty image");
    ProfScope prof("normal_map");
    dim3 block(kNmTX, kNmTY), grid(divup(cols, kNmTX), divup(rows, kNmTY));
    normal_map_kernel<<<grid, block, 0, stream()>>>(vmap, nmap, cols, rows);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_transform_maps(const kfx_xyz* vmap_src, const kfx_xyz* nmap_src, const kfx_matrix_3f* rotation,
                       const kfx_point_3f* translation, int cols, int rows, kfx_xyz* vmap_dst, kfx_xyz* nmap_dst) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(vmap_src && nmap_src && rotation && translation && vmap_dst && nmap_dst,
                "kfx_transform_maps: null pointer");
    KFX_REQUIRE(cols > 0 && rows > 0, "kfx_transform_maps: empty image");
    ProfScope prof("transform_maps");
    dim3 block(64, 4), grid(divup(cols, 64), divup(rows, 4));
    transform_maps_kernel<<<grid, block, 0, stream()>>>(vmap_src, nmap_src, *rotation, *translation, cols, rows,
                                                        vmap_dst, nmap_dst);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_resize_map(const kfx_xyz* in, int in_cols, int in_rows, kfx_xyz* out, int normalize) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(in && out, "kfx_resize_map: null pointer");
    KFX_REQUIRE(in_cols >= 2 && in_rows >= 2, "kfx_resize_map: source smaller than 2x2");
    ProfScope prof("resize_map");
    const int oc = in_cols / 2, orow = in_rows / 2;
    dim3 block(64, 4), grid(divup(oc, 64), divup(orow, 4));
    resize_map_kernel<<<grid, block, 0, stream()>>>(in, in_cols, out, oc, orow, normalize);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_maps_pyramid(const uint16_t* const* depth, kfx_xyz* const* vmap, kfx_xyz* const* nmap,
                     const kfx_intrinsics* intr, int levels) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(depth && vmap && nmap && intr, "kfx_maps_pyramid: null pointer");
    KFX_REQUIRE(levels >= 1 && levels <= KFX_ICP_LEVELS, "kfx_maps_pyramid: levels %d outside [1, %d]", levels,
                KFX_ICP_LEVELS);
    MapsParams P;
    memset(&P, 0, sizeof(P));
    P.levels = levels;
    int tiles = 0;
    for (int l = 0; l < levels; ++l) {
        KFX_REQUIRE(depth[l] && vmap[l] && nmap[l], "kfx_maps_pyramid: null map at level %d", l);
        KFX_REQUIRE(vmap[l] != nmap[l], "kfx_maps_pyramid: vertex and normal maps alias at level %d", l);
        KFX_REQUIRE(intr[l].cols > 0 && intr[l].rows > 0, "kfx_maps_pyramid: empty level %d", l);
        MapsLevel& L = P.lv[l];
        L.depth = depth[l];
        L.vmap = vmap[l];
        L.nmap = nmap[l];
        L.in = intr[l];
        L.fx_inv = 1.0f / intr[l].fx;
        L.fy_inv = 1.0f / intr[l].fy;
        L.tiles_x = (int)divup(intr[l].cols, kNmTX);
        L.first_tile = tiles;
        tiles += L.tiles_x * (int)divup(intr[l].rows, kNmTY);
    }
    ProfScope prof("maps_pyramid");
    maps_kernel<<<tiles, dim3(kNmTX, kNmTY), 0, stream()>>>(P);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_maps_pyramid_from_depth(const uint16_t* depth, uint16_t* const* depth_out, kfx_xyz* const* vmap,
                                kfx_xyz* const* nmap, const kfx_intrinsics* intr, int levels, float sigma_color) {
    KFX_REQUIRE_INIT();
    return kfx::maps_pyramid_launch(depth, depth_out, vmap, nmap, intr, levels, sigma_color, nullptr);
}

int kfx_resize_pyramid(const kfx_xyz* vmap0, const kfx_xyz* nmap0, int cols, int rows, kfx_xyz* vmap1,
                       kfx_xyz* nmap1, kfx_xyz* vmap2, kfx_xyz* nmap2) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(vmap0 && nmap0 && vmap1 && nmap1 && vmap2 && nmap2, "kfx_resize_pyramid: null pointer");
    KFX_REQUIRE(cols >= 4 && rows >= 4 && cols % 4 == 0 && rows % 4 == 0,
                "kfx_resize_pyramid: level 0 must be at least 4 x 4 and a multiple of 4");
    const int c1 = cols / 2, r1 = rows / 2, c2 = cols / 4, r2 = rows / 4;
    const int n = c1 * r1 + c2 * r2;
    ProfScope prof("resize_pyramid");
    resize_pyramid_kernel<<<(unsigned)divup(n, 256), 256, 0, stream()>>>(vmap0, nmap0, cols, vmap1, nmap1, c1, r1, vmap2,
                                                                          nmap2, c2, r2);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

}  // extern "C"
