// kfx_icp.hip — projective point-to-plane ICP normal equations.
//
// Two paths share the per-pixel correspondence and Jacobian row:
//
// 1. The reference's three-stage path (Estimate_Combined,
//    common/src/opencl_wrappers.cpp:364-438): estimate_combined writes one
//    single-precision partial sum per 32x8 pixel block for each of the 27
//    entries of the upper triangle of [A | b] into gbuf; transform_estimation
//    (GPU/opencl/transform_estimation.cl:4-53) reduces each entry with 256
//    lanes; the host downloads the 27 floats and unpacks A and b.
//
// 2. The fused path (kfx_icp_iterate): two launches per ICP iteration and
//    no host round trip.  The single-precision products of each pixel are
//    summed in double precision per thread, then over the wave and the
//    workgroup in a fixed order; a one-workgroup launch adds the workgroup
//    totals in double precision in a fixed order, solves the 6x6 system in
//    double precision over one wave and composes the increment into the
//    device-resident pose.  No host round trip between iterations, so a whole frame's 19
//    iterations can be enqueued (or captured in a graph) at once, and the
//    result does not depend on the order in which workgroups finish.
#include <math.h>

#include <type_traits>

#include <stdio.h>

#include "kfx_common.h"

namespace kfx {

struct IcpFrame {
    float rprev_inv[9];
    float tprev[3];
    float fx, fy, cx, cy;
    int cols, rows;
    float dist_thr, angle_thr;
    int prev_from_state;     // take rprev_inv / tprev from the state (set by begin_frame)
};

// Device-resident ICP state of the fused path.
struct IcpState {
    double R[9], t[3];       // current estimate, camera -> volume
    float Rf[9], tf[3];      // single-precision copy read by the Jacobian
    double A[36], b[6];      // last normal equations
    double x[6];             // last increment
    float Rpi[9], tp[3];     // previous frame's pose: R^T and t (begin_frame)
    int status;              // KFX_OK or KFX_ERR_SINGULAR (last solve)
    int singular_count;
    int converged_tag;       // tag of the level that has converged (-1 none)
    int iterations;          // solves performed
    int fault;               // a barrier wait of the one-launch frame ran out: the state is not valid
    int pad;
};

// Projective association of one pixel (current frame, camera -> volume by
// Rc, tc; previous frame by rprev_inv, tprev): g = Rc v + tc, projected into
// the previous frame and rounded to the nearest pixel j.  False when the
// current vertex or normal is missing, the point is behind the previous
// camera or it projects outside the image.
__device__ __forceinline__ bool icp_project(const kfx_xyz& v, const kfx_xyz& n, const float* Rc, const float* tc,
                                            const IcpFrame& f, float g[3], int& j) {
    if (isnan(n.x) || isnan(v.x)) return false;
    const float gx = Rc[0] * v.x + Rc[1] * v.y + Rc[2] * v.z + tc[0];
    const float gy = Rc[3] * v.x + Rc[4] * v.y + Rc[5] * v.z + tc[1];
    const float gz = Rc[6] * v.x + Rc[7] * v.y + Rc[8] * v.z + tc[2];
    g[0] = gx; g[1] = gy; g[2] = gz;
    const float dx = gx - f.tprev[0], dy = gy - f.tprev[1], dz = gz - f.tprev[2];
    const float px = f.rprev_inv[0] * dx + f.rprev_inv[1] * dy + f.rprev_inv[2] * dz;
    const float py = f.rprev_inv[3] * dx + f.rprev_inv[4] * dy + f.rprev_inv[5] * dz;
    const float pz = f.rprev_inv[6] * dx + f.rprev_inv[7] * dy + f.rprev_inv[8] * dz;
    // In front: pz in [2^-126, 2^126), where rcp_rn is the correctly rounded
    // reciprocal; each quotient (p f) / pz is then formed without a division
    // (Markstein's correction with the exact FMA residual gives the correctly
    // rounded quotient from RN(1/pz)).
    if (!(pz >= kRcpMin && pz < kRcpMax)) return false;
    const float y = rcp_rn(pz);
    const float au = px * f.fx, av = py * f.fy;
    const float qu0 = au * y, qv0 = av * y;
    const float fu = rintf(fmaf(fmaf(-qu0, pz, au), y, qu0) + f.cx);
    const float fv = rintf(fmaf(fmaf(-qv0, pz, av), y, qv0) + f.cy);
    if (!(fu >= 0.0f && fv >= 0.0f && fu < (float)f.cols && fv < (float)f.rows)) return false;
    j = (int)fv * f.cols + (int)fu;
    return true;
}

// Jacobian row and residual of an associated pair: zero unless the previous
// vertex and normal exist, the points lie within dist_thr and the normals'
// cross product is shorter than angle_thr (sine of the angle).  row[0..2] =
// g x n_prev, row[3..5] = n_prev, row[6] = n_prev . (v_prev - g).
__device__ __forceinline__ void icp_residual(const kfx_xyz& n, const float g[3], const kfx_xyz& np,
                                             const kfx_xyz& vp, const float* Rc, const IcpFrame& f, bool ok,
                                             float row[7]) {
    for (int k = 0; k < 7; ++k) row[k] = 0.0f;
    if (!ok || isnan(np.x) || isnan(vp.x)) return;
    const float gx = g[0], gy = g[1], gz = g[2];
    const float ex = vp.x - gx, ey = vp.y - gy, ez = vp.z - gz;
    // sqrt_rn: the correctly rounded root (sqrtf's value) without the
    // library's range scaling.
    if (!(sqrt_rn(ex * ex + ey * ey + ez * ez) <= f.dist_thr)) return;
    const float cnx = Rc[0] * n.x + Rc[1] * n.y + Rc[2] * n.z;
    const float cny = Rc[3] * n.x + Rc[4] * n.y + Rc[5] * n.z;
    const float cnz = Rc[6] * n.x + Rc[7] * n.y + Rc[8] * n.z;
    const float sx = cny * np.z - cnz * np.y, sy = cnz * np.x - cnx * np.z, sz = cnx * np.y - cny * np.x;
    if (!(sqrt_rn(sx * sx + sy * sy + sz * sz) < f.angle_thr)) return;
    row[0] = gy * np.z - gz * np.y;
    row[1] = gz * np.x - gx * np.z;
    row[2] = gx * np.y - gy * np.x;
    row[3] = np.x;
    row[4] = np.y;
    row[5] = np.z;
    row[6] = np.x * ex + np.y * ey + np.z * ez;
}

// The 27 products: row[i] * row[j] for i <= j, j in [i, 6], row[6] = b
// (unpacking order of opencl_wrappers.cpp:426-437).
__device__ __forceinline__ void icp_row(const kfx_xyz* __restrict__ vcurr, const kfx_xyz* __restrict__ ncurr,
                                        const kfx_xyz* __restrict__ vprev, const kfx_xyz* __restrict__ nprev,
                                        const float* Rc, const float* tc, const IcpFrame& f, int x, int y,
                                        float row[7]) {
    const int i = y * f.cols + x;
    const kfx_xyz n = ncurr[i], v = vcurr[i];
    float g[3];
    int j = 0;
    const bool ok = icp_project(v, n, Rc, tc, f, g, j);
    kfx_xyz np{NAN, NAN, NAN}, vp{NAN, NAN, NAN};
    if (ok) { np = nprev[j]; vp = vprev[j]; }
    icp_residual(n, g, np, vp, Rc, f, ok, row);
}

// ------------------------------------------------------------------------
// Reference path.
// ------------------------------------------------------------------------
constexpr int kEcBX = 32, kEcBY = 8;

__global__ void __launch_bounds__(kEcBX * kEcBY)
estimate_combined_kernel(const kfx_xyz* __restrict__ vcurr, const kfx_xyz* __restrict__ ncurr,
                         const kfx_xyz* __restrict__ vprev, const kfx_xyz* __restrict__ nprev, IcpFrame f,
                         kfx_matrix_3f Rc_m, kfx_point_3f tc_p, float* __restrict__ gbuf, int stride) {
    __shared__ float red[kEcBX * kEcBY / 64][27];
    const int x = blockIdx.x * kEcBX + threadIdx.x;
    const int y = blockIdx.y * kEcBY + threadIdx.y;
    const float Rc[9] = {Rc_m.data[0].x, Rc_m.data[0].y, Rc_m.data[0].z, Rc_m.data[1].x, Rc_m.data[1].y,
                         Rc_m.data[1].z, Rc_m.data[2].x, Rc_m.data[2].y, Rc_m.data[2].z};
    const float tc[3] = {tc_p.x, tc_p.y, tc_p.z};
    float row[7] = {0, 0, 0, 0, 0, 0, 0};
    if (x < f.cols && y < f.rows) icp_row(vcurr, ncurr, vprev, nprev, Rc, tc, f, x, y, row);
    const int tid = threadIdx.y * kEcBX + threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j, ++k) {
            float v = row[i] * row[j];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) red[wave][k] = v;
        }
    __syncthreads();
    if (tid < 27) {
        float s = 0.0f;
        for (int w = 0; w < kEcBX * kEcBY / 64; ++w) s += red[w][tid];
        gbuf[tid * stride + blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
}

// transform_estimation.cl:4-53 with the row stride as a parameter (the
// reference hard-codes 1200, correct only for 640x480 level 0).
__global__ void __launch_bounds__(256) transform_estimation_kernel(const float* __restrict__ gbuf,
                                                                   float* __restrict__ out, int buf_length,
                                                                   int stride) {
    __shared__ float smem[256];
    const int tid = threadIdx.x;
    const int start = blockIdx.x * stride;
    const int end = start + buf_length;
    float sum = 0.0f;
    for (int id = start + tid; id < end; id += 256) sum += gbuf[id];
    smem[tid] = sum;
    __syncthreads();
    float val = smem[tid];
    if (tid < 128) smem[tid] = val = val + smem[tid + 128];
    __syncthreads();
    if (tid < 64) smem[tid] = val = val + smem[tid + 64];
    __syncthreads();
    for (int off = 32; off >= 1; off >>= 1) {
        if (tid < 32) smem[tid] = val = val + smem[tid + off];
        __syncthreads();
    }
    if (tid == 0) out[blockIdx.x] = smem[0];
}

__global__ void clean_buff_kernel(float* __restrict__ g, int width) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= width) return;
    for (int i = 0; i < 27; ++i) g[i * width + id] = 0.0f;
}

// ------------------------------------------------------------------------
// Fused path.
// ------------------------------------------------------------------------
constexpr int kIcpThreads = 256;
constexpr int kIcpWaves = kIcpThreads / 64;
constexpr int kIcpEntries = 27;
constexpr unsigned kIcpTwoPixelMin = 150000;  // images at least this large take two pixels per thread

// Sums padded to 32 entries, spread over a wave by a butterfly that halves
// the entries each step: after the step with xor offset 32, 16, 8, 4, 2 a
// lane keeps the half selected by that bit of its lane id, added to its
// partner's copy.  The last step (offset 1) adds the two lanes left per
// entry.  Lane l ends with the wave total of entry l >> 1, in a fixed
// association order (each addition is commutative, so which lane of a pair
// holds which operand does not matter).
template <int kHalf, typename T>
__device__ __forceinline__ void reduce_scatter_step(T (&v)[32], int lane) {
    if constexpr (kHalf >= 8) {
#pragma unroll
        for (int i = 0; i < kHalf; ++i) {
            T x, y;
            swap_halves<kHalf>(v[i], v[i + kHalf], x, y);
            v[i] = x + y;
        }
    } else {
        const bool upper = (lane & (2 * kHalf)) != 0;
#pragma unroll
        for (int i = 0; i < kHalf; ++i) {
            const T send = upper ? v[i] : v[i + kHalf];
            const T keep = upper ? v[i + kHalf] : v[i];
            v[i] = keep + dpp_xor<2 * kHalf>(send);
        }
    }
}

template <typename T>
__device__ __forceinline__ T wave_reduce_scatter32(T (&v)[32], int lane) {
    reduce_scatter_step<16>(v, lane);
    reduce_scatter_step<8>(v, lane);
    reduce_scatter_step<4>(v, lane);
    reduce_scatter_step<2>(v, lane);
    reduce_scatter_step<1>(v, lane);
    return v[0] + dpp_xor<1>(v[0]);
}

// Error-free sum (Knuth TwoSum): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated (Kahan-style) accumulation of v into (s, c).
__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    float e;
    two_sum(s, v, s, e);
    c += e;
}

// Merge of two compensated sums (s, c) += (s2, c2).  The TwoSum error term is
// exact whichever operand comes first, so the merge is commutative.
__device__ __forceinline__ void pair_merge(float& s, float& c, float s2, float c2) {
    float e;
    two_sum(s, s2, s, e);
    c = c + c2 + e;
}

// The butterfly of wave_reduce_scatter32 on compensated single-precision
// pairs: lane l ends with the wave total (s, c) of entry l >> 1.
template <int kHalf>
__device__ __forceinline__ void reduce_scatter_step(float (&s)[32], float (&c)[32], int lane) {
    if constexpr (kHalf >= 8) {
#pragma unroll
        for (int i = 0; i < kHalf; ++i) {
            float xs, ys, xc, yc;
            swap_halves<kHalf>(s[i], s[i + kHalf], xs, ys);
            swap_halves<kHalf>(c[i], c[i + kHalf], xc, yc);
            pair_merge(xs, xc, ys, yc);
            s[i] = xs;
            c[i] = xc;
        }
    } else {
        const bool upper = (lane & (2 * kHalf)) != 0;
#pragma unroll
        for (int i = 0; i < kHalf; ++i) {
            const float send_s = upper ? s[i] : s[i + kHalf];
            const float send_c = upper ? c[i] : c[i + kHalf];
            float keep_s = upper ? s[i + kHalf] : s[i];
            float keep_c = upper ? c[i + kHalf] : c[i];
            pair_merge(keep_s, keep_c, dpp_xor<2 * kHalf>(send_s), dpp_xor<2 * kHalf>(send_c));
            s[i] = keep_s;
            c[i] = keep_c;
        }
    }
}

__device__ __forceinline__ float2 wave_reduce_scatter32(float (&s)[32], float (&c)[32], int lane) {
    reduce_scatter_step<16>(s, c, lane);
    reduce_scatter_step<8>(s, c, lane);
    reduce_scatter_step<4>(s, c, lane);
    reduce_scatter_step<2>(s, c, lane);
    reduce_scatter_step<1>(s, c, lane);
    float a = s[0], e = c[0];
    pair_merge(a, e, dpp_xor<1>(s[0]), dpp_xor<1>(c[0]));
    return make_float2(a, e);
}

// sin and cos of an ICP increment angle with |x| <= 1/8: the Taylor series
// (sin to x^9, cos to x^10; the first omitted terms are below 3e-17 relative)
// in Horner form with fused multiply-adds.
__device__ __forceinline__ void sincos_series(double x, double& s, double& c) {
    const double z = x * x;
    const double ps = fma(z, fma(z, fma(z, 1.0 / 362880.0, -1.0 / 5040.0), 1.0 / 120.0), -1.0 / 6.0);
    s = fma(x * z, ps, x);
    const double pc = fma(z, fma(z, fma(z, fma(z, -1.0 / 3628800.0, 1.0 / 40320.0), -1.0 / 720.0), 1.0 / 24.0),
                          -0.5);
    c = fma(z, pc, 1.0);
}

// Increment angles beyond the series' range (rare: an ICP step of more than
// 7 degrees) are brought into it and the series' results carried back, with
// the series' own constants so the kernels that inline the solve need few
// extra registers for it.  reduce_angle: past |x| = 3, r = x - k 2pi
// (k = rint(x / 2pi), 2pi split into two doubles, each product subtracted
// with one rounding); then r halved m <= 5 times to |r| <= 1/8 (a
// non-finite x gives r = NaN).  double_angle: m steps (s, c) <- (2 s c,
// (c - s)(c + s)), each adding a few ulps.
__device__ __forceinline__ void reduce_angle(double& r, int& m) {
    m = 0;
    if (!isfinite(r)) {
        r = r - r;
        return;
    }
    if (fabs(r) > 3.0) {
        constexpr double kInvTwoPi = 0x1.45f306dc9c883p-3;
        constexpr double kTwoPi1 = 0x1.921fb54442d18p+2, kTwoPi2 = 0x1.1a62633145c07p-52;
        const double k = rint(r * kInvTwoPi);
        r = fma(-k, kTwoPi1, r);
        r = fma(-k, kTwoPi2, r);
    }
    while (fabs(r) > 0.125 && m < 6) {
        r *= 0.5;
        ++m;
    }
}

__device__ __forceinline__ void double_angle(double& s, double& c, int m) {
    for (int i = 0; i < m; ++i) {
        const double s2 = 2.0 * s * c;
        c = (c - s) * (c + s);
        s = s2;
    }
}

// Lane l's value of v, in every lane (l a constant; every lane active).
__device__ __forceinline__ double lane_value(double v, int l) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)u, l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Index of entry (r, c), r <= c <= 6, among the 27 sums (row r of the upper
// triangle of [A | b] after rows 0 .. r-1).
__host__ __device__ constexpr int sum_index(int r, int c) { return 7 * r - r * (r - 1) / 2 + (c - r); }

// The normal equations A x = b of the 27 sums solved in double precision,
// then the pose update of the estimate R0, t0 (st->R, st->t: the state's
// own, or a copy the caller loaded ahead; read after the elimination).
// Called by every lane of one wave (lane 0 .. 63).
//
// Gauss-Jordan elimination on [A | b] without pivoting (A is symmetric
// positive definite; its pivots are those of the LDL^T factorisation):
// lane j <= 6 holds column j of [A | b] (lanes 7 .. 63 a copy of column 6);
// step k broadcasts column k (readlane), every lane scales its row-k entry
// by 1 / A_kk (reciprocal with two Newton steps) and eliminates entry k from
// the other rows with fused multiply-adds.  A pivot that is not finite and
// above 1e-300 makes the system singular: status KFX_ERR_SINGULAR, x = 0,
// the pose unchanged.  After six steps column 6 holds x.  Each step's chain
// is one broadcast, the reciprocal and two dependent operations, where a
// one-lane Cholesky factorisation and two substitutions run several hundred
// dependent double-precision operations.
__device__ void solve_and_update(IcpState* st, const double* sums, const double* R0, const double* t0,
                                 int iterations, int tag, double conv_rot, double conv_trans, int lane) {
    if (lane < kIcpEntries) {  // the normal equations of the state
        int r = 0;
#pragma unroll
        for (int q = 1; q < 6; ++q) r += lane >= sum_index(q, q) ? 1 : 0;
        const int c = lane - sum_index(r, r) + r;
        const double v = sums[lane];
        if (c == 6) {
            st->b[r] = v;
        } else {
            st->A[r * 6 + c] = v;
            st->A[c * 6 + r] = v;
        }
    }
    const int j = lane < 6 ? lane : 6;
    double col[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) col[i] = sums[i <= j ? sum_index(i, j) : sum_index(j, i)];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double ck[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) ck[i] = lane_value(col[i], k);
        const double d = ck[k];
        ok = ok && (d > 1e-300) && isfinite(d);
        const double dv = ok ? d : 1.0;
        double y = __builtin_amdgcn_rcp(dv);
#pragma unroll
        for (int nr = 0; nr < 2; ++nr) y = fma(y, fma(-dv, y, 1.0), y);
        const double pk = col[k] * y;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            if (i != k) col[i] = fma(-ck[i], pk, col[i]);
        col[k] = pk;
    }
    if (lane == 0) st->iterations = iterations + 1;
    if (!ok) {  // uniform
        if (lane == 0) {
            st->status = KFX_ERR_SINGULAR;
            st->singular_count += 1;
#pragma unroll
            for (int i = 0; i < 6; ++i) st->x[i] = 0.0;
        }
        return;
    }
    double x[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) x[i] = lane_value(col[i], 6);
    // Increment: R_inc = Rz(x2) Ry(x1) Rx(x0), t_inc = x[3..5];
    // t <- R_inc t + t_inc, R <- R_inc R.  Lane 3 r + c < 9 forms R[r][c],
    // lane 9 + r forms t[r] (each with the same operations in the same order
    // as one lane forming all twelve).
    double r0 = x[0], r1 = x[1], r2 = x[2];
    int m0 = 0, m1 = 0, m2 = 0;
    const bool wide = !(fmax(fabs(x[0]), fmax(fabs(x[1]), fabs(x[2]))) <= 0.125);  // uniform
    if (wide) {
        reduce_angle(r0, m0);
        reduce_angle(r1, m1);
        reduce_angle(r2, m2);
    }
    double sa, ca, sb, cb, sg, cg;  // the three series interleave
    sincos_series(r0, sa, ca);
    sincos_series(r1, sb, cb);
    sincos_series(r2, sg, cg);
    if (wide) {
        double_angle(sa, ca, m0);
        double_angle(sb, cb, m1);
        double_angle(sg, cg, m2);
    }
    const bool is_t = lane >= 9;
    const int r = is_t ? lane - 9 : lane / 3;  // row of R_inc (lanes >= 12: unused)
    const int c = lane - 3 * (lane / 3);        // column of R0 (R lanes)
    double ri0, ri1, ri2;
    if (r == 0) {
        ri0 = cg * cb; ri1 = cg * sb * sa - sg * ca; ri2 = cg * sb * ca + sg * sa;
    } else if (r == 1) {
        ri0 = sg * cb; ri1 = sg * sb * sa + cg * ca; ri2 = sg * sb * ca - cg * sa;
    } else {
        ri0 = -sb; ri1 = cb * sa; ri2 = cb * ca;
    }
    double o0, o1, o2;
    if (is_t) {
        o0 = t0[0]; o1 = t0[1]; o2 = t0[2];
    } else if (c == 0) {
        o0 = R0[0]; o1 = R0[3]; o2 = R0[6];
    } else if (c == 1) {
        o0 = R0[1]; o1 = R0[4]; o2 = R0[7];
    } else {
        o0 = R0[2]; o1 = R0[5]; o2 = R0[8];
    }
    double v = ri0 * o0 + ri1 * o1 + ri2 * o2;
    if (is_t) v = v + (r == 0 ? x[3] : r == 1 ? x[4] : x[5]);
    if (lane < 9) {
        st->R[lane] = v;
        st->Rf[lane] = (float)v;
    } else if (lane < 12) {
        st->t[r] = v;
        st->tf[r] = (float)v;
    }
    if (lane != 0) return;
#pragma unroll
    for (int i = 0; i < 6; ++i) st->x[i] = x[i];
    st->status = KFX_OK;
    if (conv_rot > 0.0) {
        const double rot = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        const double tr = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
        if (rot < conv_rot && tr < conv_trans) st->converged_tag = tag;
    }
}

// Precision of the per-workgroup sums (the cross-workgroup sum, the solve and
// the pose update are double precision in every mode).
enum IcpSumMode : int {
    kSumKahan = KFX_ICP_SUM_FP32_KAHAN,  // FP32 products, compensated FP32 sums
    kSumPlain = KFX_ICP_SUM_FP32,        // FP32 products, plain FP32 sums
    kSumDouble = KFX_ICP_SUM_FP64,       // FP32 products summed in FP64
};

// Sum of 32 per-thread values over a group of kIcpThreads threads (local
// ids lt = 0 .. kIcpThreads - 1; the group is the workgroup, or a quarter of
// a larger one) in a fixed order: the butterfly over each wave, then the
// four wave totals in wave order.  Local thread e < 27 returns entry e (the
// others return 0).  Every thread of the workgroup calls it (it holds a
// workgroup barrier); wsum is the group's own.  A wave with live == false
// (uniform: every value of every lane is +0) skips its butterfly, whose
// totals would all be +0.
template <int kWaves = kIcpWaves, typename T>
__device__ __forceinline__ T group_sum27(T (&acc)[32], T (*wsum)[32], int lt, bool live = true) {
    const int lane = lt & 63, wave = lt >> 6;
    T wtot = 0;
    if (live) wtot = wave_reduce_scatter32(acc, lane);
    if ((lane & 1) == 0) wsum[wave][lane >> 1] = wtot;
    __syncthreads();
    T s = 0;
    if (lt < kIcpEntries) {
        s = wsum[0][lt];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += wsum[w][lt];
    }
    return s;
}

template <int kWaves = kIcpWaves, typename T>
__device__ __forceinline__ T workgroup_sum27(T (&acc)[32], T (*wsum)[32]) {
    return group_sum27<kWaves>(acc, wsum, threadIdx.x);
}

// The same for compensated single-precision pairs.
__device__ __forceinline__ float2 group_sum27(float (&s)[32], float (&c)[32], float2 (*wsum)[32], int lt,
                                             bool live = true) {
    const int lane = lt & 63, wave = lt >> 6;
    float2 wtot = make_float2(0.0f, 0.0f);
    if (live) wtot = wave_reduce_scatter32(s, c, lane);
    if ((lane & 1) == 0) wsum[wave][lane >> 1] = wtot;
    __syncthreads();
    float a = 0.0f, e = 0.0f;
    if (lt < kIcpEntries) {
        a = wsum[0][lt].x;
        e = wsum[0][lt].y;
#pragma unroll
        for (int w = 1; w < kIcpWaves; ++w) pair_merge(a, e, wsum[w][lt].x, wsum[w][lt].y);
    }
    return make_float2(a, e);
}

// Current-frame loads of block blk's P pixels per thread (they do not
// depend on the pose): local thread lt takes pixels blk * kIcpThreads * P +
// q * kIcpThreads + lt (q < P).
template <int P>
__device__ __forceinline__ void icp_load_current(const kfx_xyz* __restrict__ vcurr,
                                                 const kfx_xyz* __restrict__ ncurr, int npix, int blk, int lt,
                                                 kfx_xyz (&v)[P], kfx_xyz (&n)[P]) {
    const int base = blk * (kIcpThreads * P) + lt;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int p = base + q * kIcpThreads;
        const int ps = p < npix ? p : 0;
        n[q] = ncurr[ps];
        v[q] = vcurr[ps];
    }
}

// Partial stores for a reader on another XCD within the same launch: a
// relaxed atomic store at agent scope goes past the XCD's L2 to the device's
// coherence point (no write-back of the L2 needed).
__device__ __forceinline__ void store_coherent(void* dst, uint64_t bits) {
    __hip_atomic_store(static_cast<unsigned long long*>(dst), (unsigned long long)bits, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t load_coherent(const void* src) {
    return (uint64_t)__hip_atomic_load(static_cast<const unsigned long long*>(src), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
}

// (sum, compensation) of an entry in the FP32 modes (compensation 0 in the
// plain mode), its sum in the FP64 mode.
template <int M>
using IcpGroupSum = std::conditional_t<M == kSumDouble, double, float2>;

// The 27 sums of block blk's pixels by a group of kIcpThreads threads (local
// id lt) at the pose (Rc, tc): the projections of all P pixels, then the
// previous-frame gathers of all P before any residual, so a wave waits on
// memory once more.  The single-precision products of a pixel (icp_row's
// arithmetic) are accumulated per thread (mode M), then over the group in a
// fixed order; local thread e < 27 returns entry e.
template <int P, int M>
__device__ __forceinline__ IcpGroupSum<M> icp_group_sums(const kfx_xyz (&v)[P], const kfx_xyz (&n)[P],
                                                         const kfx_xyz* __restrict__ vprev,
                                                         const kfx_xyz* __restrict__ nprev, const IcpFrame& f,
                                                         const float* Rc, const float* tc, int blk, int lt,
                                                         double (*wsum_d)[32], float2 (*wsum_f)[32]) {
    const int npix = f.cols * f.rows;
    const int base = blk * (kIcpThreads * P) + lt;
    kfx_xyz vp[P], np[P];
    float g[P][3];
    bool ok[P];
    int j[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int p = base + q * kIcpThreads;
        ok[q] = (p < npix) && icp_project(v[q], n[q], Rc, tc, f, g[q], j[q]);
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int js = ok[q] ? j[q] : 0;
        np[q] = nprev[js];
        vp[q] = vprev[js];
    }
    float rows[P][7];
#pragma unroll
    for (int q = 0; q < P; ++q) icp_residual(n[q], g[q], np[q], vp[q], Rc, f, ok[q], rows[q]);
    // A wave none of whose pixels is associated sums only +0 (the
    // accumulation below starts with p + 0).
    bool any_ok = false;
#pragma unroll
    for (int q = 0; q < P; ++q) any_ok = any_ok || ok[q];
    const bool live = __any(any_ok);
    // Accumulation from zero; the first pixel's step is written out: 0 + p
    // is p + 0 (which turns -0 into +0), and TwoSum(0, p) has a zero error.
    if constexpr (M == kSumDouble) {
        double acc[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) acc[k] = 0.0;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            int k = 0;
#pragma unroll
            for (int i = 0; i < 6; ++i)
#pragma unroll
                for (int jj = i; jj < 7; ++jj, ++k) {
                    const double prod = (double)(rows[q][i] * rows[q][jj]);
                    acc[k] = q == 0 ? prod + 0.0 : acc[k] + prod;
                }
        }
        return group_sum27(acc, wsum_d, lt, live);
    } else {
        float sum[32], comp[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) { sum[k] = 0.0f; comp[k] = 0.0f; }
#pragma unroll
        for (int q = 0; q < P; ++q) {
            int k = 0;
#pragma unroll
            for (int i = 0; i < 6; ++i)
#pragma unroll
                for (int jj = i; jj < 7; ++jj, ++k) {
                    const float prod = rows[q][i] * rows[q][jj];
                    if (q == 0) sum[k] = prod + 0.0f;
                    else if constexpr (M == kSumKahan) kahan_add(sum[k], comp[k], prod);
                    else sum[k] += prod;
                }
        }
        if constexpr (M == kSumKahan) return group_sum27(sum, comp, wsum_f, lt, live);
        else return make_float2(group_sum27(sum, reinterpret_cast<float (*)[32]>(wsum_f), lt, live), 0.0f);
    }
}

// The workgroup's sums (icp_group_sums of block blockIdx.x) stored as the
// partials: [27][workgroup] float2 (sum, compensation) in the FP32 modes,
// double in the FP64 mode; entry-major, so the cross-workgroup sum's loads
// of one entry over consecutive workgroups coalesce.
template <int P, int M, bool kCoherent = false>
__device__ __forceinline__ void icp_block_sums(const kfx_xyz (&v)[P], const kfx_xyz (&n)[P],
                                               const kfx_xyz* __restrict__ vprev, const kfx_xyz* __restrict__ nprev,
                                               const IcpFrame& f, const float* Rc, const float* tc,
                                               void* __restrict__ partials, double (*wsum_d)[32],
                                               float2 (*wsum_f)[32]) {
    const IcpGroupSum<M> r =
        icp_group_sums<P, M>(v, n, vprev, nprev, f, Rc, tc, blockIdx.x, threadIdx.x, wsum_d, wsum_f);
    if (threadIdx.x < kIcpEntries) {
        if constexpr (M == kSumDouble) {
            double* dst = static_cast<double*>(partials) + threadIdx.x * gridDim.x + blockIdx.x;
            if constexpr (kCoherent) store_coherent(dst, __double_as_longlong(r));
            else *dst = r;
        } else {
            float2* dst = static_cast<float2*>(partials) + threadIdx.x * gridDim.x + blockIdx.x;
            if constexpr (kCoherent)
                store_coherent(dst, ((uint64_t)__float_as_uint(r.y) << 32) | __float_as_uint(r.x));
            else *dst = r;
        }
    }
}

__device__ __forceinline__ void icp_prev_from_state(IcpFrame& f, const IcpState* st) {
    if (f.prev_from_state) {
        for (int i = 0; i < 9; ++i) f.rprev_inv[i] = st->Rpi[i];
        for (int i = 0; i < 3; ++i) f.tprev[i] = st->tp[i];
    }
}

// One ICP iteration, first half: the 27 sums of each workgroup at the
// state's pose.  Workgroups never wait for each other: the cross-workgroup
// sum is the next launch (icp_solve_kernel), so no workgroup needs a
// device-wide fence or a shared counter.
template <int P, int M>
__global__ void __launch_bounds__(kIcpThreads)
icp_sums_kernel(const kfx_xyz* __restrict__ vcurr, const kfx_xyz* __restrict__ ncurr,
                const kfx_xyz* __restrict__ vprev, const kfx_xyz* __restrict__ nprev, IcpFrame f,
                const IcpState* __restrict__ st, void* __restrict__ partials, int tag) {
    __shared__ double wsum_d[kIcpWaves][32];
    __shared__ float2 wsum_f[kIcpWaves][32];
    // The current-frame loads do not depend on the state: they are issued
    // before the convergence test and the pose reads.
    kfx_xyz v[P], n[P];
    icp_load_current<P>(vcurr, ncurr, f.cols * f.rows, blockIdx.x, threadIdx.x, v, n);
    if (tag >= 0 && st->converged_tag == tag) return;  // uniform across the grid
    icp_prev_from_state(f, st);
    float Rc[9], tc[3];
    for (int i = 0; i < 9; ++i) Rc[i] = st->Rf[i];
    for (int i = 0; i < 3; ++i) tc[i] = st->tf[i];
    icp_block_sums<P, M>(v, n, vprev, nprev, f, Rc, tc, partials, wsum_d, wsum_f);
}

// The cross-workgroup sum of nblk workgroups' partials into total[0..26]
// (LDS), by a workgroup of kIcpThreads threads: thread t adds the totals of
// workgroups t, t + kIcpThreads, ... in double precision in that order (up
// to kRound workgroups' loads issued together per round), then the
// fixed-order workgroup sum.  The result depends on nblk and the partials
// only, not on kRound or on the order in which the workgroups ran.
template <bool kCoherent>
__device__ __forceinline__ float2 partial_f2(const float2* p) {
    if constexpr (kCoherent) {
        const uint64_t b = load_coherent(p);
        return make_float2(__uint_as_float((uint32_t)b), __uint_as_float((uint32_t)(b >> 32)));
    } else {
        return *p;
    }
}
template <bool kCoherent>
__device__ __forceinline__ double partial_d(const double* p) {
    if constexpr (kCoherent) return __longlong_as_double((long long)load_coherent(p));
    else return *p;
}

template <int M, int kRound, bool kCoherent = false>
__device__ __forceinline__ void cross_block_sum(const void* __restrict__ partials, int nblk, double (*wsum)[32],
                                                double* total) {
    double acc[32];
#pragma unroll
    for (int e = 0; e < 32; ++e) acc[e] = 0.0;
    int b = threadIdx.x;
    if constexpr (M != kSumDouble && kRound >= 3) {
        const float2* pf = static_cast<const float2*>(partials);
        for (; b + 2 * kIcpThreads < nblk; b += 3 * kIcpThreads) {
            float2 t0[kIcpEntries], t1[kIcpEntries], t2[kIcpEntries];
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) t0[e] = partial_f2<kCoherent>(&pf[e * nblk + b]);
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) t1[e] = partial_f2<kCoherent>(&pf[e * nblk + b + kIcpThreads]);
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) t2[e] = partial_f2<kCoherent>(&pf[e * nblk + b + 2 * kIcpThreads]);
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) acc[e] += (double)t0[e].x + (double)t0[e].y;
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) acc[e] += (double)t1[e].x + (double)t1[e].y;
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) acc[e] += (double)t2[e].x + (double)t2[e].y;
        }
    }
    if constexpr (M != kSumDouble && kRound >= 2) {
        const float2* pf = static_cast<const float2*>(partials);
        for (; b + kIcpThreads < nblk; b += 2 * kIcpThreads) {
            float2 t0[kIcpEntries], t1[kIcpEntries];
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) t0[e] = partial_f2<kCoherent>(&pf[e * nblk + b]);
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) t1[e] = partial_f2<kCoherent>(&pf[e * nblk + b + kIcpThreads]);
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) acc[e] += (double)t0[e].x + (double)t0[e].y;
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) acc[e] += (double)t1[e].x + (double)t1[e].y;
        }
    }
    for (; b < nblk; b += kIcpThreads) {
        if constexpr (M == kSumDouble) {
            const double* pd = static_cast<const double*>(partials) + b;
            double t[kIcpEntries];
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) t[e] = partial_d<kCoherent>(&pd[e * nblk]);
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) acc[e] += t[e];
        } else {
            const float2* pf = static_cast<const float2*>(partials) + b;
            float2 t[kIcpEntries];
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) t[e] = partial_f2<kCoherent>(&pf[e * nblk]);
#pragma unroll
            for (int e = 0; e < kIcpEntries; ++e) acc[e] += (double)t[e].x + (double)t[e].y;
        }
    }
    const double s = workgroup_sum27<kIcpWaves>(acc, wsum);
    if (threadIdx.x < kIcpEntries) total[threadIdx.x] = s;
    __syncthreads();
}

// One ICP iteration, second half (one workgroup): the cross-workgroup sum
// (three workgroups' loads per round: all of them in one round at 640x480),
// then the first wave's solve and pose update.  The estimate is loaded
// before the sum so that its latency hides behind the partials' loads.
// Thread 0 copies the state to host memory (8-byte stores).
__device__ __forceinline__ void icp_publish(const IcpState* st, IcpState* mirror) {
    static_assert(sizeof(IcpState) % 8 == 0, "state copy by 8-byte words");
    const uint2* src = reinterpret_cast<const uint2*>(st);
    uint2* dst = reinterpret_cast<uint2*>(mirror);
    for (int i = 0; i < (int)(sizeof(IcpState) / 8); ++i) dst[i] = src[i];
}

// With mirror (pinned host memory), thread 0 also publishes the resulting
// state there, so the host reads it after the stream without a copy command.
template <int M>
__global__ void __launch_bounds__(kIcpThreads)
icp_solve_kernel(IcpState* __restrict__ st, const void* __restrict__ partials, int nblk, int tag,
                 double conv_rot, double conv_trans, IcpState* __restrict__ mirror) {
    __shared__ double wsum[kIcpWaves][32];
    __shared__ double total[32];
    if (tag >= 0 && st->converged_tag == tag) {
        if (mirror != nullptr && threadIdx.x == 0) icp_publish(st, mirror);
        return;
    }
    double R0[9], t0[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) R0[i] = st->R[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) t0[i] = st->t[i];
    const int iterations = st->iterations;
    cross_block_sum<M, 3>(partials, nblk, wsum, total);
    if (threadIdx.x < 64) solve_and_update(st, total, R0, t0, iterations, tag, conv_rot, conv_trans, threadIdx.x);
    if (mirror == nullptr) return;  // uniform
    __syncthreads();  // the state's stores by the first wave's lanes are visible
    if (threadIdx.x == 0) icp_publish(st, mirror);
}

// cross_block_sum over coherent loads, in a register budget that leaves the
// one-launch iteration four waves per SIMD: thread t adds the partials of
// workgroups t, t + kIcpThreads, t + 2 kIcpThreads, ... per entry, nine
// entries at a time (their loads issued together), in double precision and
// in that order, then the fixed-order workgroup sum.  The same totals as
// cross_block_sum: each entry's per-thread sum runs over the same
// workgroups in the same order.
template <int M>
__device__ __forceinline__ void cross_block_sum_coherent(const void* __restrict__ partials, int nblk,
                                                         double (*wsum)[32], double* total) {
    constexpr int kGroup = 9;
    static_assert(kIcpEntries % kGroup == 0, "entry groups");
    double acc[32];
#pragma unroll
    for (int e = 0; e < 32; ++e) acc[e] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kIcpThreads) {
#pragma unroll
        for (int g = 0; g < kIcpEntries; g += kGroup) {
            double t[kGroup];
#pragma unroll
            for (int e = 0; e < kGroup; ++e) {
                if constexpr (M == kSumDouble) {
                    t[e] = partial_d<true>(static_cast<const double*>(partials) + (size_t)(g + e) * nblk + b);
                } else {
                    const float2 v = partial_f2<true>(static_cast<const float2*>(partials) + (size_t)(g + e) * nblk + b);
                    t[e] = (double)v.x + (double)v.y;
                }
            }
#pragma unroll
            for (int e = 0; e < kGroup; ++e) acc[g + e] += t[e];
        }
    }
    const double s = workgroup_sum27<kIcpWaves>(acc, wsum);
    if (threadIdx.x < kIcpEntries) total[threadIdx.x] = s;
    __syncthreads();
}

// One ICP iteration in one launch: the 27 sums of each workgroup at the
// state's pose, then the cross-workgroup sum, the solve and the pose update
// by the workgroup that finishes last.  The partials go to the device's
// coherence point (agent-scope atomic stores, no L2 write-back), each
// workgroup waits until its stores have completed before taking a ticket
// (one agent-scope atomic add), and the workgroup with the last ticket reads
// the partials back with agent-scope atomic loads.  The sum is in workgroup
// order whichever workgroup finishes last, so the state equals the
// two-launch form's bit for bit.  The last workgroup resets the ticket for
// the next launch, which the kernel boundary orders.
constexpr int kWaitVmemZero = 0x0F70;  // s_waitcnt vmcnt(0): every load and store of the wave has completed

template <int P, int M>
__global__ void __launch_bounds__(kIcpThreads) __attribute__((amdgpu_waves_per_eu(4)))
icp_iter_kernel(const kfx_xyz* __restrict__ vcurr, const kfx_xyz* __restrict__ ncurr,
                const kfx_xyz* __restrict__ vprev, const kfx_xyz* __restrict__ nprev, IcpFrame f,
                IcpState* __restrict__ st, void* __restrict__ partials, unsigned* __restrict__ ticket, int tag,
                double conv_rot, double conv_trans, IcpState* __restrict__ mirror) {
    __shared__ double wsum_d[kIcpWaves][32];
    __shared__ float2 wsum_f[kIcpWaves][32];
    __shared__ double total[32];
    __shared__ unsigned my_ticket;
    kfx_xyz v[P], n[P];
    icp_load_current<P>(vcurr, ncurr, f.cols * f.rows, blockIdx.x, threadIdx.x, v, n);
    if (tag >= 0 && st->converged_tag == tag) {  // uniform across the grid
        if (mirror != nullptr && blockIdx.x == 0 && threadIdx.x == 0) icp_publish(st, mirror);
        return;
    }
    icp_prev_from_state(f, st);
    float Rc[9], tc[3];
    for (int i = 0; i < 9; ++i) Rc[i] = st->Rf[i];
    for (int i = 0; i < 3; ++i) tc[i] = st->tf[i];
    icp_block_sums<P, M, true>(v, n, vprev, nprev, f, Rc, tc, partials, wsum_d, wsum_f);
    __builtin_amdgcn_s_waitcnt(kWaitVmemZero);
    __syncthreads();
    if (threadIdx.x == 0)
        my_ticket = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (my_ticket != gridDim.x - 1) return;  // uniform across the workgroup
    if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double R0[9], t0[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) R0[i] = st->R[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) t0[i] = st->t[i];
    const int iterations = st->iterations;
    cross_block_sum_coherent<M>(partials, gridDim.x, wsum_d, total);
    if (threadIdx.x < 64) solve_and_update(st, total, R0, t0, iterations, tag, conv_rot, conv_trans, threadIdx.x);
    if (mirror == nullptr) return;  // uniform
    __syncthreads();  // the state's stores by the first wave's lanes are visible
    if (threadIdx.x == 0) icp_publish(st, mirror);
}

// ------------------------------------------------------------------------
// A frame's ICP in one launch.
//
// kFrameBlocks workgroups, all resident at once (the host checks the
// occupancy first), run every iteration of every level.  Workgroup w of
// kFrameGroups groups of kIcpThreads threads computes the sums of the
// two-launch form's workgroups (blocks) w, w + kFrameBlocks, ... (its groups
// take them in parallel, in rounds) and adds their totals in double
// precision in that order: exactly what thread w of icp_solve_kernel's
// cross-workgroup sum adds.  The cross-workgroup sum then runs in two
// stages over agent-scope stores and loads (they pass the XCDs' L2):
//   1. every workgroup stores its 27 values; the first wave of workgroup q
//      (q < 4), lane l holding workgroup 64 q + l's values, runs the
//      butterfly of wave q of icp_solve_kernel's workgroup sum and stores
//      the 27 wave totals;
//   2. every workgroup reads the 4 x 27 wave totals and adds them in wave
//      order, then runs the solve and pose update itself
//      (identical in every workgroup) and goes on to the next iteration.
// The states equal the begin-frame and two-launch iterations' bit for bit.
//
// The values are their own barrier: a reader polls them until none holds
// the empty marker (no counter, whose atomics serialise at the coherence
// point, and no flag behind the values, a second round trip there; the
// two stages keep the polled bytes small).  Iterations use three buffers
// of each stage in turn, numbered on from one launch to the next (the
// epoch, read by every workgroup before its first store and advanced by
// workgroup 0 after its last read).  In iteration k a writer first empties
// its entries of buffer k + 1, which hold iteration k - 2's values: every
// reader of those has finished with them, since this writer has seen
// iteration k - 1's wave totals, which depend (through the leaders) on
// every workgroup's iteration k - 1 values, each stored after its
// iteration k - 2 reads.  The writer waits for the emptying stores to
// complete before it stores iteration k's values, so a reader that has
// seen those finds its entries of buffer k + 1 empty until iteration k + 1
// fills them.  The buffers of a launch's first iteration were emptied by
// the last iteration of the launch before (or at creation or reset).
// ------------------------------------------------------------------------
constexpr int kFrameGroups = 3;                                // three blocks of 512 pixels per round
constexpr int kFrameThreads = kFrameGroups * kIcpThreads;
constexpr int kFrameBlocks = kIcpThreads;                      // one workgroup per cross-sum thread
constexpr int kFrameBuffers = 3;
constexpr int kStage1Entries = kIcpEntries * kFrameBlocks;     // [27][kFrameBlocks] per buffer
constexpr int kStage2Entries = kIcpWaves * 32;                 // [4 waves][32] per buffer
constexpr int kGateEntries = 4 * kFrameBlocks;                 // the gate's workgroup sums
constexpr int kFrameEntries = kFrameBuffers * (kStage1Entries + kStage2Entries) + kGateEntries;
constexpr unsigned kPollLimit = 1u << 21;                      // about 0.1 s of polls
// Empty-entry marker: a NaN whose payload no arithmetic produces (the
// generated NaN is 0x7ff8000000000000, and a propagated single-precision
// NaN has the low 29 bits clear).
constexpr uint64_t kEmptyEntry = 0xfff4f1a600000001ull;

struct IcpFrameLevel {
    const kfx_xyz *vc, *nc, *vp, *np;
    IcpFrame f;      // previous pose taken from the state
    int iterations;
    int nblk;        // the two-launch form's workgroups at this level
    int two;         // two pixels per thread
};

struct IcpFrameArgs {
    IcpFrameLevel lvl[KFX_ICP_LEVELS];
    IcpState* st;
    uint64_t* entries;   // stage 1 [kFrameBuffers][27][kFrameBlocks], then stage 2 [kFrameBuffers][4][32]:
                         // double bits or kEmptyEntry
    unsigned* epoch;     // iterations run by earlier launches (buffer numbering)
    IcpState* mirror;    // pinned host copy of the final state, or null
    int32_t* host_done;  // pinned flag raised after the host copy (with mirror), or null
    int begin;           // start of a frame (icp_begin_frame_kernel's update first)
    int tagged;          // level l's iterations carry tag l (early exit)
    double conv_rot, conv_trans;
    FrameGate gate;          // the fallback gate (the kernel instance with the gate)
    uint64_t* gate_entries;  // [kFrameBlocks][4] workgroup gate sums (double bits) or kEmptyEntry
};

template <int M>
__device__ __forceinline__ double group_total(const IcpGroupSum<M>& g) {
    if constexpr (M == kSumDouble) return g;
    else return (double)g.x + (double)g.y;
}

// State copy by 8-byte words, spread over the workgroup (over the first
// wave: first_wave = true).
__device__ __forceinline__ void copy_state(IcpState* dst, const IcpState* src, bool first_wave = false) {
    static_assert(sizeof(IcpState) % 8 == 0, "state copy by 8-byte words");
    const uint2* s = reinterpret_cast<const uint2*>(src);
    uint2* d = reinterpret_cast<uint2*>(dst);
    const int step = first_wave ? 64 : (int)blockDim.x;
    for (int i = threadIdx.x; i < (int)(sizeof(IcpState) / 8); i += step) d[i] = s[i];
}

// The fallback gate at the final pose (depth_consistency_kernel and
// consistency_reduce_kernel in one pass, with their sums in their order):
// group g of workgroup w takes the gate's 256-pixel blocks w + kFrameBlocks
// (3 r + g) (round r), sums each block as the gate kernel does (the wave
// butterfly, then the waves in order) and the workgroup adds its blocks in
// increasing order, as the reduce kernel's thread w does; workgroup 0 then
// waits for every workgroup's 4 sums, runs the reduce kernel's tree over
// them and writes the totals and the decision, and empties the entries for
// the next launch (every workgroup has stored its own by then).  Rounds are
// taken in pairs, both depth gathers issued before either is used.  (The
// model vertices are loaded here, not ahead during the last solve: staging
// them in LDS there cost more in the iterations' register pressure than it
// saved.)

__device__ __forceinline__ int gate_pixel(int w, int r, int t, int nblk) {
    const int blk = w + kFrameBlocks * (kFrameGroups * r + t / kIcpThreads);
    return blk < nblk ? blk * 256 + t % kIcpThreads : 0x7fffffff;
}

__device__ __forceinline__ void frame_gate(const IcpFrameArgs& a, const IcpState& s, int w, int grp, int lt,
                                           int& fault) {
    __shared__ double gred[2][kFrameGroups][kIcpWaves][4];
    __shared__ double gtree[4][kFrameBlocks];
    const FrameGate& G = a.gate;
    const int P = G.intr.cols * G.intr.rows, nblk = (P + 255) / 256;
    const int rounds = (nblk + kFrameBlocks * kFrameGroups - 1) / (kFrameBlocks * kFrameGroups);
    float pose[12];
    for (int i = 0; i < 9; ++i) pose[i] = s.Rf[i];
    for (int i = 0; i < 3; ++i) pose[9 + i] = s.tf[i];
    const int lane = lt & 63, wave = lt >> 6;
    double acc = 0.0;  // thread k < 4: entry k over this workgroup's blocks
    for (int r0 = 0; r0 < rounds; r0 += 2) {
        double v[2][4];
        int idx[2];
        float cz[2];
        for (int h = 0; h < 2; ++h) {
            const int r = r0 + h, p = gate_pixel(w, r, threadIdx.x, nblk);
            const kfx_xyz x = (r < rounds && p < P) ? G.vmap_model[p] : kfx_xyz{NAN, NAN, NAN};
            gate_project(x, pose, G.intr, v[h], idx[h], cz[h]);
        }
        uint16_t d[2];
        for (int h = 0; h < 2; ++h) d[h] = idx[h] >= 0 ? G.depth[idx[h]] : (uint16_t)0;
        for (int h = 0; h < 2; ++h) {
            if (idx[h] >= 0) gate_depth(d[h], cz[h], G.max_diff, v[h]);
            for (int k = 0; k < 4; ++k) v[h][k] = wave_sum_xor(v[h][k], lane);
            if (lane == 0)
                for (int k = 0; k < 4; ++k) gred[h][grp][wave][k] = v[h][k];
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            const int k = threadIdx.x;
            for (int h = 0; h < 2; ++h)
                for (int q = 0; q < kFrameGroups; ++q)
                    if (w + kFrameBlocks * (kFrameGroups * (r0 + h) + q) < nblk)
                        acc += ((gred[h][q][0][k] + gred[h][q][1][k]) + gred[h][q][2][k]) + gred[h][q][3][k];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) store_coherent(&a.gate_entries[w * 4 + threadIdx.x], (uint64_t)__double_as_longlong(acc));
    if (w != 0) return;
    if (threadIdx.x < kFrameBlocks) {
        uint64_t x[4];
        bool ready = false;
        for (unsigned polls = 0; !ready; ++polls) {
            for (int k = 0; k < 4; ++k) x[k] = load_coherent(&a.gate_entries[threadIdx.x * 4 + k]);
            ready = x[0] != kEmptyEntry && x[1] != kEmptyEntry && x[2] != kEmptyEntry && x[3] != kEmptyEntry;
            if (!ready && polls + 1 == kPollLimit) {
                fault = 1;
                break;
            }
        }
        for (int k = 0; k < 4; ++k) gtree[k][threadIdx.x] = __longlong_as_double((long long)x[k]);
    }
    // The reduce kernel's halving tree: h = 128 and 64 through LDS, then
    // h = 32 .. 1 in the first wave.
    __syncthreads();
    if (threadIdx.x < 128)
        for (int k = 0; k < 4; ++k) gtree[k][threadIdx.x] += gtree[k][threadIdx.x + 128];
    __syncthreads();
    if (threadIdx.x < 64) {
        double t[4];
        for (int k = 0; k < 4; ++k) t[k] = wave_sum_xor(gtree[k][threadIdx.x] + gtree[k][threadIdx.x + 64], lane);
        if (threadIdx.x < 4) G.stats[threadIdx.x] = threadIdx.x == 0 ? t[0] : threadIdx.x == 1 ? t[1] : threadIdx.x == 2 ? t[2] : t[3];
        if (threadIdx.x == 0) {
            const int32_t f = gate_decision(t[0], t[1], t[3], G.fallback_m);
            *G.flag = f;
            if (G.host_flag != nullptr) *G.host_flag = f;
        }
    }
    if (threadIdx.x < kFrameBlocks)
        for (int k = 0; k < 4; ++k) store_coherent(&a.gate_entries[threadIdx.x * 4 + k], kEmptyEntry);
}

template <int M, bool kGate>
__global__ void __launch_bounds__(kFrameThreads) icp_frame_kernel(IcpFrameArgs a) {
    __shared__ double wsum_d[kFrameGroups][kIcpWaves][32];
    __shared__ float2 wsum_f[kFrameGroups][kIcpWaves][32];
    __shared__ IcpGroupSum<M> red[kFrameGroups][32];
    __shared__ double total[32];
    __shared__ IcpState s;
    __shared__ int fault;
    const int grp = threadIdx.x / kIcpThreads, lt = threadIdx.x % kIcpThreads;
    const int w = blockIdx.x;
    const bool leader = w < kIcpWaves;
    const unsigned epoch = __hip_atomic_load(a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    copy_state(&s, a.st);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (a.begin) {
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) s.Rpi[3 * r + c] = (float)s.R[3 * c + r];
            for (int i = 0; i < 3; ++i) s.tp[i] = (float)s.t[i];
            s.status = KFX_OK;
            s.singular_count = 0;
            s.converged_tag = -1;
            s.iterations = 0;
        }
        fault = 0;
    }
    __syncthreads();
    unsigned k = 0;  // iterations so far in this launch
    const int per_round = kFrameBlocks * kFrameGroups;
    auto stage1 = [&](unsigned i) { return a.entries + (size_t)((epoch + i) % kFrameBuffers) * kStage1Entries; };
    auto stage2 = [&](unsigned i) {
        return a.entries + (size_t)kFrameBuffers * kStage1Entries + ((epoch + i) % kFrameBuffers) * kStage2Entries;
    };
    for (int l = KFX_ICP_LEVELS - 1; l >= 0; --l) {
        const IcpFrameLevel& L = a.lvl[l];
        const int tag = a.tagged ? l : -1;
        const int npix = L.f.cols * L.f.rows;
        // With one round (every block of the level in one pass of the
        // groups) a group's current-frame pixels stay in registers for all
        // of the level's iterations.
        const bool one_round = L.nblk <= per_round;
        kfx_xyz v[2], n[2];
        auto load = [&](int blk) {
            if (L.two) icp_load_current<2>(L.vc, L.nc, npix, blk, lt, v, n);
            else icp_load_current<1>(L.vc, L.nc, npix, blk, lt, reinterpret_cast<kfx_xyz(&)[1]>(v),
                                     reinterpret_cast<kfx_xyz(&)[1]>(n));
        };
        if (one_round && L.iterations > 0) load(w + kFrameBlocks * grp);
        for (int it = 0; it < L.iterations; ++it) {
            if (tag >= 0 && s.converged_tag == tag) break;  // the same in every workgroup
            uint64_t* cur1 = stage1(k);
            uint64_t* cur2 = stage2(k);
            if (threadIdx.x < kIcpEntries) {
                store_coherent(&stage1(k + 1)[threadIdx.x * kFrameBlocks + w], kEmptyEntry);
                if (leader) store_coherent(&stage2(k + 1)[w * 32 + threadIdx.x], kEmptyEntry);
            }
            IcpFrame f = L.f;
            icp_prev_from_state(f, &s);
            float Rc[9], tc[3];
            for (int i = 0; i < 9; ++i) Rc[i] = s.Rf[i];
            for (int i = 0; i < 3; ++i) tc[i] = s.tf[i];
            double acc = 0.0;  // threads 0..26: entry threadIdx.x over this workgroup's blocks
            for (int r0 = 0; r0 < L.nblk; r0 += per_round) {
                const int blk = w + r0 + kFrameBlocks * grp;
                if (!one_round) load(blk);
                const IcpGroupSum<M> g =
                    L.two ? icp_group_sums<2, M>(v, n, L.vp, L.np, f, Rc, tc, blk, lt, wsum_d[grp], wsum_f[grp])
                          : icp_group_sums<1, M>(reinterpret_cast<kfx_xyz(&)[1]>(v),
                                                 reinterpret_cast<kfx_xyz(&)[1]>(n), L.vp, L.np, f, Rc, tc, blk,
                                                 lt, wsum_d[grp], wsum_f[grp]);
                if (lt < kIcpEntries) red[grp][lt] = g;
                __syncthreads();
                if (threadIdx.x < kIcpEntries) {
#pragma unroll
                    for (int q = 0; q < kFrameGroups; ++q)
                        if (w + r0 + kFrameBlocks * q < L.nblk) acc += group_total<M>(red[q][threadIdx.x]);
                }
                __syncthreads();
            }
            if (threadIdx.x < kIcpEntries) {
                __builtin_amdgcn_s_waitcnt(kWaitVmemZero);  // the emptying stores above have completed
                store_coherent(&cur1[threadIdx.x * kFrameBlocks + w], (uint64_t)__double_as_longlong(acc));
            }
            // Stage 1 (leaders' first wave): lane l polls workgroup
            // 64 w + l's values, then the wave's butterfly.
            if (leader && threadIdx.x < 64) {
                const int src = w * 64 + threadIdx.x;
                uint64_t x[32];
#pragma unroll
                for (int e = 0; e < 32; ++e) x[e] = 0;  // +0.0
                bool ready = false;
                for (unsigned polls = 0; !ready; ++polls) {
#pragma unroll
                    for (int e = 0; e < kIcpEntries; ++e) x[e] = load_coherent(&cur1[e * kFrameBlocks + src]);
                    ready = true;
#pragma unroll
                    for (int e = 0; e < kIcpEntries; ++e) ready = ready && x[e] != kEmptyEntry;
                    if (!ready && polls + 1 == kPollLimit) {
                        fault = 1;
                        break;
                    }
                }
                double xd[32];
#pragma unroll
                for (int e = 0; e < 32; ++e) xd[e] = __longlong_as_double((long long)x[e]);
                const double wtot = wave_reduce_scatter32(xd, threadIdx.x);
                if ((threadIdx.x & 1) == 0 && (threadIdx.x >> 1) < kIcpEntries)
                    store_coherent(&cur2[w * 32 + (threadIdx.x >> 1)], (uint64_t)__double_as_longlong(wtot));
            }
            // Stage 2 and the solve in the first wave: lane e < 27 polls
            // the four wave totals of entry e and adds them in wave order,
            // then the wave solves; one barrier after it.
            if (threadIdx.x < 64) {
                if (threadIdx.x < kIcpEntries) {
                    const int e = threadIdx.x;
                    uint64_t x[kIcpWaves];
                    for (unsigned polls = 1;; ++polls) {
                        bool ready = true;
#pragma unroll
                        for (int q = 0; q < kIcpWaves; ++q) {
                            x[q] = load_coherent(&cur2[q * 32 + e]);
                            ready = ready && x[q] != kEmptyEntry;
                        }
                        if (ready) break;
                        if (polls == kPollLimit) {
                            fault = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    double t = __longlong_as_double((long long)x[0]);
#pragma unroll
                    for (int q = 1; q < kIcpWaves; ++q) t += __longlong_as_double((long long)x[q]);
                    total[e] = t;
                }
                // The other lanes read total[] after these stores: one wave,
                // whose LDS operations complete in order.
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                solve_and_update(&s, total, s.R, s.t, s.iterations, tag, a.conv_rot, a.conv_trans, threadIdx.x);
            }
            __syncthreads();
            ++k;
        }
    }
    if constexpr (kGate) frame_gate(a, s, w, grp, lt, fault);
    if (w == 0) {
        if (threadIdx.x == 0) {
            s.fault = fault;
            // Every workgroup read the epoch before its first store, and
            // the last wave totals (read by workgroup 0) depend on every
            // workgroup's last values.
            __hip_atomic_store(a.epoch, epoch + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        copy_state(a.st, &s);
        if (a.mirror != nullptr && threadIdx.x < 64) {
            // The host copy by the first wave, which wrote the gate's host
            // outputs too; then the host's completion flag, released at
            // system scope after all of them (the host reads the frame's
            // result from there while later work runs on).
            copy_state(a.mirror, &s, true);
            if (a.host_done != nullptr && threadIdx.x == 0)
                __hip_atomic_store(a.host_done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Every entry of the frame buffers empty (creation, reset).
__global__ void icp_frame_empty_kernel(uint64_t* entries, unsigned* epoch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kFrameEntries) entries[i] = kEmptyEntry;
    if (i == 0) *epoch = 0;
}

// Start of a frame: the current estimate becomes the previous pose (as R^T,
// t in single precision) and the per-frame counters reset.
__global__ void icp_begin_frame_kernel(IcpState* st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) st->Rpi[3 * r + c] = (float)st->R[3 * c + r];
    for (int i = 0; i < 3; ++i) st->tp[i] = (float)st->t[i];
    st->status = KFX_OK;
    st->singular_count = 0;
    st->converged_tag = -1;
    st->iterations = 0;
    st->fault = 0;
}

// The solve and pose update of an iteration applied to 27 sums in device
// memory (one wave).
__global__ void __launch_bounds__(64) icp_solve_sums_kernel(IcpState* __restrict__ st,
                                                             const double* __restrict__ sums) {
    __shared__ double total[32];
    if (threadIdx.x < kIcpEntries) total[threadIdx.x] = sums[threadIdx.x];
    double R0[9], t0[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) R0[i] = st->R[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) t0[i] = st->t[i];
    const int iterations = st->iterations;
    __syncthreads();
    solve_and_update(st, total, R0, t0, iterations, -1, 0.0, 0.0, threadIdx.x);
}

__global__ void icp_translate_kernel(IcpState* st, double dx, double dy, double dz) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->t[0] += dx; st->t[1] += dy; st->t[2] += dz;
    for (int i = 0; i < 3; ++i) st->tf[i] = (float)st->t[i];
}

__global__ void icp_reset_kernel(IcpState* st, kfx_pose pose) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i < 9; ++i) { st->R[i] = pose.R[i]; st->Rf[i] = (float)pose.R[i]; }
    for (int i = 0; i < 3; ++i) { st->t[i] = pose.t[i]; st->tf[i] = (float)pose.t[i]; }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) st->Rpi[3 * r + c] = (float)pose.R[3 * c + r];
    for (int i = 0; i < 3; ++i) st->tp[i] = (float)pose.t[i];
    st->status = KFX_OK;
    st->singular_count = 0;
    st->converged_tag = -1;
    st->iterations = 0;
    st->fault = 0;
}

IcpFrame make_frame(const kfx_matrix_3f* rprev_inv, const kfx_point_3f* tprev, const kfx_intrinsics* intr,
                    float dist_thr, float angle_thr) {
    IcpFrame f;
    f.prev_from_state = (rprev_inv == nullptr);
    for (int r = 0; r < 3; ++r) {
        f.rprev_inv[3 * r + 0] = rprev_inv ? rprev_inv->data[r].x : 0.0f;
        f.rprev_inv[3 * r + 1] = rprev_inv ? rprev_inv->data[r].y : 0.0f;
        f.rprev_inv[3 * r + 2] = rprev_inv ? rprev_inv->data[r].z : 0.0f;
    }
    f.tprev[0] = tprev ? tprev->x : 0.0f;
    f.tprev[1] = tprev ? tprev->y : 0.0f;
    f.tprev[2] = tprev ? tprev->z : 0.0f;
    f.fx = intr->fx; f.fy = intr->fy; f.cx = intr->cx; f.cy = intr->cy;
    f.cols = intr->cols; f.rows = intr->rows;
    f.dist_thr = dist_thr;
    f.angle_thr = angle_thr;
    return f;
}

}  // namespace kfx

struct kfx_icp {
    kfx::IcpState* state;
    void* partials;    // [27][workgroup] sums of the fused iteration (float2 or double)
    int sum_mode;      // KFX_ICP_SUM_*
    kfx::IcpState* host_state;  // pinned host copy for kfx_icp_read
    int max_blocks;
    bool publish_next;  // the next iteration's solve writes host_state
    bool published;     // a publishing solve was enqueued since the last reset or translate
    unsigned* ticket;   // [0] one-launch iteration ticket (zero between launches), [1] frame buffer epoch
    int launch_mode;    // KFX_ICP_LAUNCH_*
    uint64_t* frame_entries;  // the one-launch frame's stage 1 and stage 2 buffers (kFrameEntries words)
    bool frame_resident;     // the device holds all kFrameBlocks workgroups of icp_frame_kernel at once
    int32_t* host_done;      // pinned flag the publishing frame launch raises after the host copy
    bool done_raised;        // the last publishing launch was a frame launch (it raises host_done)
};

using namespace kfx;

template <int M>
static void icp_iterate_launch(kfx_icp* h, const kfx_xyz* vc, const kfx_xyz* nc, const kfx_xyz* vp, const kfx_xyz* np,
                               const IcpFrame& f, int nblk, bool two, int tag, double conv_rot, double conv_trans,
                               bool publish) {
    IcpState* mirror = publish ? h->host_state : nullptr;
    if (h->launch_mode == KFX_ICP_LAUNCH_ONE) {
        char name[32];
        snprintf(name, sizeof(name), "icp_iter@%d", f.cols);
        ProfScope prof(name);
        if (two)
            icp_iter_kernel<2, M><<<nblk, kIcpThreads, 0, stream()>>>(vc, nc, vp, np, f, h->state, h->partials,
                                                                      h->ticket, tag, conv_rot, conv_trans, mirror);
        else
            icp_iter_kernel<1, M><<<nblk, kIcpThreads, 0, stream()>>>(vc, nc, vp, np, f, h->state, h->partials,
                                                                      h->ticket, tag, conv_rot, conv_trans, mirror);
        return;
    }
    {
        // Named by image width, so a profile separates the pyramid levels.
        char name[32];
        snprintf(name, sizeof(name), "icp_sums@%d", f.cols);
        ProfScope prof(name);
        if (two)
            icp_sums_kernel<2, M><<<nblk, kIcpThreads, 0, stream()>>>(vc, nc, vp, np, f, h->state, h->partials, tag);
        else
            icp_sums_kernel<1, M><<<nblk, kIcpThreads, 0, stream()>>>(vc, nc, vp, np, f, h->state, h->partials, tag);
    }
    ProfScope prof("icp_solve");
    icp_solve_kernel<M><<<1, kIcpThreads, 0, stream()>>>(h->state, h->partials, nblk, tag, conv_rot,
                                                                conv_trans, mirror);
}

// Whether the device holds every workgroup of icp_frame_kernel at once (its
// grid barrier needs them all resident): workgroups per CU at the kernel's
// resources times the CUs.
static bool frame_kernel_resident() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return false;
    int per_cu = 1 << 30;
    const void* kernels[] = {reinterpret_cast<const void*>(&icp_frame_kernel<kSumKahan, false>),
                             reinterpret_cast<const void*>(&icp_frame_kernel<kSumPlain, false>),
                             reinterpret_cast<const void*>(&icp_frame_kernel<kSumDouble, false>),
                             reinterpret_cast<const void*>(&icp_frame_kernel<kSumKahan, true>),
                             reinterpret_cast<const void*>(&icp_frame_kernel<kSumPlain, true>),
                             reinterpret_cast<const void*>(&icp_frame_kernel<kSumDouble, true>)};
    for (const void* k : kernels) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kFrameThreads, 0) != hipSuccess) return false;
        per_cu = n < per_cu ? n : per_cu;
    }
    return (long long)per_cu * cus >= kFrameBlocks;
}

int kfx::icp_frame_gated(kfx_icp* h, const kfx_icp_level* levels, int begin_frame, float distance_threshold,
                         float angle_threshold, int tagged, double converge_rotation, double converge_translation,
                         const FrameGate* gate, const Workspace* gate_ws) {
    // A publish request applies to this call only, whether or not it
    // passes its checks.
    const bool publish = h != nullptr && h->publish_next;
    if (h != nullptr) h->publish_next = false;
    if (publish) h->done_raised = false;  // until a frame launch raises the flag
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && levels, "kfx_icp_frame: null pointer");
    KFX_REQUIRE(gate == nullptr || (gate->vmap_model && gate->depth && gate->stats && gate->flag && gate_ws &&
                                    gate->intr.cols > 0 && gate->intr.rows > 0),
                "kfx_icp_frame: incomplete gate");
    int last = -1;  // the finest level with iterations: its last one publishes
    for (int l = KFX_ICP_LEVELS - 1; l >= 0; --l) {
        const kfx_icp_level& L = levels[l];
        KFX_REQUIRE(L.iterations >= 0, "kfx_icp_frame: level %d: negative iteration count", l);
        if (L.iterations == 0) continue;
        last = l;
        KFX_REQUIRE(L.vmap_current && L.nmap_current && L.vmap_previous && L.nmap_previous,
                    "kfx_icp_frame: level %d: null map", l);
        KFX_REQUIRE(L.intr.cols > 0 && L.intr.rows > 0, "kfx_icp_frame: level %d: empty image", l);
        const unsigned npix = (unsigned)(L.intr.cols * L.intr.rows);
        KFX_REQUIRE(divup(npix, (unsigned)kIcpThreads) <= (unsigned)h->max_blocks,
                    "kfx_icp_frame: level %d: image of %u pixels exceeds the handle's capacity", l, npix);
    }
    if (h->launch_mode != KFX_ICP_LAUNCH_FRAME || !h->frame_resident) {
        // The same frame as separate launches.
        if (begin_frame) KFX_TRY(kfx_icp_begin_frame(h));
        for (int l = KFX_ICP_LEVELS - 1; l >= 0; --l)
            for (int it = 0; it < levels[l].iterations; ++it) {
                if (publish && l == last && it == levels[l].iterations - 1) h->publish_next = true;
                KFX_TRY(kfx_icp_iterate(h, levels[l].vmap_current, levels[l].nmap_current, levels[l].vmap_previous,
                                        levels[l].nmap_previous, nullptr, nullptr, &levels[l].intr,
                                        distance_threshold, angle_threshold, tagged ? l : -1, converge_rotation,
                                        converge_translation));
            }
        if (gate != nullptr)
            KFX_TRY(depth_consistency(*gate_ws, gate->vmap_model, &h->state->Rf[0], gate->depth, &gate->intr,
                                      gate->max_diff, gate->stats, gate->fallback_m, gate->flag, gate->host_flag));
        return KFX_OK;
    }
    IcpFrameArgs a;
    for (int l = 0; l < KFX_ICP_LEVELS; ++l) {
        const kfx_icp_level& L = levels[l];
        IcpFrameLevel& d = a.lvl[l];
        d.vc = L.vmap_current;
        d.nc = L.nmap_current;
        d.vp = L.vmap_previous;
        d.np = L.nmap_previous;
        d.iterations = L.iterations;
        d.two = 0;
        d.nblk = 0;
        if (L.iterations > 0) {
            d.f = make_frame(nullptr, nullptr, &L.intr, distance_threshold, angle_threshold);
            const unsigned npix = (unsigned)(L.intr.cols * L.intr.rows);
            d.two = npix >= kIcpTwoPixelMin ? 1 : 0;
            d.nblk = (int)divup(npix, (unsigned)(kIcpThreads * (d.two ? 2 : 1)));
        } else {
            d.f = IcpFrame{};
        }
    }
    a.st = h->state;
    a.entries = h->frame_entries;
    a.epoch = h->ticket + 1;
    a.mirror = (publish && last >= 0) ? h->host_state : nullptr;
    a.host_done = a.mirror != nullptr ? h->host_done : nullptr;
    a.begin = begin_frame ? 1 : 0;
    a.tagged = tagged ? 1 : 0;
    a.conv_rot = converge_rotation;
    a.conv_trans = converge_translation;
    a.gate = gate != nullptr ? *gate : FrameGate{};
    a.gate_entries = h->frame_entries + kFrameBuffers * (kStage1Entries + kStage2Entries);
    {
        ProfScope prof("icp_frame");
        auto launch = [&](auto mode, auto with_gate) {
            icp_frame_kernel<decltype(mode)::value, decltype(with_gate)::value>
                <<<kFrameBlocks, kFrameThreads, 0, stream()>>>(a);
        };
        auto launch_mode = [&](auto with_gate) {
            switch (h->sum_mode) {
                case KFX_ICP_SUM_FP32: launch(std::integral_constant<int, kSumPlain>{}, with_gate); break;
                case KFX_ICP_SUM_FP64: launch(std::integral_constant<int, kSumDouble>{}, with_gate); break;
                default: launch(std::integral_constant<int, kSumKahan>{}, with_gate); break;
            }
        };
        if (gate != nullptr) launch_mode(std::true_type{});
        else launch_mode(std::false_type{});
        KFX_CHECK_LAUNCH();
    }
    if (a.mirror != nullptr) {
        h->published = true;
        h->done_raised = true;
    }
    return KFX_OK;
}

// The host copy of the state as a result (after the copy is complete).
static int read_host_state(kfx_icp* h, kfx_icp_result* out) {
    const IcpState& s = *h->host_state;
    if (s.fault != 0)
        return set_error(KFX_ERR_HIP, "%s: a grid barrier wait of the one-launch frame ran out (the state is not valid)",
                         __func__);
    for (int i = 0; i < 9; ++i) out->pose.R[i] = s.R[i];
    for (int i = 0; i < 3; ++i) out->pose.t[i] = s.t[i];
    for (int i = 0; i < 36; ++i) out->A[i] = s.A[i];
    for (int i = 0; i < 6; ++i) { out->b[i] = s.b[i]; out->x[i] = s.x[i]; }
    out->status = s.status;
    out->singular_count = s.singular_count;
    out->iterations = s.iterations;
    out->converged_tag = s.converged_tag;
    return KFX_OK;
}

void kfx::icp_clear_done(kfx_icp* h) { __atomic_store_n(h->host_done, 0, __ATOMIC_RELEASE); }

int kfx::icp_read_published_early(kfx_icp* h, kfx_icp_result* out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && out, "icp_read_published_early: null pointer");
    if (!h->done_raised || !h->published) return kfx_icp_read_published(h, out);
    // Spin on the flag; every 256 polls ask whether the stream has ended
    // (then the flag must be up) or failed, so a failed launch cannot hold
    // the host here.
    for (unsigned polls = 1; __atomic_load_n(h->host_done, __ATOMIC_ACQUIRE) == 0; ++polls) {
        if (polls % 256 != 0) continue;
        const hipError_t q = hipStreamQuery(stream());
        if (q == hipErrorNotReady) continue;
        KFX_CHECK_HIP(q);
        if (__atomic_load_n(h->host_done, __ATOMIC_ACQUIRE) == 0)
            return set_error(KFX_ERR_INVALID_ARG, "%s: the stream ended without the frame launch's completion flag",
                             __func__);
        break;
    }
    return read_host_state(h, out);
}

extern "C" {

int kfx_estimate_combined(const kfx_matrix_3f* rotation_current, const kfx_point_3f* translation_current,
                          const kfx_xyz* vmap_current, const kfx_xyz* nmap_current,
                          const kfx_matrix_3f* rotation_previous_inverse, const kfx_point_3f* translation_previous,
                          const kfx_intrinsics* intr, const kfx_xyz* vmap_previous, const kfx_xyz* nmap_previous,
                          float distance_threshold, float angle_threshold, float* gbuf, int gbuf_stride,
                          float* mbuf, double* A, double* b) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(rotation_current && translation_current && vmap_current && nmap_current && rotation_previous_inverse &&
                    translation_previous && intr && vmap_previous && nmap_previous && gbuf && mbuf && A && b,
                "kfx_estimate_combined: null pointer");
    KFX_REQUIRE(intr->cols > 0 && intr->rows > 0, "kfx_estimate_combined: empty image");
    const int gx = (int)divup(intr->cols, kEcBX), gy = (int)divup(intr->rows, kEcBY);
    const int buf_len = gx * gy;
    KFX_REQUIRE(gbuf_stride >= buf_len, "kfx_estimate_combined: gbuf stride %d < %d partials per entry",
                gbuf_stride, buf_len);
    IcpFrame f = make_frame(rotation_previous_inverse, translation_previous, intr, distance_threshold,
                            angle_threshold);
    {
        ProfScope prof("estimate_combined");
        estimate_combined_kernel<<<dim3(gx, gy), dim3(kEcBX, kEcBY), 0, stream()>>>(
            vmap_current, nmap_current, vmap_previous, nmap_previous, f, *rotation_current, *translation_current,
            gbuf, gbuf_stride);
        KFX_CHECK_LAUNCH();
    }
    {
        ProfScope prof("transform_estimation");
        transform_estimation_kernel<<<27, 256, 0, stream()>>>(gbuf, mbuf, buf_len, gbuf_stride);
        KFX_CHECK_LAUNCH();
    }
    float h[27];
    KFX_CHECK_HIP(hipMemcpyAsync(h, mbuf, sizeof(h), hipMemcpyDeviceToHost, stream()));
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    int shift = 0;
    for (int k = 0; k < 6; ++k)
        for (int j = k; j < 7; ++j) {
            const double value = h[shift++];
            if (j == 6) b[k] = value;
            else A[j * 6 + k] = A[k * 6 + j] = value;
        }
    return KFX_OK;
}

int kfx_clean_buff(float* gbuf, int width) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(gbuf != nullptr && width > 0, "kfx_clean_buff: bad arguments");
    clean_buff_kernel<<<divup(width, 256), 256, 0, stream()>>>(gbuf, width);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_icp_create(int max_pixels, kfx_icp** out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(out != nullptr && max_pixels > 0, "kfx_icp_create: bad arguments");
    kfx_icp* h = new kfx_icp();
    h->max_blocks = (int)divup((unsigned)max_pixels, (unsigned)kIcpThreads);
    hipError_t e = hipMalloc((void**)&h->state, sizeof(IcpState));
    h->sum_mode = KFX_ICP_SUM_FP64;
    h->host_state = nullptr;
    h->partials = nullptr;
    h->publish_next = false;
    h->published = false;
    h->ticket = nullptr;
    h->launch_mode = KFX_ICP_LAUNCH_FRAME;
    h->frame_entries = nullptr;
    h->frame_resident = frame_kernel_resident();
    h->host_done = nullptr;
    h->done_raised = false;
    if (e == hipSuccess) e = hipMalloc((void**)&h->partials, sizeof(double) * kIcpEntries * (size_t)h->max_blocks);
    if (e == hipSuccess)
        e = hipMalloc((void**)&h->frame_entries, sizeof(uint64_t) * kFrameEntries);
    if (e == hipSuccess) e = hipMalloc((void**)&h->ticket, 2 * sizeof(unsigned));
    if (e == hipSuccess) e = hipMemsetAsync(h->ticket, 0, 2 * sizeof(unsigned), stream());
    if (e == hipSuccess) {
        icp_frame_empty_kernel<<<divup(kFrameEntries, 256), 256, 0, stream()>>>(
            h->frame_entries, h->ticket + 1);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipHostMalloc((void**)&h->host_state, sizeof(IcpState), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&h->host_done, 64, hipHostMallocDefault);
    if (e == hipSuccess) *h->host_done = 0;
    if (e == hipSuccess) e = hipMemsetAsync(h->state, 0, sizeof(IcpState), stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    if (e != hipSuccess) {
        (void)hipFree(h->state);
        (void)hipFree(h->partials);
        (void)hipFree(h->frame_entries);
        (void)hipFree(h->ticket);
        if (h->host_state) (void)hipHostFree(h->host_state);
        if (h->host_done) (void)hipHostFree(h->host_done);
        delete h;
        return set_error(KFX_ERR_HIP, "kfx_icp_create: %s", hipGetErrorString(e));
    }
    *out = h;
    return KFX_OK;
}

int kfx_icp_destroy(kfx_icp* h) {
    if (h == nullptr) return KFX_OK;
    KFX_REQUIRE_INIT();
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    (void)hipFree(h->state);
    (void)hipFree(h->partials);
    (void)hipFree(h->frame_entries);
    (void)hipFree(h->ticket);
    (void)hipHostFree(h->host_state);
    (void)hipHostFree(h->host_done);
    delete h;
    return KFX_OK;
}

int kfx_icp_reset(kfx_icp* h, const kfx_pose* pose) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && pose, "kfx_icp_reset: null pointer");
    icp_reset_kernel<<<1, 64, 0, stream()>>>(h->state, *pose);
    KFX_CHECK_LAUNCH();
    // The one-launch frame's buffers empty again (after a frame whose
    // polls ran out they may not be).
    icp_frame_empty_kernel<<<divup(kFrameEntries, 256), 256, 0, stream()>>>(
        h->frame_entries, h->ticket + 1);
    KFX_CHECK_LAUNCH();
    h->published = false;
    return KFX_OK;
}

int kfx_icp_iterate(kfx_icp* h, const kfx_xyz* vmap_current, const kfx_xyz* nmap_current,
                    const kfx_xyz* vmap_previous, const kfx_xyz* nmap_previous,
                    const kfx_matrix_3f* rotation_previous_inverse, const kfx_point_3f* translation_previous,
                    const kfx_intrinsics* intr, float distance_threshold, float angle_threshold, int level_tag,
                    double converge_rotation, double converge_translation) {
    // A publish request applies to this call only, whether or not it
    // passes its checks.
    const bool publish = h != nullptr && h->publish_next;
    if (h != nullptr) h->publish_next = false;
    if (publish) h->done_raised = false;  // until a frame launch raises the flag
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && vmap_current && nmap_current && vmap_previous && nmap_previous && intr,
                "kfx_icp_iterate: null pointer");
    KFX_REQUIRE((rotation_previous_inverse == nullptr) == (translation_previous == nullptr),
                "kfx_icp_iterate: give both previous-pose arguments or neither");
    KFX_REQUIRE(intr->cols > 0 && intr->rows > 0, "kfx_icp_iterate: empty image");
    const unsigned npix = (unsigned)(intr->cols * intr->rows);
    KFX_REQUIRE(divup(npix, (unsigned)kIcpThreads) <= (unsigned)h->max_blocks,
                "kfx_icp_iterate: image of %u pixels exceeds the handle's capacity", npix);
    // Two pixels per thread on full-resolution images (fewer, fuller
    // workgroups), one on the coarse levels (enough workgroups to spread).
    const bool two = npix >= kIcpTwoPixelMin;
    const int nblk = (int)divup(npix, (unsigned)(kIcpThreads * (two ? 2 : 1)));
    IcpFrame f = make_frame(rotation_previous_inverse, translation_previous, intr, distance_threshold,
                            angle_threshold);
    auto launch = [&](auto mode) {
        icp_iterate_launch<decltype(mode)::value>(h, vmap_current, nmap_current, vmap_previous, nmap_previous, f,
                                                  nblk, two, level_tag, converge_rotation, converge_translation,
                                                  publish);
    };
    switch (h->sum_mode) {
        case KFX_ICP_SUM_FP32: launch(std::integral_constant<int, kSumPlain>{}); break;
        case KFX_ICP_SUM_FP64: launch(std::integral_constant<int, kSumDouble>{}); break;
        default: launch(std::integral_constant<int, kSumKahan>{}); break;
    }
    KFX_CHECK_LAUNCH();
    if (publish) h->published = true;
    return KFX_OK;
}

int kfx_icp_set_launch_mode(kfx_icp* h, int mode) {
    KFX_REQUIRE(h != nullptr, "kfx_icp_set_launch_mode: null handle");
    KFX_REQUIRE(mode == KFX_ICP_LAUNCH_FRAME || mode == KFX_ICP_LAUNCH_ONE || mode == KFX_ICP_LAUNCH_TWO,
                "kfx_icp_set_launch_mode: unknown mode %d", mode);
    h->launch_mode = mode;
    return KFX_OK;
}

int kfx_icp_frame_resident(kfx_icp* h, int* resident) {
    KFX_REQUIRE(h && resident, "kfx_icp_frame_resident: null pointer");
    *resident = h->frame_resident ? 1 : 0;
    return KFX_OK;
}

int kfx_icp_frame(kfx_icp* h, const kfx_icp_level* levels, int begin_frame, float distance_threshold,
                  float angle_threshold, int tagged, double converge_rotation, double converge_translation) {
    return kfx::icp_frame_gated(h, levels, begin_frame, distance_threshold, angle_threshold, tagged,
                                converge_rotation, converge_translation, nullptr, nullptr);
}

int kfx_icp_set_sum_mode(kfx_icp* h, int mode) {
    KFX_REQUIRE(h != nullptr, "kfx_icp_set_sum_mode: null handle");
    KFX_REQUIRE(mode == KFX_ICP_SUM_FP32_KAHAN || mode == KFX_ICP_SUM_FP32 || mode == KFX_ICP_SUM_FP64,
                "kfx_icp_set_sum_mode: unknown mode %d", mode);
    h->sum_mode = mode;
    return KFX_OK;
}

int kfx_icp_begin_frame(kfx_icp* h) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h != nullptr, "kfx_icp_begin_frame: null handle");
    ProfScope prof("icp_begin_frame");
    icp_begin_frame_kernel<<<1, 64, 0, stream()>>>(h->state);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_icp_solve_sums(kfx_icp* h, const double sums[27]) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && sums, "kfx_icp_solve_sums: null pointer");
    // The partial-sum buffer holds at least 27 doubles; the copy and the
    // solve are ordered on the stream after any iteration enqueued before.
    KFX_CHECK_HIP(hipMemcpyAsync(h->partials, sums, sizeof(double) * kIcpEntries, hipMemcpyHostToDevice, stream()));
    icp_solve_sums_kernel<<<1, 64, 0, stream()>>>(h->state, static_cast<const double*>(h->partials));
    KFX_CHECK_LAUNCH();
    h->published = false;
    return KFX_OK;
}

int kfx_icp_translate(kfx_icp* h, double dx, double dy, double dz) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h != nullptr, "kfx_icp_translate: null handle");
    icp_translate_kernel<<<1, 64, 0, stream()>>>(h->state, dx, dy, dz);
    KFX_CHECK_LAUNCH();
    h->published = false;
    return KFX_OK;
}

int kfx_icp_read(kfx_icp* h, kfx_icp_result* out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && out, "kfx_icp_read: null pointer");
    KFX_CHECK_HIP(hipMemcpyAsync(h->host_state, h->state, sizeof(IcpState), hipMemcpyDeviceToHost, stream()));
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    return read_host_state(h, out);
}

int kfx_icp_publish_next(kfx_icp* h) {
    KFX_REQUIRE(h != nullptr, "kfx_icp_publish_next: null handle");
    h->publish_next = true;
    return KFX_OK;
}

int kfx_icp_read_published(kfx_icp* h, kfx_icp_result* out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(h && out, "kfx_icp_read_published: null pointer");
    KFX_REQUIRE(h->published,
                "kfx_icp_read_published: no publishing iteration was enqueued since the last reset or translate");
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    return read_host_state(h, out);
}

int kfx_icp_pose_device(kfx_icp* h, const float** dev_pose, const float** dev_prev_pose_inv) {
    KFX_REQUIRE(h && dev_pose && dev_prev_pose_inv, "kfx_icp_pose_device: null pointer");
    static_assert(offsetof(IcpState, tf) == offsetof(IcpState, Rf) + 9 * sizeof(float), "pose layout");
    static_assert(offsetof(IcpState, tp) == offsetof(IcpState, Rpi) + 9 * sizeof(float), "prev pose layout");
    *dev_pose = &h->state->Rf[0];
    *dev_prev_pose_inv = &h->state->Rpi[0];
    return KFX_OK;
}

}  // extern "C"
