// rgbd.cpp — CPU restatement of the RGBD photometric odometry
// (Estimate_Isometry_By_RGBD, common/src/opencl_wrappers.cpp:1328-1466;
// kernels scan_list_of_valid_correspondence.cl:50-110 and
// compute_least_squares_equation.cl:37-85; solve and twist exponential
// opencl_wrappers.cpp:1282-1326).  The reference's path is not functional;
// the algorithm restated is the dense RGBD odometry of the OpenCV rgbd module
// that its code follows (weights 1 / (sigma + |diff|), Sobel scale 1/8,
// Jacobian rows of compute_least_squares_equation.cl:59-71, |det| test,
// exp of the 4x4 twist).  Single precision where the GPU works in single
// precision (frame preparation, correspondences), in the same operation
// order; double precision sums in pixel order.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include <math.h>
#include <string.h>

#include <vector>

#include "oracle.h"

namespace {

int reflect101(int i, int n) {
    if (i < 0) return n > 1 ? -i : 0;
    if (i >= n) return n > 1 ? 2 * n - 2 - i : 0;
    return i;
}

struct Level {
    const float *I, *D, *dx, *dy;
    int cols, rows;
};

Level level_of(const float* packed, int cols, int rows, int l) {
    const float* p = packed;
    for (int k = 0; k < l; ++k) p += 4 * (size_t)(cols >> k) * (rows >> k);
    const size_t n = (size_t)(cols >> l) * (rows >> l);
    return Level{p, p + n, p + 2 * n, p + 3 * n, cols >> l, rows >> l};
}

bool solve6(const double A[36], const double b[6], double x[6]) {
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
    if (!(fabs(det) >= 1e-6) || !std::isfinite(det)) return false;
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

}  // namespace

extern "C" {

void orc_rgbd_prepare(const kfx_color* rgb, const uint16_t* depth, int cols, int rows, int levels, float* out) {
    const size_t n0 = (size_t)cols * rows;
    float* I0 = out;
    float* D0 = out + n0;
    for (size_t i = 0; i < n0; ++i) {
        I0[i] = (0.299f * (float)rgb[i].r + 0.587f * (float)rgb[i].g) + 0.114f * (float)rgb[i].b;
        D0[i] = (float)depth[i] * 0.001f;
    }
    for (int l = 0; l < levels; ++l) {
        float* base = out;
        for (int k = 0; k < l; ++k) base += 4 * (size_t)(cols >> k) * (rows >> k);
        const int c = cols >> l, r = rows >> l;
        const size_t n = (size_t)c * r;
        float *I = base, *D = base + n, *dx = base + 2 * n, *dy = base + 3 * n;
        if (l > 0) {
            const Level p = level_of(out, cols, rows, l - 1);
            for (int y = 0; y < r; ++y)
                for (int x = 0; x < c; ++x) {
                    const float* r0 = p.I + (2 * y) * p.cols + 2 * x;
                    const float* r1 = r0 + p.cols;
                    I[y * c + x] = 0.25f * ((r0[0] + r0[1]) + (r1[0] + r1[1]));
                    D[y * c + x] = p.D[(2 * y) * p.cols + 2 * x];
                }
        }
        for (int y = 0; y < r; ++y)
            for (int x = 0; x < c; ++x) {
                const int xm = reflect101(x - 1, c), xp = reflect101(x + 1, c);
                const float* rm = I + reflect101(y - 1, r) * c;
                const float* r0 = I + y * c;
                const float* rp = I + reflect101(y + 1, r) * c;
                dx[y * c + x] = ((rm[xp] + 2.0f * r0[xp]) + rp[xp]) - ((rm[xm] + 2.0f * r0[xm]) + rp[xm]);
                dy[y * c + x] = ((rp[xm] + 2.0f * rp[x]) + rp[xp]) - ((rm[xm] + 2.0f * rm[x]) + rm[xp]);
            }
    }
}

int orc_rgbd_corr(const float* /*i_prev*/, const float* d_prev, const float* d_curr, const float* dx_curr,
                  const float* dy_curr, const kfx_intrinsics* in, const float* Ri, const float* ti,
                  float min_gradient_sq, float max_depth_diff, int32_t* corr) {
    const int W = in->cols, H = in->rows;
    std::vector<float> best((size_t)W * H, 0.0f);
    for (int i = 0; i < W * H; ++i) corr[i] = -1;
    int pairs = 0;
    for (int v1 = 0; v1 < H; ++v1)
        for (int u1 = 0; u1 < W; ++u1) {
            const int i1 = v1 * W + u1;
            const float d1 = d_curr[i1], gx = dx_curr[i1], gy = dy_curr[i1];
            if (!(d1 > 0.0f) || gx * gx + gy * gy < min_gradient_sq) continue;
            const float x1 = ((float)u1 - in->cx) / in->fx * d1, y1 = ((float)v1 - in->cy) / in->fy * d1;
            const float px = (Ri[0] * x1 + Ri[1] * y1) + Ri[2] * d1 + ti[0];
            const float py = (Ri[3] * x1 + Ri[4] * y1) + Ri[5] * d1 + ti[1];
            const float pz = (Ri[6] * x1 + Ri[7] * y1) + Ri[8] * d1 + ti[2];
            if (!(pz > 0.0f)) continue;
            const float inv_z = 1.0f / pz;
            const float fu = rintf(px * in->fx * inv_z + in->cx), fv = rintf(py * in->fy * inv_z + in->cy);
            if (!(fu >= 0.0f && fu < (float)W && fv >= 0.0f && fv < (float)H)) continue;
            const int i0 = (int)fv * W + (int)fu;
            const float d0 = d_prev[i0];
            if (!(d0 > 0.0f) || !(fabsf(pz - d0) <= max_depth_diff)) continue;
            // The nearest moved point wins; at equal depth the later pixel.
            if (corr[i0] >= 0 && pz > best[i0]) continue;
            if (corr[i0] < 0) ++pairs;
            corr[i0] = i1;
            best[i0] = pz;
        }
    return pairs;
}

int orc_rgbd_sums(const float* i_prev, const float* d_prev, const float* i_curr, const float* dx_curr,
                  const float* dy_curr, const kfx_intrinsics* in, const int32_t* corr, const double* R,
                  const double* t, double* sigma_out, double sums[27]) {
    const int W = in->cols, H = in->rows;
    int count = 0;
    double ss = 0.0;
    for (int i0 = 0; i0 < W * H; ++i0) {
        if (corr[i0] < 0) continue;
        const double diff = (double)(i_prev[i0] - i_curr[corr[i0]]);
        ss += diff * diff;
        ++count;
    }
    for (int k = 0; k < 27; ++k) sums[k] = 0.0;
    const double sigma = count > 0 ? sqrt(ss / (double)count) : 0.0;
    *sigma_out = sigma;
    for (int i0 = 0; i0 < W * H; ++i0) {
        const int i1 = corr[i0];
        if (i1 < 0) continue;
        const double diff = (double)(i_prev[i0] - i_curr[i1]);
        double w = sigma + fabs(diff);
        w = w > 2.220446049250313e-16 ? 1.0 / w : 1.0;
        const double ws = w / 8.0;
        const int u0 = i0 % W, v0 = i0 / W;
        const float d0 = d_prev[i0];
        const float x0 = ((float)u0 - in->cx) / in->fx * d0, y0 = ((float)v0 - in->cy) / in->fy * d0;
        const float X = (float)(R[0] * x0 + R[1] * y0 + R[2] * d0 + t[0]);
        const float Y = (float)(R[3] * x0 + R[4] * y0 + R[5] * d0 + t[1]);
        const float Z = (float)(R[6] * x0 + R[7] * y0 + R[8] * d0 + t[2]);
        const double inv_z = 1.0 / (double)Z;
        const double c3 = ws * (double)dx_curr[i1] * (double)in->fx * inv_z;
        const double c4 = ws * (double)dy_curr[i1] * (double)in->fy * inv_z;
        const double c5 = -(c3 * (double)X + c4 * (double)Y) * inv_z;
        const double C[6] = {-(double)Z * c4 + (double)Y * c5, (double)Z * c3 - (double)X * c5,
                             -(double)Y * c3 + (double)X * c4, c3, c4, c5};
        int k = 0;
        for (int a = 0; a < 6; ++a)
            for (int c = a; c < 6; ++c) sums[k++] += C[a] * C[c];
        for (int a = 0; a < 6; ++a) sums[21 + a] += C[a] * w * diff;
    }
    return count;
}

void orc_rgbd_twist_exp(const double ksi[6], double R[9], double t[3]) {
    const double wx = ksi[0], wy = ksi[1], wz = ksi[2];
    const double a2 = wx * wx + wy * wy + wz * wz, a = sqrt(a2);
    double A, B, Cc;
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
    double K2[9], V[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) K2[3 * r + c] = K[3 * r] * K[c] + K[3 * r + 1] * K[3 + c] + K[3 * r + 2] * K[6 + c];
    for (int i = 0; i < 9; ++i) {
        const double I = (i % 4 == 0) ? 1.0 : 0.0;
        R[i] = I + A * K[i] + B * K2[i];
        V[i] = I + B * K[i] + Cc * K2[i];
    }
    for (int r = 0; r < 3; ++r) t[r] = V[3 * r] * ksi[3] + V[3 * r + 1] * ksi[4] + V[3 * r + 2] * ksi[5];
}

int orc_rgbd_estimate(const kfx_color* rgb_prev, const uint16_t* depth_prev, const kfx_color* rgb_curr,
                      const uint16_t* depth_curr, const kfx_intrinsics* intr, int levels, const int* iterations,
                      float min_gradient, float max_depth_diff, const double* R0, const double* t0, double* Rout,
                      double* tout, int* solves_out) {
    const int W = intr->cols, H = intr->rows;
    size_t total = 0;
    for (int l = 0; l < levels; ++l) total += 4 * (size_t)(W >> l) * (H >> l);
    std::vector<float> fp(total), fc(total);
    orc_rgbd_prepare(rgb_prev, depth_prev, W, H, levels, fp.data());
    orc_rgbd_prepare(rgb_curr, depth_curr, W, H, levels, fc.data());
    double R[9], t[3];
    memcpy(R, R0, sizeof(R));
    memcpy(t, t0, sizeof(t));
    int ok = 0, solves = 0;
    std::vector<int32_t> corr((size_t)W * H);
    for (int l = levels - 1; l >= 0; --l) {
        const Level P = level_of(fp.data(), W, H, l), Cu = level_of(fc.data(), W, H, l);
        const float d = (float)(1 << l);
        const kfx_intrinsics li{intr->fx / d, intr->fy / d, intr->cx / d, intr->cy / d, P.cols, P.rows};
        for (int it = 0; it < iterations[l]; ++it) {
            float Ri[9], ti[3];
            for (int r = 0; r < 3; ++r) {
                for (int c = 0; c < 3; ++c) Ri[3 * r + c] = (float)R[3 * c + r];
                ti[r] = (float)(-(R[r] * t[0] + R[3 + r] * t[1] + R[6 + r] * t[2]));
            }
            const int pairs = orc_rgbd_corr(P.I, P.D, Cu.D, Cu.dx, Cu.dy, &li, Ri, ti, min_gradient * min_gradient,
                                            max_depth_diff, corr.data());
            if (pairs < 120) break;
            double sigma, s[27];
            orc_rgbd_sums(P.I, P.D, Cu.I, Cu.dx, Cu.dy, &li, corr.data(), R, t, &sigma, s);
            double A[36], b[6], ksi[6];
            int k = 0;
            for (int r = 0; r < 6; ++r)
                for (int c = r; c < 6; ++c, ++k) A[r * 6 + c] = A[c * 6 + r] = s[k];
            for (int r = 0; r < 6; ++r) b[r] = s[21 + r];
            if (!solve6(A, b, ksi)) break;
            double Rc[9], tc[3], Rn[9], tn[3];
            orc_rgbd_twist_exp(ksi, Rc, tc);
            for (int r = 0; r < 3; ++r) {
                for (int c = 0; c < 3; ++c)
                    Rn[3 * r + c] = Rc[3 * r] * R[c] + Rc[3 * r + 1] * R[3 + c] + Rc[3 * r + 2] * R[6 + c];
                tn[r] = Rc[3 * r] * t[0] + Rc[3 * r + 1] * t[1] + Rc[3 * r + 2] * t[2] + tc[r];
            }
            memcpy(R, Rn, sizeof(R));
            memcpy(t, tn, sizeof(t));
            ok = 1;
            ++solves;
        }
    }
    memcpy(Rout, R, sizeof(R));
    memcpy(tout, t, sizeof(t));
    *solves_out = solves;
    return ok;
}

}  // extern "C"
