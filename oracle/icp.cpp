// icp.cpp — oracle for the ICP normal equations, the reference's block
// reduction and the 6x6 solve / pose update.  Test infrastructure only (see
// oracle.h).
#include <math.h>
#include <string.h>

#include <algorithm>

#include "oracle.h"

namespace {

// Per-pixel correspondence and row in single precision, operation for
// operation as the GPU (KinectFusion estimate_combined convention; the
// reference kernel is absent, its argument contract is
// opencl_wrappers.cpp:386-398).
bool row_f32(const kfx_xyz* vc, const kfx_xyz* nc, const kfx_xyz* vp, const kfx_xyz* np, const float* Rc,
             const float* tc, const float* Rpi, const float* tp, const kfx_intrinsics* in, float dthr, float athr,
             int x, int y, float row[7]) {
    for (int k = 0; k < 7; ++k) row[k] = 0.0f;
    const int i = y * in->cols + x;
    const kfx_xyz n = nc[i];
    if (isnan(n.x)) return false;
    const kfx_xyz v = vc[i];
    if (isnan(v.x)) return false;
    const float gx = Rc[0] * v.x + Rc[1] * v.y + Rc[2] * v.z + tc[0];
    const float gy = Rc[3] * v.x + Rc[4] * v.y + Rc[5] * v.z + tc[1];
    const float gz = Rc[6] * v.x + Rc[7] * v.y + Rc[8] * v.z + tc[2];
    const float dx = gx - tp[0], dy = gy - tp[1], dz = gz - tp[2];
    const float px = Rpi[0] * dx + Rpi[1] * dy + Rpi[2] * dz;
    const float py = Rpi[3] * dx + Rpi[4] * dy + Rpi[5] * dz;
    const float pz = Rpi[6] * dx + Rpi[7] * dy + Rpi[8] * dz;
    // In front: pz in [2^-126, 2^126) (the kernels' exact reciprocal range).
    if (!(pz >= 0x1p-126f && pz < 0x1p126f)) return false;
    const float fu = rintf(px * in->fx / pz + in->cx);
    const float fv = rintf(py * in->fy / pz + in->cy);
    if (!(fu >= 0.0f && fv >= 0.0f && fu < (float)in->cols && fv < (float)in->rows)) return false;
    const int j = (int)fv * in->cols + (int)fu;
    const kfx_xyz q = np[j];
    if (isnan(q.x)) return false;
    const kfx_xyz w = vp[j];
    if (isnan(w.x)) return false;
    const float ex = w.x - gx, ey = w.y - gy, ez = w.z - gz;
    if (!(sqrtf(ex * ex + ey * ey + ez * ez) <= dthr)) return false;
    const float cnx = Rc[0] * n.x + Rc[1] * n.y + Rc[2] * n.z;
    const float cny = Rc[3] * n.x + Rc[4] * n.y + Rc[5] * n.z;
    const float cnz = Rc[6] * n.x + Rc[7] * n.y + Rc[8] * n.z;
    const float sx = cny * q.z - cnz * q.y, sy = cnz * q.x - cnx * q.z, sz = cnx * q.y - cny * q.x;
    if (!(sqrtf(sx * sx + sy * sy + sz * sz) < athr)) return false;
    row[0] = gy * q.z - gz * q.y;
    row[1] = gz * q.x - gx * q.z;
    row[2] = gx * q.y - gy * q.x;
    row[3] = q.x;
    row[4] = q.y;
    row[5] = q.z;
    row[6] = q.x * ex + q.y * ey + q.z * ez;
    return true;
}

// Same rule in double precision on double-precision poses.
bool row_f64(const kfx_xyz* vc, const kfx_xyz* nc, const kfx_xyz* vp, const kfx_xyz* np, const double* Rc,
             const double* tc, const double* Rpi, const double* tp, const kfx_intrinsics* in, double dthr,
             double athr, int x, int y, double row[7]) {
    for (int k = 0; k < 7; ++k) row[k] = 0.0;
    const int i = y * in->cols + x;
    const kfx_xyz n = nc[i], v = vc[i];
    if (isnan(n.x) || isnan(v.x)) return false;
    double g[3], p[3], d[3];
    for (int r = 0; r < 3; ++r) g[r] = Rc[3 * r] * v.x + Rc[3 * r + 1] * v.y + Rc[3 * r + 2] * v.z + tc[r];
    for (int r = 0; r < 3; ++r) d[r] = g[r] - tp[r];
    for (int r = 0; r < 3; ++r) p[r] = Rpi[3 * r] * d[0] + Rpi[3 * r + 1] * d[1] + Rpi[3 * r + 2] * d[2];
    if (!(p[2] > 0.0)) return false;
    const double fu = rint(p[0] * in->fx / p[2] + in->cx), fv = rint(p[1] * in->fy / p[2] + in->cy);
    if (!(fu >= 0 && fv >= 0 && fu < in->cols && fv < in->rows)) return false;
    const int j = (int)fv * in->cols + (int)fu;
    const kfx_xyz q = np[j], w = vp[j];
    if (isnan(q.x) || isnan(w.x)) return false;
    const double e[3] = {w.x - g[0], w.y - g[1], w.z - g[2]};
    if (!(sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) <= dthr)) return false;
    double cn[3];
    for (int r = 0; r < 3; ++r) cn[r] = Rc[3 * r] * n.x + Rc[3 * r + 1] * n.y + Rc[3 * r + 2] * n.z;
    const double s[3] = {cn[1] * q.z - cn[2] * q.y, cn[2] * q.x - cn[0] * q.z, cn[0] * q.y - cn[1] * q.x};
    if (!(sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) < athr)) return false;
    row[0] = g[1] * q.z - g[2] * q.y;
    row[1] = g[2] * q.x - g[0] * q.z;
    row[2] = g[0] * q.y - g[1] * q.x;
    row[3] = q.x;
    row[4] = q.y;
    row[5] = q.z;
    row[6] = q.x * e[0] + q.y * e[1] + q.z * e[2];
    return true;
}

}  // namespace

// Sums of the 27 products over all pixels.  precision 0: single-precision
// rows and products (as the GPU), accumulated exactly enough in double;
// precision 1: everything in double.  Returns the number of correspondences.
extern "C" int orc_icp_sums(const kfx_xyz* vc, const kfx_xyz* nc, const kfx_xyz* vp, const kfx_xyz* np,
                            const double* Rc, const double* tc, const double* Rprev_inv, const double* tprev,
                            const kfx_intrinsics* in, double dthr, double athr, int precision, double out[27]) {
    double acc[27];
    memset(acc, 0, sizeof(acc));
    int count = 0;
    float Rcf[9], tcf[3], Rpf[9], tpf[3];
    for (int i = 0; i < 9; ++i) { Rcf[i] = (float)Rc[i]; Rpf[i] = (float)Rprev_inv[i]; }
    for (int i = 0; i < 3; ++i) { tcf[i] = (float)tc[i]; tpf[i] = (float)tprev[i]; }
    for (int y = 0; y < in->rows; ++y)
        for (int x = 0; x < in->cols; ++x) {
            int k = 0;
            if (precision == 0) {
                float r[7];
                if (!row_f32(vc, nc, vp, np, Rcf, tcf, Rpf, tpf, in, (float)dthr, (float)athr, x, y, r)) continue;
                for (int i = 0; i < 6; ++i)
                    for (int j = i; j < 7; ++j, ++k) acc[k] += (double)(r[i] * r[j]);
            } else {
                double r[7];
                if (!row_f64(vc, nc, vp, np, Rc, tc, Rprev_inv, tprev, in, dthr, athr, x, y, r)) continue;
                for (int i = 0; i < 6; ++i)
                    for (int j = i; j < 7; ++j, ++k) acc[k] += r[i] * r[j];
            }
            ++count;
        }
    memcpy(out, acc, sizeof(acc));
    return count;
}

// The reference-path reduction, bit for bit: per 32x8 block, each wave of 64
// lanes (two image rows) sums its products with the xor butterfly
// (v[l] += v[l ^ off], off = 32..1), the four wave totals are added in order;
// then transform_estimation.cl:4-53 reduces each entry's partials (thread t
// sums t, t+256, ...; tree 128, 64, then lockstep 32..1).  mbuf receives 27
// floats.
extern "C" void orc_estimate_combined_f32(const kfx_xyz* vc, const kfx_xyz* nc, const kfx_xyz* vp,
                                          const kfx_xyz* np, const float* Rc, const float* tc, const float* Rpi,
                                          const float* tp, const kfx_intrinsics* in, float dthr, float athr,
                                          float mbuf[27]) {
    const int gx = (in->cols + 31) / 32, gy = (in->rows + 7) / 8, nb = gx * gy;
    float* gbuf = new float[27 * (size_t)nb];
    for (int by = 0; by < gy; ++by)
        for (int bx = 0; bx < gx; ++bx) {
            float wave_tot[4][27];
            for (int w = 0; w < 4; ++w) {
                float v[64][27];
                for (int l = 0; l < 64; ++l) {
                    const int tid = w * 64 + l, tx = tid % 32, ty = tid / 32;
                    const int x = bx * 32 + tx, y = by * 8 + ty;
                    float r[7] = {0, 0, 0, 0, 0, 0, 0};
                    if (x < in->cols && y < in->rows) row_f32(vc, nc, vp, np, Rc, tc, Rpi, tp, in, dthr, athr, x, y, r);
                    int k = 0;
                    for (int i = 0; i < 6; ++i)
                        for (int j = i; j < 7; ++j, ++k) v[l][k] = r[i] * r[j];
                }
                for (int off = 32; off > 0; off >>= 1) {
                    float nv[64][27];
                    for (int l = 0; l < 64; ++l)
                        for (int k = 0; k < 27; ++k) nv[l][k] = v[l][k] + v[l ^ off][k];
                    memcpy(v, nv, sizeof(v));
                }
                for (int k = 0; k < 27; ++k) wave_tot[w][k] = v[0][k];
            }
            for (int k = 0; k < 27; ++k) {
                float s = 0.0f;
                for (int w = 0; w < 4; ++w) s += wave_tot[w][k];
                gbuf[k * nb + by * gx + bx] = s;
            }
        }
    for (int k = 0; k < 27; ++k) {
        float a[256];
        for (int t = 0; t < 256; ++t) {
            float s = 0.0f;
            for (int id = t; id < nb; id += 256) s += gbuf[k * nb + id];
            a[t] = s;
        }
        for (int t = 0; t < 128; ++t) a[t] = a[t] + a[t + 128];
        for (int t = 0; t < 64; ++t) a[t] = a[t] + a[t + 64];
        for (int off = 32; off >= 1; off >>= 1) {
            float nv[32];
            for (int t = 0; t < 32; ++t) nv[t] = a[t] + a[t + off];
            memcpy(a, nv, sizeof(nv));
        }
        mbuf[k] = a[0];
    }
    delete[] gbuf;
}

// Cholesky solve of A x = b (A from the 27 sums, unpacked as
// opencl_wrappers.cpp:426-437) and the KinectFusion pose increment
// R_inc = Rz(x2) Ry(x1) Rx(x0), t <- R_inc t + x[3..5], R <- R_inc R, in
// double precision.  Returns 0, or 1 when A is not positive definite (pose
// unchanged).
extern "C" int orc_icp_solve_update(const double sums[27], double* R, double* t, double x_out[6]) {
    double A[36], b[6];
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j, ++k) {
            if (j == 6) b[i] = sums[k];
            else A[j * 6 + i] = A[i * 6 + j] = sums[k];
        }
    double L[36] = {0};
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = A[i * 6 + j];
            for (int p = 0; p < j; ++p) s -= L[i * 6 + p] * L[j * 6 + p];
            if (i == j) {
                if (!(s > 1e-300) || !std::isfinite(s)) {
                    for (int q = 0; q < 6; ++q) x_out[q] = 0.0;
                    return 1;
                }
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
    for (int i = 0; i < 6; ++i) x_out[i] = x[i];
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
    return 0;
}
