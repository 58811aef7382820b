// preprocess.cpp — oracle for the per-frame depth preprocessing: bilateral
// filter, depth pyramid, vertex/normal maps, map transform and resize.
// Test infrastructure only (see oracle.h).
#include <math.h>

#include <algorithm>
#include <cstdlib>

#include "oracle.h"

// compute_bilateral_filtering.cl:8-57 with the intended (normalised) output.
// Weights use a double-precision exponential rounded to single precision;
// sums are single precision in the kernel's loop order.
extern "C" void orc_bilateral(const uint16_t* src, uint16_t* dst, int cols, int rows, float sigma_space,
                              float sigma_color, uint32_t max_distance) {
    const float a = 0.5f / (sigma_space * sigma_space);
    const float b = 0.5f / (sigma_color * sigma_color);
    const int R = 6, D = 2 * R + 1;
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            const int value = src[r * cols + c];
            const int max_x = std::min(c - D / 2 + D, cols - 1), max_y = std::min(r - D / 2 + D, rows - 1);
            const int min_x = std::max(c - D / 2, 0), min_y = std::max(r - D / 2, 0);
            float sum1 = 0.0f, sum2 = 0.0f;
            for (int cy = min_y; cy < max_y; ++cy)
                for (int cx = min_x; cx < max_x; ++cx) {
                    const int tmp = src[cy * cols + cx];
                    const int space2 = (c - cx) * (c - cx) + (r - cy) * (r - cy);
                    const long long color2 = (long long)(value - tmp) * (value - tmp);
                    const float arg = (float)space2 * a + (float)color2 * b;
                    const float w = (float)exp(-(double)arg);
                    sum1 += (float)tmp * w;
                    sum2 += w;
                }
            int res = (int)rint(sum1 / sum2);
            res = std::max(0, std::min(res, 32767));
            if ((uint32_t)res > max_distance) res = 0;
            dst[r * cols + c] = (uint16_t)res;
        }
}

// KinectFusion pyrDown (the reference's compute_depth_pyramid is absent; call
// site opencl_wrappers.cpp:244-271).  Pinned by constant images, a step edge
// and the 3-sigma rejection (tests/test_oracle_preprocess.py).
extern "C" void orc_pyr_down(const uint16_t* src, int scols, int srows, uint16_t* dst, float sigma_color) {
    const int dc = scols / 2, dr = srows / 2;
    const float wt[3] = {0.375f, 0.25f, 0.0625f};
    for (int y = 0; y < dr; ++y)
        for (int x = 0; x < dc; ++x) {
            const int center = src[(2 * y) * scols + 2 * x];
            float sum = 0.0f, wall = 0.0f;
            for (int yi = std::max(0, 2 * y - 2) - 2 * y; yi < std::min(srows, 2 * y + 3) - 2 * y; ++yi)
                for (int xi = std::max(0, 2 * x - 2) - 2 * x; xi < std::min(scols, 2 * x + 3) - 2 * x; ++xi) {
                    const int val = src[(2 * y + yi) * scols + 2 * x + xi];
                    if ((float)std::abs(val - center) < 3.0f * sigma_color) {
                        const float w = wt[std::abs(xi)] * wt[std::abs(yi)];
                        sum += (float)val * w;
                        wall += w;
                    }
                }
            dst[y * dc + x] = (uint16_t)(int)(sum / wall);
        }
}

// KinectFusion computeVmap (compute_vertex_map absent; call site
// opencl_wrappers.cpp:281-305).
extern "C" void orc_vertex_map(const uint16_t* depth, kfx_xyz* vmap, const kfx_intrinsics* in) {
    const float fxi = 1.0f / in->fx, fyi = 1.0f / in->fy;
    for (int v = 0; v < in->rows; ++v)
        for (int u = 0; u < in->cols; ++u) {
            const int i = v * in->cols + u;
            const float z = (float)depth[i] / 1000.0f;
            if (z != 0.0f)
                vmap[i] = kfx_xyz{z * ((float)u - in->cx) * fxi, z * ((float)v - in->cy) * fyi, z};
            else
                vmap[i] = kfx_xyz{NAN, NAN, NAN};
        }
}

// Cyclic Jacobi eigenvalue iteration for a symmetric 3x3 matrix in double
// precision (Golub & Van Loan, Matrix Computations, 4th ed., sec. 8.5);
// returns the eigenvector of the smallest eigenvalue.
static void jacobi_smallest(double A[3][3], double out[3]) {
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (fabs(A[p][q]) < 1e-300) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {  // A <- J^T A J
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int m = 0;
    for (int k = 1; k < 3; ++k)
        if (A[k][k] < A[m][m]) m = k;
    const double n = sqrt(V[0][m] * V[0][m] + V[1][m] * V[1][m] + V[2][m] * V[2][m]);
    for (int k = 0; k < 3; ++k) out[k] = V[k][m] / n;
}

// KinectFusion computeNmapEigen (compute_normal_map_eigen absent; call site
// opencl_wrappers.cpp:315-343), in double precision: window [u-3, min(u+4,
// cols-1)) x [v-3, min(v+4, rows-1)), at least 24 valid vertices, smallest
// eigenvector of the covariance about the centroid, oriented toward the
// camera.
extern "C" void orc_normal_map_f64(const kfx_xyz* vmap, kfx_xyz* nmap, int cols, int rows) {
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (int v = 0; v < rows; ++v)
        for (int u = 0; u < cols; ++u) {
            const int i = v * cols + u;
            nmap[i] = kfx_xyz{NAN, NAN, NAN};
            if (isnan(vmap[i].x)) continue;
            const int ty = std::min(v + 4, rows - 1), tx = std::min(u + 4, cols - 1);
            double c[3] = {0, 0, 0};
            int count = 0;
            for (int cy = std::max(v - 3, 0); cy < ty; ++cy)
                for (int cx = std::max(u - 3, 0); cx < tx; ++cx) {
                    const kfx_xyz p = vmap[cy * cols + cx];
                    if (isnan(p.x)) continue;
                    c[0] += p.x; c[1] += p.y; c[2] += p.z;
                    ++count;
                }
            if (count < 24) continue;
            for (double& k : c) k /= count;
            double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int cy = std::max(v - 3, 0); cy < ty; ++cy)
                for (int cx = std::max(u - 3, 0); cx < tx; ++cx) {
                    const kfx_xyz p = vmap[cy * cols + cx];
                    if (isnan(p.x)) continue;
                    const double d[3] = {p.x - c[0], p.y - c[1], p.z - c[2]};
                    for (int r = 0; r < 3; ++r)
                        for (int s = 0; s < 3; ++s) A[r][s] += d[r] * d[s];
                }
            double n[3];
            jacobi_smallest(A, n);
            const kfx_xyz p = vmap[i];
            if (n[0] * p.x + n[1] * p.y + n[2] * p.z > 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
            nmap[i] = kfx_xyz{(float)n[0], (float)n[1], (float)n[2]};
        }
}

// transform_map.cl:6-65.
extern "C" void orc_transform_maps(const kfx_xyz* vs, const kfx_xyz* ns, const kfx_matrix_3f* R,
                                   const kfx_point_3f* t, int cols, int rows, kfx_xyz* vd, kfx_xyz* nd) {
    auto mul = [&](kfx_xyz a) {
        kfx_xyz o;
        float s;
        s = 0.0f; s = s + R->data[0].x * a.x; s = s + R->data[0].y * a.y; s = s + R->data[0].z * a.z; o.x = s;
        s = 0.0f; s = s + R->data[1].x * a.x; s = s + R->data[1].y * a.y; s = s + R->data[1].z * a.z; o.y = s;
        s = 0.0f; s = s + R->data[2].x * a.x; s = s + R->data[2].y * a.y; s = s + R->data[2].z * a.z; o.z = s;
        return o;
    };
    for (int i = 0; i < cols * rows; ++i) {
        kfx_xyz vo{NAN, NAN, NAN}, no{NAN, NAN, NAN};
        if (!isnan(vs[i].x)) {
            vo = mul(vs[i]);
            vo.x = vo.x + t->x; vo.y = vo.y + t->y; vo.z = vo.z + t->z;
        }
        if (!isnan(ns[i].x)) no = mul(ns[i]);
        vd[i] = vo;
        nd[i] = no;
    }
}

// KinectFusion resizeMap (resize_vertex_map / resize_normal_map absent; call
// sites opencl_wrappers.cpp:97-146).
extern "C" void orc_resize_map(const kfx_xyz* in, int icols, int irows, kfx_xyz* out, int normalize) {
    const int oc = icols / 2, orows = irows / 2;
    for (int y = 0; y < orows; ++y)
        for (int x = 0; x < oc; ++x) {
            const kfx_xyz a = in[(2 * y) * icols + 2 * x], b = in[(2 * y) * icols + 2 * x + 1];
            const kfx_xyz c = in[(2 * y + 1) * icols + 2 * x], d = in[(2 * y + 1) * icols + 2 * x + 1];
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
            out[y * oc + x] = o;
        }
}
