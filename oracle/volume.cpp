// volume.cpp — oracle for the cyclic TSDF volume: addressing, codec,
// scale_depth, integration and raycast.  Test infrastructure only (see
// oracle.h).
#include <math.h>
#include <string.h>

#include <algorithm>

#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

extern "C" void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
extern "C" int orc_get_threads(void) { return g_threads; }

// raycast.cl:29-47.  int arithmetic for the per-axis terms, size_t sum,
// start_index added, folded with '%' only when strictly greater than
// buffer_length.
extern "C" uint32_t orc_cyclic_address(int x, int y, int z, const kfx_round_buffer* rb) {
    const int px = (x + rb->origin_GRID.x) % rb->voxels_size.x;
    const int py = (y + rb->origin_GRID.y) % rb->voxels_size.y;
    const int pz = (z + rb->origin_GRID.z) % rb->voxels_size.z;
    size_t a = (size_t)(rb->voxels_size.x * rb->voxels_size.y * pz) + (size_t)(rb->voxels_size.x * py) + (size_t)px;
    a += (size_t)rb->start_index;
    if (a > (size_t)rb->buffer_length) a %= (size_t)rb->buffer_length;
    return (uint32_t)a;
}

// raycast.cl:9-16: (float)value / (float)MAX_DEPTH_VALUE.
extern "C" float orc_tsdf_decode(int16_t v) { return (float)v / 32767.0f; }

// raycast.cl:18-27: max(-32767, min(32767, convert_int_rtz(tsdf * 32767))).
extern "C" int16_t orc_tsdf_encode(float tsdf) {
    float s = tsdf * 32767.0f;
    double t = trunc((double)s);
    if (t > 32767.0) t = 32767.0;
    if (t < -32767.0) t = -32767.0;
    return (int16_t)t;
}

static int16_t encode_f64(double tsdf) {
    double t = trunc(tsdf * 32767.0);
    if (t > 32767.0) t = 32767.0;
    if (t < -32767.0) t = -32767.0;
    return (int16_t)t;
}

extern "C" void orc_scale_depth_f32(const uint16_t* depth, float* scaled, const kfx_intrinsics* in) {
    for (int v = 0; v < in->rows; ++v)
        for (int u = 0; u < in->cols; ++u) {
            const float xl = ((float)u - in->cx) / in->fx;
            const float yl = ((float)v - in->cy) / in->fy;
            const float lambda = sqrtf(xl * xl + yl * yl + 1.0f);
            const int i = v * in->cols + u;
            scaled[i] = (float)depth[i] * lambda / 1000.0f;
        }
}

extern "C" void orc_scale_depth_f64(const uint16_t* depth, double* scaled, const kfx_intrinsics* in) {
    for (int v = 0; v < in->rows; ++v)
        for (int u = 0; u < in->cols; ++u) {
            const double xl = ((double)u - in->cx) / in->fx;
            const double yl = ((double)v - in->cy) / in->fy;
            const int i = v * in->cols + u;
            scaled[i] = (double)depth[i] * sqrt(xl * xl + yl * yl + 1.0) / 1000.0;
        }
}

// Projective TSDF update, KinectFusion form (the reference's call site
// opencl_wrappers.cpp:47-61 passes depth_scaled, volume, trunc, R^-1, t,
// intrinsics, cell size, round_buffer over an (Nx, Ny) launch):
//   g   = voxel centre - t;  c = R^-1 g;  skip unless c.z > 0
//   (u, v) = rint(f * c.xy / c.z + c0);  skip unless inside the image
//   D = scaled(u, v);  skip if D == 0
//   sdf = D - |g|;  skip unless sdf >= -trunc
//   tsdf = min(1, sdf / trunc);  running average with weight +1 (cap 128).
// Operation order matches the single-precision kernel term for term.
extern "C" uint64_t orc_integrate_f32(const float* scaled, kfx_tsdf* vol, const kfx_intrinsics* in,
                                      const kfx_matrix_3f* Ri, const kfx_point_3f* t, float trunc,
                                      const kfx_point_3f* cs, const kfx_round_buffer* rb) {
    const int nx = rb->voxels_size.x, ny = rb->voxels_size.y, nz = rb->voxels_size.z;
    const float r[9] = {Ri->data[0].x, Ri->data[0].y, Ri->data[0].z, Ri->data[1].x, Ri->data[1].y,
                        Ri->data[1].z, Ri->data[2].x, Ri->data[2].y, Ri->data[2].z};
    const float inv_trunc = 1.0f / trunc;
    uint64_t touched = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 64) reduction(+ : touched) num_threads(g_threads)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const float gx = ((float)x + 0.5f) * cs->x - t->x;
            const float gy = ((float)y + 0.5f) * cs->y - t->y;
            const float part = gx * gx + gy * gy;
            for (int z = 0; z < nz; ++z) {
                const float gz = ((float)z + 0.5f) * cs->z - t->z;
                const float camx = r[0] * gx + r[1] * gy + r[2] * gz;
                const float camy = r[3] * gx + r[4] * gy + r[5] * gz;
                const float camz = r[6] * gx + r[7] * gy + r[8] * gz;
                // In front: camz in [2^-126, 2^126) (the kernels' exact
                // reciprocal range; wider than any volume).
                if (!(camz >= 0x1p-126f && camz < 0x1p126f)) continue;
                const float inv_z = 1.0f / camz;
                const float fu = rintf(camx * in->fx * inv_z + in->cx);
                const float fv = rintf(camy * in->fy * inv_z + in->cy);
                if (!(fu >= 0.0f && fu < (float)in->cols && fv >= 0.0f && fv < (float)in->rows)) continue;
                const float dp = scaled[(int)fv * in->cols + (int)fu];
                if (dp == 0.0f) continue;
                const float sdf = dp - sqrtf(gz * gz + part);
                if (!(sdf >= -trunc)) continue;
                const float tsdf = std::min(1.0f, sdf * inv_trunc);
                kfx_tsdf& cell = vol[orc_cyclic_address(x, y, z, rb)];
                const float tp = orc_tsdf_decode(cell.value);
                // A stored weight outside [0, 128] (a volume not written by
                // this rule) counts as the nearest of 0 and 128.
                const int w = std::max(0, std::min((int)cell.weight, 128));
                const float avg = (tp * (float)w + tsdf) / (float)(w + 1);
                cell.value = orc_tsdf_encode(avg);
                cell.weight = (int16_t)std::min(w + 1, 128);
                ++touched;
            }
        }
    return touched;
}

extern "C" uint64_t orc_integrate_f64(const double* scaled, kfx_tsdf* vol, const kfx_intrinsics* in,
                                      const double* r, const double* t, double trunc, const double* cs,
                                      const kfx_round_buffer* rb) {
    const int nx = rb->voxels_size.x, ny = rb->voxels_size.y, nz = rb->voxels_size.z;
    uint64_t touched = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 64) reduction(+ : touched) num_threads(g_threads)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const double gx = ((double)x + 0.5) * cs[0] - t[0];
            const double gy = ((double)y + 0.5) * cs[1] - t[1];
            for (int z = 0; z < nz; ++z) {
                const double gz = ((double)z + 0.5) * cs[2] - t[2];
                const double camx = r[0] * gx + r[1] * gy + r[2] * gz;
                const double camy = r[3] * gx + r[4] * gy + r[5] * gz;
                const double camz = r[6] * gx + r[7] * gy + r[8] * gz;
                if (!(camz > 0.0)) continue;
                // More than a pixel outside the image on either axis: skipped
                // before the divisions (the test below rejects it as well).
                const double ux = camx * in->fx, uy = camy * in->fy;
                if (ux < -(in->cx + 1.0) * camz || ux > ((double)in->cols - in->cx + 1.0) * camz ||
                    uy < -(in->cy + 1.0) * camz || uy > ((double)in->rows - in->cy + 1.0) * camz)
                    continue;
                const double fu = rint(camx * in->fx / camz + in->cx);
                const double fv = rint(camy * in->fy / camz + in->cy);
                if (!(fu >= 0.0 && fu < (double)in->cols && fv >= 0.0 && fv < (double)in->rows)) continue;
                const double dp = scaled[(int)fv * in->cols + (int)fu];
                if (dp == 0.0) continue;
                const double sdf = dp - sqrt(gx * gx + gy * gy + gz * gz);
                if (!(sdf >= -trunc)) continue;
                const double tsdf = std::min(1.0, sdf / trunc);
                kfx_tsdf& cell = vol[orc_cyclic_address(x, y, z, rb)];
                const double tp = (double)cell.value / 32767.0;
                const int w = std::max(0, std::min((int)cell.weight, 128));
                cell.value = encode_f64((tp * w + tsdf) / (w + 1));
                cell.weight = (int16_t)std::min(w + 1, 128);
                ++touched;
            }
        }
    return touched;
}

// Colour volume update (Integrate_Colors, call site
// opencl_wrappers.cpp:181-214; kernel integrate_color_volume absent).  The
// published KinectFusion colour-volume rule (PCL gpu/kinfu,
// color_volume.cu, updateColorVolume) with the projection of the TSDF
// update above:
//   g = (i + 1/2) * cell;  c = R^-1 (g - t);  skip unless c.z > 0
//   (u, v) = rint(f * c.xy / c.z + c0);  skip unless inside the image
//   p = vmap(u, v);  skip if p.x is NaN
//   skip unless |p - g| < trunc
//   per channel: rint((rgb * w + rgb_frame) / (w + 1)) in FP32, clamped to
//   [0, 255];  w = min(w + 1, max_weight).
// Returns the number of voxels the rule touched.
extern "C" uint64_t orc_integrate_colors(kfx_color* vol, const kfx_color* rgb, const kfx_xyz* vmap,
                                         const kfx_intrinsics* in, const kfx_matrix_3f* Ri, const kfx_point_3f* t,
                                         float trunc, int max_weight, const kfx_point_3f* cs,
                                         const kfx_round_buffer* rb) {
    const int nx = rb->voxels_size.x, ny = rb->voxels_size.y, nz = rb->voxels_size.z;
    const float r[9] = {Ri->data[0].x, Ri->data[0].y, Ri->data[0].z, Ri->data[1].x, Ri->data[1].y,
                        Ri->data[1].z, Ri->data[2].x, Ri->data[2].y, Ri->data[2].z};
    uint64_t touched = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 64) reduction(+ : touched) num_threads(g_threads)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            for (int z = 0; z < nz; ++z) {
                const float vx = ((float)x + 0.5f) * cs->x, vy = ((float)y + 0.5f) * cs->y,
                            vz = ((float)z + 0.5f) * cs->z;
                const float gx = vx - t->x, gy = vy - t->y, gz = vz - t->z;
                const float camx = r[0] * gx + r[1] * gy + r[2] * gz;
                const float camy = r[3] * gx + r[4] * gy + r[5] * gz;
                const float camz = r[6] * gx + r[7] * gy + r[8] * gz;
                // In front: camz in [2^-126, 2^126) (the kernels' exact
                // reciprocal range; wider than any volume).
                if (!(camz >= 0x1p-126f && camz < 0x1p126f)) continue;
                const float inv_z = 1.0f / camz;
                const float fu = rintf(camx * in->fx * inv_z + in->cx);
                const float fv = rintf(camy * in->fy * inv_z + in->cy);
                if (!(fu >= 0.0f && fu < (float)in->cols && fv >= 0.0f && fv < (float)in->rows)) continue;
                const int pix = (int)fv * in->cols + (int)fu;
                const kfx_xyz p = vmap[pix];
                if (isnan(p.x)) continue;
                const float dx = p.x - vx, dy = p.y - vy, dz = p.z - vz;
                if (!(sqrtf(dx * dx + dy * dy + dz * dz) < trunc)) continue;
                kfx_color& cell = vol[orc_cyclic_address(x, y, z, rb)];
                const kfx_color c = rgb[pix];
                const int w = cell.w;
                const float den = (float)(w + 1);
                const float nr = ((float)(cell.r * w) + 1.0f * (float)c.r) / den;
                const float ng = ((float)(cell.g * w) + 1.0f * (float)c.g) / den;
                const float nb = ((float)(cell.b * w) + 1.0f * (float)c.b) / den;
                cell.r = (uint8_t)std::min(255, std::max(0, (int)rintf(nr)));
                cell.g = (uint8_t)std::min(255, std::max(0, (int)rintf(ng)));
                cell.b = (uint8_t)std::min(255, std::max(0, (int)rintf(nb)));
                cell.w = (uint8_t)std::min(max_weight, w + 1);
                ++touched;
            }
        }
    return touched;
}

// ---------------------------------------------------------------------------
// Raycast, single precision in the order of raycast.cl.
// ---------------------------------------------------------------------------
namespace {

struct V3 { float x, y, z; };

inline float dot3(V3 a, V3 b) {
    float r = 0.0f;  // elementary_point_template.h:135-143 accumulates from 0
    r = r + a.x * b.x;
    r = r + a.y * b.y;
    r = r + a.z * b.z;
    return r;
}
inline V3 normalized3(V3 a) {  // elementary_point_template.h:166-177
    const float rs = 1.0f / sqrtf(dot3(a, a));
    return V3{a.x * rs, a.y * rs, a.z * rs};
}
// convert_int_sat_rtn(p / cell): raycast.cl:68-77.
inline int vox_floor(float p, float c) {
    double q = floor((double)(p / c));
    if (q != q) return 0;
    if (q < -2147483648.0) q = -2147483648.0;
    if (q > 2147483647.0) q = 2147483647.0;
    return (int)q;
}
inline float read_t(const kfx_tsdf* vol, const kfx_round_buffer* rb, int x, int y, int z) {
    return orc_tsdf_decode(vol[orc_cyclic_address(x, y, z, rb)].value);
}
inline V3 at(V3 o, V3 d, float t) {
    V3 s{d.x * t, d.y * t, d.z * t};
    return V3{o.x + s.x, o.y + s.y, o.z + s.z};
}

}  // namespace

extern "C" float orc_interpolate_trilinear_f32(const kfx_tsdf* vol, float px, float py, float pz,
                                               const kfx_point_3f* cs, const kfx_round_buffer* rb) {
    int gx = vox_floor(px, cs->x), gy = vox_floor(py, cs->y), gz = vox_floor(pz, cs->z);
    if (gx <= 0 || gx >= rb->voxels_size.x - 1) return NAN;
    if (gy <= 0 || gy >= rb->voxels_size.y - 1) return NAN;
    if (gz <= 0 || gz >= rb->voxels_size.z - 1) return NAN;
    const float vx = ((float)gx + 0.5f) * cs->x;
    const float vy = ((float)gy + 0.5f) * cs->y;
    const float vz = ((float)gz + 0.5f) * cs->z;
    gx = (px < vx) ? gx - 1 : gx;
    gy = (py < vy) ? gy - 1 : gy;
    gz = (pz < vz) ? gz - 1 : gz;
    float a = px / cs->x - ((float)gx + 0.5f);
    if (a < 0) { gx--; a += 1.0f; }
    float b = py / cs->y - ((float)gy + 0.5f);
    if (b < 0) { gy--; b += 1.0f; }
    float c = pz / cs->z - ((float)gz + 0.5f);
    if (c < 0) { gz--; c += 1.0f; }
    return (1 - a) * ((1 - b) * (read_t(vol, rb, gx + 0, gy + 0, gz + 0) * (1 - c) +
                                 read_t(vol, rb, gx + 0, gy + 0, gz + 1) * c) +
                      b * (read_t(vol, rb, gx + 0, gy + 1, gz + 0) * (1 - c) +
                           read_t(vol, rb, gx + 0, gy + 1, gz + 1) * c)) +
           a * ((1 - b) * (read_t(vol, rb, gx + 1, gy + 0, gz + 0) * (1 - c) +
                           read_t(vol, rb, gx + 1, gy + 0, gz + 1) * c) +
                b * (read_t(vol, rb, gx + 1, gy + 1, gz + 0) * (1 - c) +
                     read_t(vol, rb, gx + 1, gy + 1, gz + 1) * c));
}

// Returns S, the TSDF voxel reads the algorithm makes: the first sample, every
// march sample inside the volume, and 8 per trilinear evaluation that passes
// its bounds test.
extern "C" uint64_t orc_raycast_f32(const kfx_tsdf* vol, kfx_xyz* vmap, kfx_xyz* nmap, const kfx_round_buffer* rb,
                                    const kfx_point_3f* T, const kfx_matrix_3f* R, float trunc,
                                    const kfx_intrinsics* in, const kfx_point_3f* cs) {
    const int W = in->cols, H = in->rows;
    const V3 vm{rb->volume_size.x, rb->volume_size.y, rb->volume_size.z};
    const int NX = rb->voxels_size.x, NY = rb->voxels_size.y, NZ = rb->voxels_size.z;
    const float step = (float)((double)trunc * 0.8);  // raycast.cl:209
    uint64_t samples = 0;
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads) reduction(+ : samples)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int idx = W * y + x;
            vmap[idx] = kfx_xyz{NAN, NAN, NAN};
            nmap[idx] = kfx_xyz{NAN, NAN, NAN};
            const V3 start{T->x, T->y, T->z};
            const V3 rn{((float)x - in->cx) / in->fx, ((float)y - in->cy) / in->fy, 1.0f};
            const V3 rr{dot3(V3{R->data[0].x, R->data[0].y, R->data[0].z}, rn),
                        dot3(V3{R->data[1].x, R->data[1].y, R->data[1].z}, rn),
                        dot3(V3{R->data[2].x, R->data[2].y, R->data[2].z}, rn)};
            const V3 nx{rr.x + start.x, rr.y + start.y, rr.z + start.z};
            V3 d = normalized3(V3{nx.x - start.x, nx.y - start.y, nx.z - start.z});
            d.x = d.x == 0.0f ? 1e-15f : d.x;
            d.y = d.y == 0.0f ? 1e-15f : d.y;
            d.z = d.z == 0.0f ? 1e-15f : d.z;
            float tmin = std::max(std::max(((d.x > 0 ? 0.f : vm.x) - start.x) / d.x,
                                           ((d.y > 0 ? 0.f : vm.y) - start.y) / d.y),
                                  ((d.z > 0 ? 0.f : vm.z) - start.z) / d.z);
            const float tmax = std::min(std::min(((d.x > 0 ? vm.x : 0.f) - start.x) / d.x,
                                                 ((d.y > 0 ? vm.y : 0.f) - start.y) / d.y),
                                        ((d.z > 0 ? vm.z : 0.f) - start.z) / d.z);
            tmin = std::max(tmin, 0.0f);
            if (tmin >= tmax) continue;
            float tc = tmin;
            V3 q = at(start, d, tc);
            int gx = std::max(0, std::min(vox_floor(q.x, cs->x), NX - 1));
            int gy = std::max(0, std::min(vox_floor(q.y, cs->y), NY - 1));
            int gz = std::max(0, std::min(vox_floor(q.z, cs->z), NZ - 1));
            float tsdf = read_t(vol, rb, gx, gy, gz);
            samples += 1;
            const float max_time = 3 * (vm.x + vm.y + vm.z);
            for (; tc < max_time; tc += step) {
                const float prev = tsdf;
                const float tn = tc + step;
                const V3 q1 = at(start, d, tn);
                const int vx = vox_floor(q1.x, cs->x), vy = vox_floor(q1.y, cs->y), vz = vox_floor(q1.z, cs->z);
                if (!(vx >= 0 && vy >= 0 && vz >= 0 && vx < NX && vy < NY && vz < NZ)) break;
                tsdf = read_t(vol, rb, vx, vy, vz);
                samples += 1;
                if (prev < 0.f && tsdf > 0.f) break;
                if (prev > 0.f && tsdf < 0.f) {
                    const float ftdt = orc_interpolate_trilinear_f32(vol, q1.x, q1.y, q1.z, cs, rb);
                    if (isnan(ftdt)) break;
                    samples += 8;
                    const V3 q2 = at(start, d, tc);
                    const float ft = orc_interpolate_trilinear_f32(vol, q2.x, q2.y, q2.z, cs, rb);
                    if (isnan(ft)) break;
                    samples += 8;
                    const float ts = tc - step * ft / (ftdt - ft);
                    const V3 v = at(start, d, ts);
                    vmap[idx] = kfx_xyz{v.x, v.y, v.z};
                    const V3 q3 = at(start, d, tc);
                    const int hx = vox_floor(q3.x, cs->x), hy = vox_floor(q3.y, cs->y), hz = vox_floor(q3.z, cs->z);
                    if (hx > 1 && hy > 1 && hz > 1 && hx < NX - 2 && hy < NY - 2 && hz < NZ - 2) {
                        const float fx1 = orc_interpolate_trilinear_f32(vol, v.x + cs->x, v.y, v.z, cs, rb);
                        const float fx2 = orc_interpolate_trilinear_f32(vol, v.x - cs->x, v.y, v.z, cs, rb);
                        const float fy1 = orc_interpolate_trilinear_f32(vol, v.x, v.y + cs->y, v.z, cs, rb);
                        const float fy2 = orc_interpolate_trilinear_f32(vol, v.x, v.y - cs->y, v.z, cs, rb);
                        const float fz1 = orc_interpolate_trilinear_f32(vol, v.x, v.y, v.z + cs->z, cs, rb);
                        const float fz2 = orc_interpolate_trilinear_f32(vol, v.x, v.y, v.z - cs->z, cs, rb);
                        samples += 8 * (uint64_t)(!isnan(fx1) + !isnan(fx2) + !isnan(fy1) + !isnan(fy2) +
                                                  !isnan(fz1) + !isnan(fz2));
                        const V3 n = normalized3(V3{fx1 - fx2, fy1 - fy2, fz1 - fz2});
                        nmap[idx] = kfx_xyz{n.x, n.y, n.z};
                    }
                    break;
                }
            }
        }
    return samples;
}
