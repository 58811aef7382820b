// kfx_synth.cpp — deterministic synthetic input: analytic scene (ground plane,
// boxes, spheres), camera trajectories and Kinect-like depth frames.
//
// Rays are intersected with the primitives in closed form (plane, slab test,
// sphere quadratic) in double precision, so one frame of 640x480 costs a few
// milliseconds of host time and every caller (tests, benchmark, denoiser
// training) sees bit-identical inputs for a given seed.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/kfx.h"

namespace {

inline uint64_t mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform double in [0, 1) from the top 53 bits.
inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    double next() { s = mix(s); return unit(s); }
    double uniform(double a, double b) { return a + (b - a) * next(); }
    double normal() {
        double u1 = next(), u2 = next();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

void look_rotation(const double f_in[3], double R[9]) {
    // Columns: x right = normalize(down x f), y = f x x, z = f; down = +y.
    double f[3] = {f_in[0], f_in[1], f_in[2]};
    double n = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= n;
    const double down[3] = {0.0, 1.0, 0.0};
    double x[3] = {down[1] * f[2] - down[2] * f[1], down[2] * f[0] - down[0] * f[2], down[0] * f[1] - down[1] * f[0]};
    n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (double& v : x) v /= n;
    double y[3] = {f[1] * x[2] - f[2] * x[1], f[2] * x[0] - f[0] * x[2], f[0] * x[1] - f[1] * x[0]};
    for (int r = 0; r < 3; ++r) {
        R[3 * r + 0] = x[r];
        R[3 * r + 1] = y[r];
        R[3 * r + 2] = f[r];
    }
}

// Smallest positive ray parameter s of o + s d against the scene (d has
// camera z component 1, so s is the camera z-depth).  0 = no hit.  *hit_id
// (optional) receives the primitive hit: the object index, -1 for the
// ground, -2 for none.
double intersect(const kfx_synth_scene& sc, const double o[3], const double d[3], int* hit_id = nullptr) {
    double best = INFINITY;
    int id = -2;
    if (d[1] > 1e-12) {
        const double s = (sc.ground_y - o[1]) / d[1];
        if (s > 0 && s < best) { best = s; id = -1; }
    }
    for (int i = 0; i < sc.n_objects; ++i) {
        const kfx_synth_object& ob = sc.objects[i];
        if (ob.shape == KFX_SYNTH_SPHERE) {
            const double oc[3] = {o[0] - ob.center[0], o[1] - ob.center[1], o[2] - ob.center[2]};
            const double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
            const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - ob.half[0] * ob.half[0];
            const double disc = b * b - a * c;
            if (disc < 0) continue;
            const double sq = sqrt(disc);
            double s = (-b - sq) / a;
            if (s <= 0) s = (-b + sq) / a;
            if (s > 0 && s < best) { best = s; id = i; }
        } else {
            double t0 = -INFINITY, t1 = INFINITY;
            bool miss = false;
            for (int k = 0; k < 3; ++k) {
                const double lo = ob.center[k] - ob.half[k], hi = ob.center[k] + ob.half[k];
                if (fabs(d[k]) < 1e-15) {
                    if (o[k] < lo || o[k] > hi) { miss = true; break; }
                    continue;
                }
                double a = (lo - o[k]) / d[k], b = (hi - o[k]) / d[k];
                if (a > b) std::swap(a, b);
                t0 = std::max(t0, a);
                t1 = std::min(t1, b);
            }
            if (miss || t0 > t1) continue;
            const double s = t0 > 0 ? t0 : t1;
            if (s > 0 && s < best) { best = s; id = i; }
        }
    }
    if (hit_id) *hit_id = std::isfinite(best) ? id : -2;
    return std::isfinite(best) ? best : 0.0;
}

// Albedo of a surface point: object i a fixed saturated colour (hue on the
// golden-ratio sequence), the ground a 0.5 m checkerboard of two greys,
// nothing black.  The w byte is 1 as the reference's capture sets it
// (common/src/Capture.cpp:119-122).
kfx_color albedo(int id, const double p[3]) {
    kfx_color c{0, 0, 0, 1};
    if (id == -1) {
        const long k = (long)floor(p[0] / 0.5) + (long)floor(p[2] / 0.5);
        const uint8_t g = (k & 1) ? 200 : 60;
        c.r = c.g = c.b = g;
    } else if (id >= 0) {
        const double h = fmod(0.6180339887498949 * (id + 1), 1.0) * 6.0, s = 0.75, v = 0.9;
        const int sector = (int)h;
        const double f = h - sector, pp = v * (1 - s), q = v * (1 - s * f), t = v * (1 - s * (1 - f));
        double rgb[3];
        switch (sector % 6) {
            case 0: rgb[0] = v; rgb[1] = t; rgb[2] = pp; break;
            case 1: rgb[0] = q; rgb[1] = v; rgb[2] = pp; break;
            case 2: rgb[0] = pp; rgb[1] = v; rgb[2] = t; break;
            case 3: rgb[0] = pp; rgb[1] = q; rgb[2] = v; break;
            case 4: rgb[0] = t; rgb[1] = pp; rgb[2] = v; break;
            default: rgb[0] = v; rgb[1] = pp; rgb[2] = q; break;
        }
        c.r = (uint8_t)lrint(255.0 * rgb[0]);
        c.g = (uint8_t)lrint(255.0 * rgb[1]);
        c.b = (uint8_t)lrint(255.0 * rgb[2]);
    }
    return c;
}

}  // namespace

extern "C" {

uint64_t kfx_splitmix64(uint64_t x) { return mix(x); }

int kfx_synth_make_scene(uint64_t seed, int n_boxes, int n_spheres, const double region[4], double ground_y,
                         kfx_synth_scene* out) {
    if (!out || !region || n_boxes < 0 || n_spheres < 0 || n_boxes + n_spheres > KFX_SYNTH_MAX_OBJECTS)
        return KFX_ERR_INVALID_ARG;
    memset(out, 0, sizeof(*out));
    out->ground_y = ground_y;
    Rng rng(mix(seed ^ 0x5CE4E5CE4Eull));
    int k = 0;
    for (int i = 0; i < n_boxes; ++i, ++k) {
        kfx_synth_object& ob = out->objects[k];
        ob.shape = KFX_SYNTH_BOX;
        for (int a = 0; a < 3; ++a) ob.half[a] = 0.5 * rng.uniform(0.1, 0.5);
        ob.center[0] = rng.uniform(region[0], region[1]);
        ob.center[2] = rng.uniform(region[2], region[3]);
        ob.center[1] = ground_y - ob.half[1];
    }
    for (int i = 0; i < n_spheres; ++i, ++k) {
        kfx_synth_object& ob = out->objects[k];
        ob.shape = KFX_SYNTH_SPHERE;
        ob.half[0] = rng.uniform(0.1, 0.5);
        ob.half[1] = ob.half[2] = ob.half[0];
        ob.center[0] = rng.uniform(region[0], region[1]);
        ob.center[2] = rng.uniform(region[2], region[3]);
        ob.center[1] = ground_y - ob.half[0];
    }
    out->n_objects = k;
    return KFX_OK;
}

int kfx_synth_orbit(const double center[3], double radius, double dy, double angle0, double dangle, int n,
                    kfx_pose* out) {
    if (!center || !out || n < 0) return KFX_ERR_INVALID_ARG;
    for (int i = 0; i < n; ++i) {
        const double a = angle0 + dangle * i;
        double* t = out[i].t;
        t[0] = center[0] + radius * sin(a);
        t[1] = center[1] + dy;
        t[2] = center[2] - radius * cos(a);
        const double f[3] = {center[0] - t[0], center[1] - t[1], center[2] - t[2]};
        look_rotation(f, out[i].R);
    }
    return KFX_OK;
}

int kfx_synth_walk(uint64_t seed, const double start[3], double step, double pitch, int n, kfx_pose* out) {
    if (!start || !out || n < 0) return KFX_ERR_INVALID_ARG;
    Rng rng(mix(seed ^ 0x3A1C3A1Cull));
    double p[3] = {start[0], start[1], start[2]};
    double heading = 0.0, rate = 0.0;
    for (int i = 0; i < n; ++i) {
        const double yaw = heading + 0.15 * sin(0.05 * i);
        const double pt = pitch + 0.05 * sin(0.031 * i);
        const double f[3] = {cos(yaw) * cos(pt), sin(pt), sin(yaw) * cos(pt)};
        look_rotation(f, out[i].R);
        out[i].t[0] = p[0];
        out[i].t[1] = p[1] + 0.03 * sin(0.07 * i);
        out[i].t[2] = p[2];
        rate = 0.9 * rate + 0.1 * 0.05 * rng.normal();
        heading = std::max(-0.3, std::min(0.3, heading + rate));
        p[0] += step * cos(heading);
        p[2] += step * sin(heading);
    }
    return KFX_OK;
}

int kfx_synth_render(const kfx_synth_scene* sc, const kfx_pose* pose, const kfx_intrinsics* in, uint64_t seed,
                     int frame, int flags, int threads, uint16_t* depth_mm, float* clean_m) {
    if (!sc || !pose || !in || !depth_mm || in->cols <= 0 || in->rows <= 0) return KFX_ERR_INVALID_ARG;
    const int W = in->cols, H = in->rows;
    const uint64_t fkey = mix(mix(seed) ^ mix((uint64_t)frame + 0x1000));
    int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, H);
    auto rows = [&](int y0, int y1) {
        for (int v = y0; v < y1; ++v)
            for (int u = 0; u < W; ++u) {
                const double dc[3] = {((double)u - in->cx) / in->fx, ((double)v - in->cy) / in->fy, 1.0};
                const double* R = pose->R;
                const double dw[3] = {R[0] * dc[0] + R[1] * dc[1] + R[2] * dc[2],
                                      R[3] * dc[0] + R[4] * dc[1] + R[5] * dc[2],
                                      R[6] * dc[0] + R[7] * dc[1] + R[8] * dc[2]};
                const double z = intersect(*sc, pose->t, dw);
                const int i = v * W + u;
                if (clean_m) clean_m[i] = (float)z;
                double zm = z;
                if (z > 0 && (flags & KFX_SYNTH_NOISE)) {
                    const uint64_t h = mix(fkey ^ mix((uint64_t)i));
                    Rng r(h);
                    const double sigma_mm = 1.2 + 1.9 * (z - 0.4) * (z - 0.4);
                    zm = z + 1e-3 * sigma_mm * r.normal();
                    if (r.next() < 0.02) zm = 0.0;
                }
                uint16_t d = 0;
                if (zm >= 0.4 && zm <= 6.0) d = (uint16_t)lrint(zm * 1000.0);
                depth_mm[i] = d;
            }
    };
    if (nt <= 1) {
        rows(0, H);
    } else {
        std::vector<std::thread> pool;
        for (int k = 0; k < nt; ++k) pool.emplace_back(rows, H * k / nt, H * (k + 1) / nt);
        for (auto& th : pool) th.join();
    }
    return KFX_OK;
}

int kfx_synth_render_rgb(const kfx_synth_scene* sc, const kfx_pose* pose, const kfx_intrinsics* in, int threads,
                         kfx_color* rgb) {
    if (!sc || !pose || !in || !rgb || in->cols <= 0 || in->rows <= 0) return KFX_ERR_INVALID_ARG;
    const int W = in->cols, H = in->rows;
    int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, H);
    auto rows = [&](int y0, int y1) {
        for (int v = y0; v < y1; ++v)
            for (int u = 0; u < W; ++u) {
                const double dc[3] = {((double)u - in->cx) / in->fx, ((double)v - in->cy) / in->fy, 1.0};
                const double* R = pose->R;
                const double dw[3] = {R[0] * dc[0] + R[1] * dc[1] + R[2] * dc[2],
                                      R[3] * dc[0] + R[4] * dc[1] + R[5] * dc[2],
                                      R[6] * dc[0] + R[7] * dc[1] + R[8] * dc[2]};
                int id = -2;
                const double z = intersect(*sc, pose->t, dw, &id);
                const double p[3] = {pose->t[0] + z * dw[0], pose->t[1] + z * dw[1], pose->t[2] + z * dw[2]};
                rgb[v * W + u] = albedo(id, p);
            }
    };
    if (nt <= 1) {
        rows(0, H);
    } else {
        std::vector<std::thread> pool;
        for (int k = 0; k < nt; ++k) pool.emplace_back(rows, H * k / nt, H * (k + 1) / nt);
        for (auto& th : pool) th.join();
    }
    return KFX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Trajectory file and error metrics.
// ---------------------------------------------------------------------------
extern "C" {

// Text format of the reference's write_isometry_to_file
// (common/src/string_utils.cpp:26-74): per pose, its index on one line, the
// three rows of R as "axbxc " (single precision, 6 significant digits), then
// t the same way.
int kfx_trajectory_write(const char* path, const kfx_pose* poses, int n) {
    if (!path || (n > 0 && !poses) || n < 0) return KFX_ERR_INVALID_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return KFX_ERR_INVALID_ARG;
    for (int i = 0; i < n; ++i) {
        fprintf(f, "%d\n", i);
        for (int r = 0; r < 3; ++r)
            fprintf(f, "%gx%gx%g %s\n", (float)poses[i].R[3 * r], (float)poses[i].R[3 * r + 1],
                    (float)poses[i].R[3 * r + 2], r == 2 ? " " : "");
        fprintf(f, "%gx%gx%g \n", (float)poses[i].t[0], (float)poses[i].t[1], (float)poses[i].t[2]);
    }
    return fclose(f) == 0 ? KFX_OK : KFX_ERR_INVALID_ARG;
}

// Reads a file in that format; returns the number of poses read (at most
// max_poses) in *n.  As the reference's stream operators do, the 'x'
// delimiters are replaced by blanks before parsing (a bare "0x..." would
// otherwise read as a hexadecimal number).
int kfx_trajectory_read(const char* path, kfx_pose* poses, int max_poses, int* n) {
    if (!path || !n || (max_poses > 0 && !poses)) return KFX_ERR_INVALID_ARG;
    FILE* f = fopen(path, "r");
    if (!f) return KFX_ERR_INVALID_ARG;
    int count = 0;
    char line[512];
    std::vector<double> vals;
    while (count < max_poses && fgets(line, sizeof(line), f)) {
        for (char* c = line; *c; ++c)
            if (*c == 'x') *c = ' ';
        char* p = line;
        char* end = nullptr;
        for (double v = strtod(p, &end); end != p; v = strtod(p, &end)) {
            vals.push_back(v);
            p = end;
        }
        if (vals.size() == 13) {  // index, 9 rotation entries, 3 translation entries
            for (int k = 0; k < 9; ++k) poses[count].R[k] = vals[1 + k];
            for (int k = 0; k < 3; ++k) poses[count].t[k] = vals[10 + k];
            ++count;
            vals.clear();
        }
    }
    fclose(f);
    *n = count;
    return KFX_OK;
}

// ATE: RMSE of camera positions (no alignment: both trajectories share the
// first frame's world frame).  RPE over consecutive frames: RMSE of the
// relative translation error (m) and mean relative rotation error (rad).
int kfx_trajectory_errors(const kfx_pose* est, const kfx_pose* gt, int n, double* ate, double* rpe_trans,
                          double* rpe_rot) {
    if (!est || !gt || n <= 0) return KFX_ERR_INVALID_ARG;
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const double d = est[i].t[k] - gt[i].t[k];
            s += d * d;
        }
    if (ate) *ate = sqrt(s / n);
    double st = 0.0, sr = 0.0;
    int m = 0;
    for (int i = 1; i < n; ++i) {
        // relative motions D = P_{i-1}^-1 P_i for both, error E = Dgt^-1 Dest
        double De[12], Dg[12];
        const kfx_pose* P[2][2] = {{&est[i - 1], &est[i]}, {&gt[i - 1], &gt[i]}};
        double* D[2] = {De, Dg};
        for (int w = 0; w < 2; ++w) {
            const kfx_pose &a = *P[w][0], &b = *P[w][1];
            for (int r = 0; r < 3; ++r) {
                for (int c = 0; c < 3; ++c)
                    D[w][3 * r + c] = a.R[r] * b.R[c] + a.R[3 + r] * b.R[3 + c] + a.R[6 + r] * b.R[6 + c];
                D[w][9 + r] = a.R[r] * (b.t[0] - a.t[0]) + a.R[3 + r] * (b.t[1] - a.t[1]) + a.R[6 + r] * (b.t[2] - a.t[2]);
            }
        }
        double tr = 0.0, et[3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) tr += Dg[3 * c + r] * De[3 * c + r];  // trace(Dg^T De)
            et[r] = Dg[r] * (De[9] - Dg[9]) + Dg[3 + r] * (De[10] - Dg[10]) + Dg[6 + r] * (De[11] - Dg[11]);
        }
        st += et[0] * et[0] + et[1] * et[1] + et[2] * et[2];
        sr += acos(std::max(-1.0, std::min(1.0, 0.5 * (tr - 1.0))));
        ++m;
    }
    if (rpe_trans) *rpe_trans = m ? sqrt(st / m) : 0.0;
    if (rpe_rot) *rpe_rot = m ? sr / m : 0.0;
    return KFX_OK;
}

}  // extern "C"
