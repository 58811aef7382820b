// kfx_raycast.hip — per-pixel ray march through the cyclic TSDF volume.
//
// Follows the reference kernel's algorithm (GPU/opencl/raycast.cl:182-405):
// the ray starts at the camera, is clipped to the volume box [0, volume_size],
// advances in steps of 0.8 * truncation sampling the nearest voxel, stops at
// the first + to - sign change, refines the crossing linearly from two
// trilinear samples, and takes the normal as the normalised central
// difference of six trilinear samples one cell apart.  Arithmetic is plain
// IEEE single precision in the reference's operation order (this file is
// built without floating-point contraction), so the CPU oracle restates it
// bit for bit.
//
// Layout: one lane per pixel, 64 x 4 pixel tiles per workgroup, so the 64
// rays of a wave are neighbours and their samples fall in the same cache
// lines of the volume.
#include <math.h>

#include "kfx_common.h"

namespace kfx {

struct RaycastParams {
    float r[9];
    float tx, ty, tz;
    float csx, csy, csz;
    float icsx, icsy, icsz;  // 1 / cell size, correctly rounded
    float fx, fy, cx, cy;
    int cols, rows;
    float time_step;
    kfx_round_buffer rb;
    // Brick flags for the skipping march (null: the full march): one byte per
    // kRayBrick^3 brick of local voxel coordinates, x fastest (nbx x nby x
    // nbz bricks), non-zero when the brick may hold a negative value.
    const uint8_t* bricks;
    int nbx, nby;
    // Levels 1 and 2 of the output maps, built by each workgroup from its
    // tile (null: not built): c1 x r1 = (cols / 2) x (rows / 2), c2 x r2
    // half of that again.
    kfx_xyz *v1, *n1, *v2, *n2;
    int c1, r1, c2, r2;
};

struct f3 { float x, y, z; };

__device__ __forceinline__ float dot3(f3 a, f3 b) {
    float r = 0.0f;  // the reference accumulates from 0 (elementary_point_template.h:135-143)
    r = r + a.x * b.x;
    r = r + a.y * b.y;
    r = r + a.z * b.z;
    return r;
}

__device__ __forceinline__ f3 normalized3(f3 a) {
    const float rs = 1.0f / sqrtf(dot3(a, a));
    return f3{a.x * rs, a.y * rs, a.z * rs};
}

// Every read below is at coordinates inside [0, N) on each axis: march
// samples pass the bounds test first, the first sample is clamped, and
// trilinear taps lie in [g-1, g+1] with g in [1, N-2].
__device__ __forceinline__ int16_t read_raw(const kfx_tsdf* __restrict__ vol, const kfx_round_buffer& rb,
                                            int x, int y, int z) {
    return vol[cyclic_address_in_range(x, y, z, rb)].value;
}

__device__ __forceinline__ float read_tsdf(const kfx_tsdf* __restrict__ vol, const kfx_round_buffer& rb,
                                           int x, int y, int z) {
    return tsdf_decode(read_raw(vol, rb, x, y, z));
}

// Samples of the march whose loads are issued together.
constexpr int kMarchBatch = 4;

// Linear index of in-range voxel coordinates after the cyclic offset: x + o
// is below 2N, so min(x + o, x + o - N) in unsigned arithmetic is the wrap.
__device__ __forceinline__ uint32_t march_address(uint32_t x, uint32_t y, uint32_t z, const kfx_round_buffer& rb) {
    const uint32_t nx = (uint32_t)rb.voxels_size.x, ny = (uint32_t)rb.voxels_size.y, nz = (uint32_t)rb.voxels_size.z;
    uint32_t rx = x + (uint32_t)rb.origin_GRID.x, ry = y + (uint32_t)rb.origin_GRID.y,
             rz = z + (uint32_t)rb.origin_GRID.z;
    rx = min(rx, rx - nx);
    ry = min(ry, ry - ny);
    rz = min(rz, rz - nz);
    return (rz * ny + ry) * nx + rx;
}

// p / cell, correctly rounded, without a division: q0 = p * (1/cell) is
// within an ulp of the quotient, the residual p - q0 * cell is exact by FMA,
// and one correction with the correctly rounded reciprocal gives the
// correctly rounded quotient (Markstein's theorem; P. Markstein, "Computation
// of elementary functions on the IBM RISC System/6000 processor", IBM J. Res.
// Dev. 34(1), 1990).  The result equals IEEE p / cell bit for bit for all
// operands away from underflow; near zero the sign, and so the floor below,
// is still exact.
__device__ __forceinline__ float div_cell(float p, float cell, float inv_cell) {
    const float q0 = p * inv_cell;
    const float r = fmaf(-q0, cell, p);
    return fmaf(r, inv_cell, q0);
}

// floor(p / cell) with saturation (convert_int_sat_rtn).
__device__ __forceinline__ int voxel_floor(float p, float cell, float inv_cell) {
    float q = floorf(div_cell(p, cell, inv_cell));
    if (q != q) return 0;
    q = fminf(fmaxf(q, -2147483648.0f), 2147483520.0f);
    return (int)q;
}

// interpolateTrilineary (raycast.cl:95-145): NaN outside [1, N-2].  In three
// parts, so that a caller can issue the loads of several evaluations before
// using any: tri_setup (the cell and the weights; ok false where the
// reference returns NaN), tri_load (the eight taps; address 0 when !ok) and
// tri_eval.
struct TriCell {
    bool ok;
    uint32_t x0, x1, r00, r01, r10, r11;  // tap addresses: r.. + x.
    float a, b, c;
};

__device__ __forceinline__ TriCell tri_setup(f3 pt, const RaycastParams& p) {
    const kfx_round_buffer& rb = p.rb;
    TriCell t;
    int gx = voxel_floor(pt.x, p.csx, p.icsx);
    int gy = voxel_floor(pt.y, p.csy, p.icsy);
    int gz = voxel_floor(pt.z, p.csz, p.icsz);
    t.ok = !(gx <= 0 || gx >= rb.voxels_size.x - 1 || gy <= 0 || gy >= rb.voxels_size.y - 1 || gz <= 0 ||
             gz >= rb.voxels_size.z - 1);
    const float vx = ((float)gx + 0.5f) * p.csx;
    const float vy = ((float)gy + 0.5f) * p.csy;
    const float vz = ((float)gz + 0.5f) * p.csz;
    gx = (pt.x < vx) ? (gx - 1) : gx;
    gy = (pt.y < vy) ? (gy - 1) : gy;
    gz = (pt.z < vz) ? (gz - 1) : gz;
    float a = div_cell(pt.x, p.csx, p.icsx) - ((float)gx + 0.5f);
    if (a < 0.0f) { gx--; a += 1.0f; }
    float b = div_cell(pt.y, p.csy, p.icsy) - ((float)gy + 0.5f);
    if (b < 0.0f) { gy--; b += 1.0f; }
    float c = div_cell(pt.z, p.csz, p.icsz) - ((float)gz + 0.5f);
    if (c < 0.0f) { gz--; c += 1.0f; }
    t.a = a; t.b = b; t.c = c;
    // The taps gx .. gx + 1 etc. lie inside [0, N) when ok: their wrapped
    // coordinates, with the wraps taken once per axis (cyclic_address).
    const uint32_t nx = (uint32_t)rb.voxels_size.x, ny = (uint32_t)rb.voxels_size.y;
    const uint32_t nz = (uint32_t)rb.voxels_size.z;
    uint32_t x0 = (uint32_t)gx + (uint32_t)rb.origin_GRID.x, y0 = (uint32_t)gy + (uint32_t)rb.origin_GRID.y;
    uint32_t z0 = (uint32_t)gz + (uint32_t)rb.origin_GRID.z;
    x0 = min(x0, x0 - nx);
    y0 = min(y0, y0 - ny);
    z0 = min(z0, z0 - nz);
    const uint32_t x1 = x0 + 1 == nx ? 0u : x0 + 1, y1 = y0 + 1 == ny ? 0u : y0 + 1, z1 = z0 + 1 == nz ? 0u : z0 + 1;
    t.x0 = t.ok ? x0 : 0u;
    t.x1 = t.ok ? x1 : 0u;
    t.r00 = t.ok ? (z0 * ny + y0) * nx : 0u;
    t.r01 = t.ok ? (z1 * ny + y0) * nx : 0u;
    t.r10 = t.ok ? (z0 * ny + y1) * nx : 0u;
    t.r11 = t.ok ? (z1 * ny + y1) * nx : 0u;
    return t;
}

// v = {000, 001, 010, 011, 100, 101, 110, 111} (x, y, z bits).
__device__ __forceinline__ void tri_load(const kfx_tsdf* __restrict__ vol, const TriCell& t, int16_t (&v)[8]) {
    v[0] = vol[t.r00 + t.x0].value;
    v[1] = vol[t.r01 + t.x0].value;
    v[2] = vol[t.r10 + t.x0].value;
    v[3] = vol[t.r11 + t.x0].value;
    v[4] = vol[t.r00 + t.x1].value;
    v[5] = vol[t.r01 + t.x1].value;
    v[6] = vol[t.r10 + t.x1].value;
    v[7] = vol[t.r11 + t.x1].value;
}

__device__ __forceinline__ float tri_eval(const TriCell& t, const int16_t (&v)[8]) {
    if (!t.ok) return NAN;
    const float a = t.a, b = t.b, c = t.c;
    const float t000 = tsdf_decode(v[0]), t001 = tsdf_decode(v[1]), t010 = tsdf_decode(v[2]);
    const float t011 = tsdf_decode(v[3]), t100 = tsdf_decode(v[4]), t101 = tsdf_decode(v[5]);
    const float t110 = tsdf_decode(v[6]), t111 = tsdf_decode(v[7]);
    return (1.0f - a) * ((1.0f - b) * (t000 * (1.0f - c) + t001 * c) + b * (t010 * (1.0f - c) + t011 * c)) +
           a * ((1.0f - b) * (t100 * (1.0f - c) + t101 * c) + b * (t110 * (1.0f - c) + t111 * c));
}

__device__ float interpolate_trilinear(const kfx_tsdf* __restrict__ vol, f3 pt, const RaycastParams& p) {
    const TriCell t = tri_setup(pt, p);
    if (!t.ok) return NAN;
    int16_t v[8];
    tri_load(vol, t, v);
    return tri_eval(t, v);
}

__device__ __forceinline__ f3 ray_at(f3 o, f3 d, float t) {
    return f3{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
}

// The refinement of a + to - crossing between march times t_cur and t_next
// (raycast.cl:291-331) and the normal (raycast.cl:332-404): vertex and normal
// outputs, unchanged (NaN) where a trilinear sample leaves [1, N-2]; with
// kCount, the reads made are added to ns.
template <bool kCount>
__device__ __forceinline__ void raycast_hit(const kfx_tsdf* __restrict__ vol, const RaycastParams& p, f3 start,
                                            f3 dir, float hit_tc, float hit_tn, kfx_xyz& vout, kfx_xyz& nout,
                                            uint32_t& ns) {
    const kfx_round_buffer& rb = p.rb;
    const float step = p.time_step;
    const int nxv = rb.voxels_size.x, nyv = rb.voxels_size.y, nzv = rb.voxels_size.z;
    const float t_cur = hit_tc;
    // The two samples of the crossing: both cells' taps are loaded together.
    const TriCell c1 = tri_setup(ray_at(start, dir, hit_tn), p);
    const TriCell c2 = tri_setup(ray_at(start, dir, t_cur), p);
    int16_t v1[8], v2[8];
    tri_load(vol, c1, v1);
    tri_load(vol, c2, v2);
    const float ftdt = tri_eval(c1, v1);
    if (isnan(ftdt)) return;
    if (kCount) ns += 8;
    const float ft = tri_eval(c2, v2);
    if (isnan(ft)) return;
    if (kCount) ns += 8;
    const float ts = t_cur - step * ft / (ftdt - ft);
    const f3 vert = ray_at(start, dir, ts);
    vout = kfx_xyz{vert.x, vert.y, vert.z};

    const f3 q3 = ray_at(start, dir, t_cur);
    const int hx = voxel_floor(q3.x, p.csx, p.icsx), hy = voxel_floor(q3.y, p.csy, p.icsy),
              hz = voxel_floor(q3.z, p.csz, p.icsz);
    if (!(hx > 1 && hy > 1 && hz > 1 && hx < nxv - 2 && hy < nyv - 2 && hz < nzv - 2)) return;
    // The six central-difference samples one cell apart: 48 taps loaded
    // together.
    TriCell cn[6];
    {
        f3 tt = vert; tt.x += p.csx; cn[0] = tri_setup(tt, p);
        tt = vert; tt.x -= p.csx; cn[1] = tri_setup(tt, p);
        tt = vert; tt.y += p.csy; cn[2] = tri_setup(tt, p);
        tt = vert; tt.y -= p.csy; cn[3] = tri_setup(tt, p);
        tt = vert; tt.z += p.csz; cn[4] = tri_setup(tt, p);
        tt = vert; tt.z -= p.csz; cn[5] = tri_setup(tt, p);
    }
    int16_t vn[6][8];
#pragma unroll
    for (int i = 0; i < 6; ++i) tri_load(vol, cn[i], vn[i]);
    float f[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) f[i] = tri_eval(cn[i], vn[i]);
    if (kCount)
        ns += 8u * (uint32_t)(!isnan(f[0]) + !isnan(f[1]) + !isnan(f[2]) + !isnan(f[3]) + !isnan(f[4]) +
                              !isnan(f[5]));
    const f3 n = normalized3(f3{f[0] - f[1], f[2] - f[3], f[4] - f[5]});
    nout = kfx_xyz{n.x, n.y, n.z};
}

// The skipping march: the reference's samples (the same times, the same
// voxels) and the same outcome, without loading the samples that cannot
// change it.  The march stops at the first pair of consecutive samples with
// opposite signs (- to +: no hit; + to -: the crossing), so a sample that is
// not negative can only matter next to a negative one.  In a brick whose
// flag is clear (no negative value) no sample is loaded: the first sample
// in it is located exactly (and loaded only after a negative sample: a
// positive value ends the march), then every later sample whose time lies
// in the brick shrunk by kRayBrickMargin metres on each side (certainly
// inside it, whatever the rounding of its position) is stepped over with
// the running time sum alone.  When a negative sample follows unloaded
// samples, the last of them is loaded to decide the crossing.  In a flagged
// brick the samples go kMarchBatch at a time, loaded together.
constexpr int kRayBrickShift = 4;  // bricks of 16^3 voxels
constexpr int kRayBrick = 1 << kRayBrickShift;
constexpr float kRayBrickMargin = 1e-4f;

__device__ __forceinline__ bool march_skip(const kfx_tsdf* __restrict__ vol, const RaycastParams& p, f3 start,
                                           f3 dir, float t, int tsdf, float& hit_tc, float& hit_tn) {
    const kfx_round_buffer& rb = p.rb;
    const f3 vol_m{rb.volume_size.x, rb.volume_size.y, rb.volume_size.z};
    const float max_time = 3.0f * (vol_m.x + vol_m.y + vol_m.z);
    const float step = p.time_step;
    const float fnx = (float)rb.voxels_size.x, fny = (float)rb.voxels_size.y, fnz = (float)rb.voxels_size.z;
    const f3 idir{1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z};
    const f3 bsz{(float)kRayBrick * p.csx, (float)kRayBrick * p.csy, (float)kRayBrick * p.csz};
    // Brick b's shrunk box along an axis is entered at (b bsz + e - start) /
    // dir and left at (b bsz + x - start) / dir, with the faces taken by the
    // sign of dir: e = m or bsz - m, x = bsz - m or m (m the margin).
    const auto offsets = [&](float d, float bs, float st, float& e, float& x) {
        e = (d > 0.0f ? kRayBrickMargin : bs - kRayBrickMargin) - st;
        x = (d > 0.0f ? bs - kRayBrickMargin : kRayBrickMargin) - st;
    };
    f3 oe, ox;
    offsets(dir.x, bsz.x, start.x, oe.x, ox.x);
    offsets(dir.y, bsz.y, start.y, oe.y, ox.y);
    offsets(dir.z, bsz.z, start.z, oe.z, ox.z);
    bool known = true;  // tsdf holds the previous sample's value (else it is >= 0, at time t_unk)
    float t_unk = 0.0f;
    for (;;) {
        if (!(t < max_time)) return false;
        const float tn = t + step;
        const float fx = floorf(div_cell(start.x + dir.x * tn, p.csx, p.icsx));
        const float fy = floorf(div_cell(start.y + dir.y * tn, p.csy, p.icsy));
        const float fz = floorf(div_cell(start.z + dir.z * tn, p.csz, p.icsz));
        if (!((fx >= 0.0f) & (fx < fnx) & (fy >= 0.0f) & (fy < fny) & (fz >= 0.0f) & (fz < fnz))) return false;
        const uint32_t gx = (uint32_t)(int)fx, gy = (uint32_t)(int)fy, gz = (uint32_t)(int)fz;
        const uint32_t bx = gx >> kRayBrickShift, by = gy >> kRayBrickShift, bz = gz >> kRayBrickShift;
        if (p.bricks[(bz * (uint32_t)p.nby + by) * (uint32_t)p.nbx + bx] == 0) {
            // This sample is not negative; loaded only after a negative one
            // (a positive value ends the march there).
            if (known && tsdf < 0) {
                const int cur = vol[march_address(gx, gy, gz, rb)].value;
                if (cur > 0) return false;  // - to +
                tsdf = cur;                 // 0
            } else {
                known = false;
            }
            // Samples certainly inside the brick: times in [t_lo, t_hi) (the
            // box shrunk by the margin, which covers the rounding of these
            // few operations).
            const float fbx = (float)bx, fby = (float)by, fbz = (float)bz;
            const float t_lo = fmaxf(fmaxf(fmaf(fbx, bsz.x, oe.x) * idir.x, fmaf(fby, bsz.y, oe.y) * idir.y),
                                     fmaf(fbz, bsz.z, oe.z) * idir.z);
            const float t_hi = fminf(fminf(fmaf(fbx, bsz.x, ox.x) * idir.x, fmaf(fby, bsz.y, ox.y) * idir.y),
                                     fmaf(fbz, bsz.z, ox.z) * idir.z);
            float tt = tn;
            for (;;) {
                const float tnn = tt + step;
                if (!(tnn >= t_lo && tnn < t_hi)) break;
                tt = tnn;
                known = false;
            }
            if (!known) t_unk = tt;
            t = tt;
            continue;
        }
        // A flagged brick: this sample and the next kMarchBatch - 1 are
        // loaded together (up to the first that leaves the volume or the
        // time range), then tested in order, as the full march does.
        float tcs[kMarchBatch], tns[kMarchBatch];
        uint32_t addrs[kMarchBatch];
        bool ok[kMarchBatch];
        tcs[0] = t;
        tns[0] = tn;
        addrs[0] = march_address(gx, gy, gz, rb);
        ok[0] = true;
        {
            float tt = tn;
            bool alive = true;
#pragma unroll
            for (int k = 1; k < kMarchBatch; ++k) {
                tcs[k] = tt;
                tns[k] = tt + step;
                const float qx = floorf(div_cell(start.x + dir.x * tns[k], p.csx, p.icsx));
                const float qy = floorf(div_cell(start.y + dir.y * tns[k], p.csy, p.icsy));
                const float qz = floorf(div_cell(start.z + dir.z * tns[k], p.csz, p.icsz));
                alive = alive & (tt < max_time) & (qx >= 0.0f) & (qx < fnx) & (qy >= 0.0f) & (qy < fny) &
                        (qz >= 0.0f) & (qz < fnz);
                ok[k] = alive;
                addrs[k] = alive ? march_address((uint32_t)(int)qx, (uint32_t)(int)qy, (uint32_t)(int)qz, rb) : 0u;
                tt = tns[k];
            }
        }
        int val[kMarchBatch];
#pragma unroll
        for (int k = 0; k < kMarchBatch; ++k) val[k] = ok[k] ? vol[addrs[k]].value : 0;
        if (!known) {
            // The samples before were not loaded (not negative): the last of
            // them is needed only when this one is negative.
            tsdf = 0;
            if (val[0] < 0) {
                const uint32_t ux = (uint32_t)(int)floorf(div_cell(start.x + dir.x * t_unk, p.csx, p.icsx));
                const uint32_t uy = (uint32_t)(int)floorf(div_cell(start.y + dir.y * t_unk, p.csy, p.icsy));
                const uint32_t uz = (uint32_t)(int)floorf(div_cell(start.z + dir.z * t_unk, p.csz, p.icsz));
                tsdf = vol[march_address(ux, uy, uz, rb)].value;
            }
            known = true;
        }
#pragma unroll
        for (int k = 0; k < kMarchBatch; ++k) {
            if (!ok[k]) return false;  // left the volume or the time range
            const int prev = tsdf;
            tsdf = val[k];
            if (prev < 0 && tsdf > 0) return false;
            if (prev > 0 && tsdf < 0) {
                hit_tc = tcs[k];
                hit_tn = tns[k];
                return true;
            }
        }
        t = tns[kMarchBatch - 1];
    }
}

// kCount: also count the algorithm's voxel reads (first sample, march
// samples inside the volume, 8 per in-bounds trilinear evaluation) into
// *count, for the raycast's algorithmic bytes; results are unchanged.
// The ray of pixel (x, y): its vertex and normal (NaN when it hits
// nothing).
template <bool kCount, bool kSkip>
__device__ __forceinline__ void raycast_pixel(const kfx_tsdf* __restrict__ vol, const RaycastParams& p, int x, int y,
                                              unsigned long long* __restrict__ count, kfx_xyz& vout,
                                              kfx_xyz& nout) {
    const kfx_round_buffer& rb = p.rb;
    const f3 vol_m{rb.volume_size.x, rb.volume_size.y, rb.volume_size.z};
    const f3 start{p.tx, p.ty, p.tz};
    // get_ray_next, R * ray, + t, then (ray_next - start) normalised.
    const f3 rn{((float)x - p.cx) / p.fx, ((float)y - p.cy) / p.fy, 1.0f};
    const f3 rr{dot3(f3{p.r[0], p.r[1], p.r[2]}, rn), dot3(f3{p.r[3], p.r[4], p.r[5]}, rn),
                dot3(f3{p.r[6], p.r[7], p.r[8]}, rn)};
    const f3 next{rr.x + start.x, rr.y + start.y, rr.z + start.z};
    f3 dir = normalized3(f3{next.x - start.x, next.y - start.y, next.z - start.z});
    dir.x = (dir.x == 0.0f) ? 1e-15f : dir.x;
    dir.y = (dir.y == 0.0f) ? 1e-15f : dir.y;
    dir.z = (dir.z == 0.0f) ? 1e-15f : dir.z;

    // getMinTime / getMaxTime (raycast.cl:153-171).
    const float txmin = ((dir.x > 0.0f ? 0.0f : vol_m.x) - start.x) / dir.x;
    const float tymin = ((dir.y > 0.0f ? 0.0f : vol_m.y) - start.y) / dir.y;
    const float tzmin = ((dir.z > 0.0f ? 0.0f : vol_m.z) - start.z) / dir.z;
    float t_enter = fmaxf(fmaxf(txmin, tymin), tzmin);
    const float txmax = ((dir.x > 0.0f ? vol_m.x : 0.0f) - start.x) / dir.x;
    const float tymax = ((dir.y > 0.0f ? vol_m.y : 0.0f) - start.y) / dir.y;
    const float tzmax = ((dir.z > 0.0f ? vol_m.z : 0.0f) - start.z) / dir.z;
    const float t_exit = fminf(fminf(txmax, tymax), tzmax);

    t_enter = fmaxf(t_enter, 0.0f);
    if (!(t_enter < t_exit)) return;

    float time_curr = t_enter;
    const int nxv = rb.voxels_size.x, nyv = rb.voxels_size.y, nzv = rb.voxels_size.z;
    {
        const f3 q = ray_at(start, dir, time_curr);
        int gx = voxel_floor(q.x, p.csx, p.icsx), gy = voxel_floor(q.y, p.csy, p.icsy),
            gz = voxel_floor(q.z, p.csz, p.icsz);
        gx = max(0, min(gx, nxv - 1));
        gy = max(0, min(gy, nyv - 1));
        gz = max(0, min(gz, nzv - 1));
        // The march tests only the signs of the samples; the sign of
        // value / 32767 is the sign of the stored int16, so the samples stay
        // undecoded here.
        int tsdf = read_raw(vol, rb, gx, gy, gz);
        uint32_t ns = 1;
        const float max_time = 3.0f * (vol_m.x + vol_m.y + vol_m.z);
        const float step = p.time_step;
        const float fnx = (float)nxv, fny = (float)nyv, fnz = (float)nzv;

        // The reference loop, for (; t < max_time; t += step) { sample at
        // t + step; stop outside the volume; stop on - to +; refine on + to -
        // }, taken kMarchBatch samples at a time: the positions and loads of a
        // batch are issued first (loads stop at the first sample that leaves
        // the volume or the time range), then the samples are tested in order.
        // The sample times are the same running float sums as the reference's.
        // March samples are finite (finite origin and direction, bounded t),
        // so the floor is tested against [0, N) in float and converted only
        // then, which equals the saturating convert_int_sat_rtn test.
        bool found = false;
        float hit_tc = 0.0f, hit_tn = 0.0f;
        if constexpr (kSkip) found = march_skip(vol, p, start, dir, time_curr, tsdf, hit_tc, hit_tn);
        for (bool done = kSkip; !done;) {
            float tc[kMarchBatch], tn[kMarchBatch];
            bool ok[kMarchBatch];
            uint32_t addr[kMarchBatch];
            float t = time_curr;
            bool alive = true;
#pragma unroll
            for (int k = 0; k < kMarchBatch; ++k) {
                tc[k] = t;
                tn[k] = t + step;
                const float fx = floorf(div_cell(start.x + dir.x * tn[k], p.csx, p.icsx));
                const float fy = floorf(div_cell(start.y + dir.y * tn[k], p.csy, p.icsy));
                const float fz = floorf(div_cell(start.z + dir.z * tn[k], p.csz, p.icsz));
                alive = alive & (t < max_time) & (fx >= 0.0f) & (fx < fnx) & (fy >= 0.0f) & (fy < fny) &
                        (fz >= 0.0f) & (fz < fnz);
                ok[k] = alive;
                addr[k] = alive ? march_address((uint32_t)(int)fx, (uint32_t)(int)fy, (uint32_t)(int)fz, rb) : 0u;
                t = tn[k];
            }
            int16_t val[kMarchBatch];
#pragma unroll
            for (int k = 0; k < kMarchBatch; ++k) val[k] = ok[k] ? vol[addr[k]].value : (int16_t)0;
#pragma unroll
            for (int k = 0; k < kMarchBatch; ++k) {
                // !ok: the time range ended or the sample left the volume.
                if (!ok[k]) { done = true; break; }
                if (kCount) ++ns;
                const int tsdf_prev = tsdf;
                tsdf = val[k];
                if (tsdf_prev < 0 && tsdf > 0) { done = true; break; }
                if (tsdf_prev > 0 && tsdf < 0) {
                    found = true;
                    hit_tc = tc[k];
                    hit_tn = tn[k];
                    done = true;
                    break;
                }
            }
            time_curr = t;
        }
        if (found) raycast_hit<kCount>(vol, p, start, dir, hit_tc, hit_tn, vout, nout, ns);
        if (kCount) atomicAdd(count, (unsigned long long)ns);
    }
}

// Levels 1 and 2 of the maps from the workgroup's 16x16 tile (tile position
// lx, ly of this thread's pixel, its vertex v and normal n): the same
// resize4 arithmetic, in the same order, as the separate resize of the whole
// level 0, so the results are those bit for bit.  Every thread of the
// workgroup calls it.
__device__ __forceinline__ void raycast_pyramid(const RaycastParams& p, int tid, int lx, int ly, const kfx_xyz& v,
                                                const kfx_xyz& n) {
    __shared__ kfx_xyz sv[16][16], sn[16][16], s1v[8][8], s1n[8][8];
    sv[ly][lx] = v;
    sn[ly][lx] = n;
    __syncthreads();
    if (tid < 64) {
        const int ax = tid & 7, ay = tid >> 3;
        const kfx_xyz a = resize4(sv[2 * ay][2 * ax], sv[2 * ay][2 * ax + 1], sv[2 * ay + 1][2 * ax],
                                  sv[2 * ay + 1][2 * ax + 1], false);
        const kfx_xyz b = resize4(sn[2 * ay][2 * ax], sn[2 * ay][2 * ax + 1], sn[2 * ay + 1][2 * ax],
                                  sn[2 * ay + 1][2 * ax + 1], true);
        s1v[ay][ax] = a;
        s1n[ay][ax] = b;
        const int x1 = (int)blockIdx.x * 8 + ax, y1 = (int)blockIdx.y * 8 + ay;
        if (x1 < p.c1 && y1 < p.r1) {
            p.v1[y1 * p.c1 + x1] = a;
            p.n1[y1 * p.c1 + x1] = b;
        }
    }
    __syncthreads();
    if (tid < 16) {
        const int ax = tid & 3, ay = tid >> 2;
        const int x2 = (int)blockIdx.x * 4 + ax, y2 = (int)blockIdx.y * 4 + ay;
        if (x2 < p.c2 && y2 < p.r2) {
            p.v2[y2 * p.c2 + x2] = resize4(s1v[2 * ay][2 * ax], s1v[2 * ay][2 * ax + 1], s1v[2 * ay + 1][2 * ax],
                                           s1v[2 * ay + 1][2 * ax + 1], false);
            p.n2[y2 * p.c2 + x2] = resize4(s1n[2 * ay][2 * ax], s1n[2 * ay][2 * ax + 1], s1n[2 * ay + 1][2 * ax],
                                           s1n[2 * ay + 1][2 * ax + 1], true);
        }
    }
}

template <bool kDevicePose, bool kCount = false, bool kSkip = false>
__global__ void __launch_bounds__(256) raycast_kernel(const kfx_tsdf* __restrict__ vol,
                                                      kfx_xyz* __restrict__ vmap,
                                                      kfx_xyz* __restrict__ nmap, RaycastParams p,
                                                      const float* __restrict__ dpose,
                                                      const int32_t* __restrict__ dskip,
                                                      unsigned long long* __restrict__ count = nullptr) {
    if (kDevicePose && dskip != nullptr && *dskip != 0) return;  // uniform across the grid
    if (kDevicePose) {
        for (int i = 0; i < 9; ++i) p.r[i] = dpose[i];
        p.tx = dpose[9]; p.ty = dpose[10]; p.tz = dpose[11];
    }
    // A workgroup takes a 16x16 pixel tile, each wave an 8x8 quarter of it:
    // rays of one wave stay close, so their marches diverge less than a
    // 64-pixel row's.
    const int tid = threadIdx.y * blockDim.x + threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lx = 8 * (wave & 1) + (lane & 7), ly = 8 * (wave >> 1) + (lane >> 3);
    const int x = (int)blockIdx.x * 16 + lx, y = (int)blockIdx.y * 16 + ly;
    kfx_xyz vout{NAN, NAN, NAN};
    kfx_xyz nout{NAN, NAN, NAN};
    if (x < p.cols && y < p.rows) {
        raycast_pixel<kCount, kSkip>(vol, p, x, y, count, vout, nout);
        vmap[p.cols * y + x] = vout;
        nmap[p.cols * y + x] = nout;
    }
    if (p.v1 != nullptr) raycast_pyramid(p, tid, lx, ly, vout, nout);  // uniform
}

int check_round_buffer(const kfx_round_buffer* rb, const char* who);

// Exact brick flags: a workgroup takes the 16 x 16 rows (y, z) of one
// brick row (16 local y by 16 local z, every x), reading each row from
// memory in order as 16-byte quads (a row is contiguous; the cyclic offset
// only rotates it), and sets the bits of the x bricks whose voxels hold a
// negative value; the workgroup's bits then become the bricks' flags.
constexpr int kFlagThreads = 256;
__global__ void __launch_bounds__(kFlagThreads) brick_flags_kernel(const kfx_tsdf* __restrict__ vol,
                                                                   kfx_round_buffer rb,
                                                                   uint8_t* __restrict__ bricks) {
    __shared__ uint32_t words[kFlagThreads / 64][2];
    const int nx = rb.voxels_size.x, ny = rb.voxels_size.y, nz = rb.voxels_size.z;
    const int nbx = nx >> kRayBrickShift;  // at most 64 (the host checks nx <= 1024)
    const int by = blockIdx.x, bz = blockIdx.y;
    const int quads = nx / 4;
    const uint32_t ox = (uint32_t)rb.origin_GRID.x;
    uint64_t mask = 0;
    for (int i = threadIdx.x; i < kRayBrick * kRayBrick * quads; i += kFlagThreads) {
        const int q = i % quads, row = i / quads;
        const int y = by * kRayBrick + row % kRayBrick, z = bz * kRayBrick + row / kRayBrick;
        const uint32_t py = (uint32_t)wrap_once(y + rb.origin_GRID.y, ny), pz = (uint32_t)wrap_once(z + rb.origin_GRID.z, nz);
        const uint4 u = *reinterpret_cast<const uint4*>(vol + ((size_t)pz * (size_t)ny + py) * (size_t)nx + 4 * q);
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if ((int16_t)(w[k] & 0xffffu) < 0) {
                // physical x 4 q + k is local x (4 q + k - ox) mod nx
                uint32_t lx = (uint32_t)(4 * q + k) - ox;
                lx = min(lx, lx + (uint32_t)nx);
                mask |= 1ull << (lx >> kRayBrickShift);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) mask |= __shfl_xor(mask, off);
    if ((threadIdx.x & 63) == 0) {
        words[threadIdx.x >> 6][0] = (uint32_t)mask;
        words[threadIdx.x >> 6][1] = (uint32_t)(mask >> 32);
    }
    __syncthreads();
    if (threadIdx.x < nbx) {
        const int h = threadIdx.x >> 5, b = threadIdx.x & 31;
        uint32_t m = 0;
        for (int wv = 0; wv < kFlagThreads / 64; ++wv) m |= words[wv][h];
        bricks[((size_t)bz * (size_t)(ny >> kRayBrickShift) + by) * (size_t)nbx + threadIdx.x] = (m >> b) & 1u;
    }
}

size_t brick_flag_bytes(const kfx_point_3i& v) {
    if (v.x <= 0 || v.y <= 0 || v.z <= 0 || v.x % kRayBrick || v.y % kRayBrick || v.z % kRayBrick) return 0;
    return (size_t)(v.x / kRayBrick) * (size_t)(v.y / kRayBrick) * (size_t)(v.z / kRayBrick);
}

int brick_flags_build(const kfx_tsdf* volume, const kfx_round_buffer* rb, uint8_t* bricks) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(volume && bricks, "kfx_brick_flags_build: null pointer");
    int rc = check_round_buffer(rb, "kfx_brick_flags_build");
    if (rc) return rc;
    const size_t n = brick_flag_bytes(rb->voxels_size);
    KFX_REQUIRE(n > 0, "kfx_brick_flags_build: volume axes must be multiples of %d", kRayBrick);
    KFX_REQUIRE(rb->voxels_size.x <= 64 * kRayBrick, "kfx_brick_flags_build: at most %d voxels along x",
                64 * kRayBrick);
    ProfScope prof("brick_flags");
    dim3 grid(rb->voxels_size.y >> kRayBrickShift, rb->voxels_size.z >> kRayBrickShift);
    brick_flags_kernel<<<grid, kFlagThreads, 0, stream()>>>(volume, *rb, bricks);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

}  // namespace kfx

using namespace kfx;

static int raycast_launch(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                          const kfx_round_buffer* rb, const kfx_point_3f* translation,
                          const kfx_matrix_3f* rotation, const float* dev_pose, float truncation_distance,
                          const kfx_intrinsics* intr, const kfx_point_3f* cell_size, const int32_t* dev_skip,
                          unsigned long long* dev_count = nullptr, const uint8_t* bricks = nullptr,
                          kfx_xyz* const* pyramid = nullptr) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(volume && vertex_map && normal_map && translation && rotation && intr && cell_size,
                "kfx_raycast: null pointer");
    int rc = check_round_buffer(rb, "kfx_raycast");
    if (rc) return rc;
    KFX_REQUIRE(intr->cols > 0 && intr->rows > 0, "kfx_raycast: empty image");
    RaycastParams p;
    for (int r = 0; r < 3; ++r) {
        p.r[3 * r + 0] = rotation->data[r].x;
        p.r[3 * r + 1] = rotation->data[r].y;
        p.r[3 * r + 2] = rotation->data[r].z;
    }
    p.tx = translation->x; p.ty = translation->y; p.tz = translation->z;
    p.csx = cell_size->x; p.csy = cell_size->y; p.csz = cell_size->z;
    KFX_REQUIRE(p.csx > 0.0f && p.csy > 0.0f && p.csz > 0.0f, "kfx_raycast: cell size must be positive");
    p.icsx = 1.0f / p.csx; p.icsy = 1.0f / p.csy; p.icsz = 1.0f / p.csz;
    p.fx = intr->fx; p.fy = intr->fy; p.cx = intr->cx; p.cy = intr->cy;
    p.cols = intr->cols; p.rows = intr->rows;
    // time_step = Truncation_Distance * 0.8 (raycast.cl:209), the product
    // taken in double as a C compiler would with the double literal.
    p.time_step = (float)((double)truncation_distance * 0.8);
    p.rb = *rb;
    p.bricks = bricks;
    p.nbx = rb->voxels_size.x >> kRayBrickShift;
    p.nby = rb->voxels_size.y >> kRayBrickShift;
    p.v1 = p.n1 = p.v2 = p.n2 = nullptr;
    p.c1 = intr->cols / 2; p.r1 = intr->rows / 2;
    p.c2 = p.c1 / 2; p.r2 = p.r1 / 2;
    if (pyramid != nullptr) {
        KFX_REQUIRE(pyramid[0] && pyramid[1] && pyramid[2] && pyramid[3], "kfx_raycast: null pyramid map");
        p.v1 = pyramid[0]; p.n1 = pyramid[1]; p.v2 = pyramid[2]; p.n2 = pyramid[3];
    }
    KFX_REQUIRE(bricks == nullptr || brick_flag_bytes(rb->voxels_size) > 0,
                "kfx_raycast: brick flags need volume axes that are multiples of %d", kRayBrick);
    ProfScope prof("raycast");
    dim3 block(64, 4);  // 16x16 pixel tiles (the kernel's mapping)
    dim3 grid(divup(intr->cols, 16), divup(intr->rows, 16));
    if (bricks != nullptr && dev_pose)
        raycast_kernel<true, false, true><<<grid, block, 0, stream()>>>(volume, vertex_map, normal_map, p, dev_pose,
                                                                        dev_skip);
    else if (bricks != nullptr)
        raycast_kernel<false, false, true><<<grid, block, 0, stream()>>>(volume, vertex_map, normal_map, p, nullptr,
                                                                         nullptr);
    else if (dev_count)
        raycast_kernel<false, true><<<grid, block, 0, stream()>>>(volume, vertex_map, normal_map, p, nullptr, nullptr,
                                                                  dev_count);
    else if (dev_pose)
        raycast_kernel<true><<<grid, block, 0, stream()>>>(volume, vertex_map, normal_map, p, dev_pose, dev_skip);
    else
        raycast_kernel<false><<<grid, block, 0, stream()>>>(volume, vertex_map, normal_map, p, nullptr, nullptr);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

extern "C" int kfx_raycast(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                           const kfx_round_buffer* rb, const kfx_point_3f* translation,
                           const kfx_matrix_3f* rotation, float truncation_distance,
                           const kfx_intrinsics* intr, const kfx_point_3f* cell_size) {
    return raycast_launch(volume, vertex_map, normal_map, rb, translation, rotation, nullptr, truncation_distance,
                          intr, cell_size, nullptr);
}

extern "C" int kfx_raycast_posed(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                                 const kfx_round_buffer* rb, const float* dev_pose, float truncation_distance,
                                 const kfx_intrinsics* intr, const kfx_point_3f* cell_size, const int32_t* dev_skip) {
    KFX_REQUIRE(dev_pose != nullptr, "kfx_raycast_posed: null pose");
    const kfx_matrix_3f ident = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    const kfx_point_3f zero = {0, 0, 0};
    return raycast_launch(volume, vertex_map, normal_map, rb, &zero, &ident, dev_pose, truncation_distance, intr,
                          cell_size, dev_skip);
}

int kfx::raycast_posed_bricks(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                              const kfx_round_buffer* rb, const float* dev_pose, float truncation_distance,
                              const kfx_intrinsics* intr, const kfx_point_3f* cell_size, const int32_t* dev_skip,
                              const uint8_t* bricks, kfx_xyz* const* pyramid) {
    KFX_REQUIRE(dev_pose != nullptr, "kfx_raycast_posed: null pose");
    const kfx_matrix_3f ident = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    const kfx_point_3f zero = {0, 0, 0};
    return raycast_launch(volume, vertex_map, normal_map, rb, &zero, &ident, dev_pose, truncation_distance, intr,
                          cell_size, dev_skip, nullptr, bricks, pyramid);
}

extern "C" int kfx_brick_flags_size(const kfx_round_buffer* rb, size_t* bytes) {
    KFX_REQUIRE(rb && bytes, "kfx_brick_flags_size: null pointer");
    *bytes = brick_flag_bytes(rb->voxels_size);
    return KFX_OK;
}

extern "C" int kfx_brick_flags_build(const kfx_tsdf* volume, const kfx_round_buffer* rb, uint8_t* bricks) {
    return brick_flags_build(volume, rb, bricks);
}

extern "C" int kfx_raycast_bricks(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                                  const kfx_round_buffer* rb, const kfx_point_3f* translation,
                                  const kfx_matrix_3f* rotation, float truncation_distance, const kfx_intrinsics* intr,
                                  const kfx_point_3f* cell_size, const uint8_t* bricks) {
    KFX_REQUIRE(bricks != nullptr, "kfx_raycast_bricks: null brick flags");
    return raycast_launch(volume, vertex_map, normal_map, rb, translation, rotation, nullptr, truncation_distance,
                          intr, cell_size, nullptr, nullptr, bricks);
}

extern "C" int kfx_raycast_count_samples(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                                         const kfx_round_buffer* rb, const kfx_point_3f* translation,
                                         const kfx_matrix_3f* rotation, float truncation_distance,
                                         const kfx_intrinsics* intr, const kfx_point_3f* cell_size,
                                         uint64_t* samples) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(samples != nullptr, "kfx_raycast_count_samples: null output");
    unsigned long long* dcount = nullptr;
    KFX_CHECK_HIP(hipMallocAsync((void**)&dcount, sizeof(unsigned long long), stream()));
    hipError_t e = hipMemsetAsync(dcount, 0, sizeof(unsigned long long), stream());
    int rc = KFX_OK;
    if (e == hipSuccess)
        rc = raycast_launch(volume, vertex_map, normal_map, rb, translation, rotation, nullptr, truncation_distance,
                            intr, cell_size, nullptr, dcount);
    unsigned long long host = 0;
    if (e == hipSuccess && rc == KFX_OK)
        e = hipMemcpyAsync(&host, dcount, sizeof(host), hipMemcpyDeviceToHost, stream());
    hipError_t e2 = hipFreeAsync(dcount, stream());
    hipError_t e3 = hipStreamSynchronize(stream());
    if (rc) return rc;
    KFX_CHECK_HIP(e);
    KFX_CHECK_HIP(e2);
    KFX_CHECK_HIP(e3);
    *samples = host;
    return KFX_OK;
}
