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

// interpolateTrilineary (raycast.cl:95-145): NaN outside [1, N-2].
__device__ float interpolate_trilinear(const kfx_tsdf* __restrict__ vol, f3 pt, const RaycastParams& p) {
    const kfx_round_buffer& rb = p.rb;
    int gx = voxel_floor(pt.x, p.csx, p.icsx);
    int gy = voxel_floor(pt.y, p.csy, p.icsy);
    int gz = voxel_floor(pt.z, p.csz, p.icsz);
    if (gx <= 0 || gx >= rb.voxels_size.x - 1) return NAN;
    if (gy <= 0 || gy >= rb.voxels_size.y - 1) return NAN;
    if (gz <= 0 || gz >= rb.voxels_size.z - 1) return NAN;
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
    const float t000 = read_tsdf(vol, rb, gx + 0, gy + 0, gz + 0);
    const float t001 = read_tsdf(vol, rb, gx + 0, gy + 0, gz + 1);
    const float t010 = read_tsdf(vol, rb, gx + 0, gy + 1, gz + 0);
    const float t011 = read_tsdf(vol, rb, gx + 0, gy + 1, gz + 1);
    const float t100 = read_tsdf(vol, rb, gx + 1, gy + 0, gz + 0);
    const float t101 = read_tsdf(vol, rb, gx + 1, gy + 0, gz + 1);
    const float t110 = read_tsdf(vol, rb, gx + 1, gy + 1, gz + 0);
    const float t111 = read_tsdf(vol, rb, gx + 1, gy + 1, gz + 1);
    return (1.0f - a) * ((1.0f - b) * (t000 * (1.0f - c) + t001 * c) + b * (t010 * (1.0f - c) + t011 * c)) +
           a * ((1.0f - b) * (t100 * (1.0f - c) + t101 * c) + b * (t110 * (1.0f - c) + t111 * c));
}

__device__ __forceinline__ f3 ray_at(f3 o, f3 d, float t) {
    return f3{o.x + d.x * t, o.y + d.y * t, o.z + d.z * t};
}

// kCount: also count the algorithm's voxel reads (first sample, march
// samples inside the volume, 8 per in-bounds trilinear evaluation) into
// *count, for the raycast's algorithmic bytes; results are unchanged.
template <bool kDevicePose, bool kCount = false>
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
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= p.cols || y >= p.rows) return;
    const int idx = p.cols * y + x;
    kfx_xyz vout{NAN, NAN, NAN};
    kfx_xyz nout{NAN, NAN, NAN};

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
    if (!(t_enter < t_exit)) {
        vmap[idx] = vout;
        nmap[idx] = nout;
        return;
    }

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
        for (bool done = false; !done;) {
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
        if (found) {
            do {
                const float t_cur = hit_tc;
                const f3 q1 = ray_at(start, dir, hit_tn);
                const float ftdt = interpolate_trilinear(vol, q1, p);
                if (isnan(ftdt)) break;
                if (kCount) ns += 8;
                const f3 q2 = ray_at(start, dir, t_cur);
                const float ft = interpolate_trilinear(vol, q2, p);
                if (isnan(ft)) break;
                if (kCount) ns += 8;
                const float ts = t_cur - step * ft / (ftdt - ft);
                const f3 vert = ray_at(start, dir, ts);
                vout = kfx_xyz{vert.x, vert.y, vert.z};

                const f3 q3 = ray_at(start, dir, t_cur);
                const int hx = voxel_floor(q3.x, p.csx, p.icsx), hy = voxel_floor(q3.y, p.csy, p.icsy),
                          hz = voxel_floor(q3.z, p.csz, p.icsz);
                if (hx > 1 && hy > 1 && hz > 1 && hx < nxv - 2 && hy < nyv - 2 && hz < nzv - 2) {
                    f3 tt = vert;
                    tt.x += p.csx;
                    const float fx1 = interpolate_trilinear(vol, tt, p);
                    tt = vert;
                    tt.x -= p.csx;
                    const float fx2 = interpolate_trilinear(vol, tt, p);
                    tt = vert;
                    tt.y += p.csy;
                    const float fy1 = interpolate_trilinear(vol, tt, p);
                    tt = vert;
                    tt.y -= p.csy;
                    const float fy2 = interpolate_trilinear(vol, tt, p);
                    tt = vert;
                    tt.z += p.csz;
                    const float fz1 = interpolate_trilinear(vol, tt, p);
                    tt = vert;
                    tt.z -= p.csz;
                    const float fz2 = interpolate_trilinear(vol, tt, p);
                    if (kCount)
                        ns += 8u * (uint32_t)(!isnan(fx1) + !isnan(fx2) + !isnan(fy1) + !isnan(fy2) + !isnan(fz1) +
                                              !isnan(fz2));
                    const f3 n = normalized3(f3{fx1 - fx2, fy1 - fy2, fz1 - fz2});
                    nout = kfx_xyz{n.x, n.y, n.z};
                }
            } while (false);
        }
        if (kCount) atomicAdd(count, (unsigned long long)ns);
    }
    vmap[idx] = vout;
    nmap[idx] = nout;
}

int check_round_buffer(const kfx_round_buffer* rb, const char* who);

}  // namespace kfx

using namespace kfx;

static int raycast_launch(const kfx_tsdf* volume, kfx_xyz* vertex_map, kfx_xyz* normal_map,
                          const kfx_round_buffer* rb, const kfx_point_3f* translation,
                          const kfx_matrix_3f* rotation, const float* dev_pose, float truncation_distance,
                          const kfx_intrinsics* intr, const kfx_point_3f* cell_size, const int32_t* dev_skip,
                          unsigned long long* dev_count = nullptr) {
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
    ProfScope prof("raycast");
    dim3 block(64, 4);
    dim3 grid(divup(intr->cols, 64), divup(intr->rows, 4));
    if (dev_count)
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
