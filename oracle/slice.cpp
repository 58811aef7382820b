// slice.cpp — oracle for the cyclic shift: black-zone clean and the
// reference's three-pass slice extraction, restated literally (predicate
// planes, exclusive scan, scatter).  Test infrastructure only (see oracle.h).
#include <string.h>

#include <vector>

#include "oracle.h"

namespace {

// clean_color_slice.cl:29-40.
bool black_zone(int x, int y, int z, const kfx_point_3i* mn, const kfx_point_3i* mx, const kfx_round_buffer* rb) {
    const int ix = x + rb->origin_GRID.x, iy = y + rb->origin_GRID.y, iz = z + rb->origin_GRID.z;
    return (ix >= mn->x && ix <= mx->x) || (iy >= mn->y && iy <= mx->y) || (iz >= mn->z && iz <= mx->z);
}

float fetch(const kfx_tsdf* vol, const kfx_round_buffer* rb, int x, int y, int z, int* w) {
    const kfx_tsdf v = vol[orc_cyclic_address(x, y, z, rb)];
    *w = v.weight;
    return orc_tsdf_decode(v.value);
}

bool sign_change(float F, float Fn, int Wn) {
    return (Wn != 0 && Fn != 1.f) && ((F > 0 && Fn < 0) || (F < 0 && Fn > 0));
}

bool color_set(kfx_color c) { return c.r != 0 || c.g != 0 || c.b != 0 || c.w != 0; }

}  // namespace

// clean_color_slice.cl:41-67 pattern for either volume.
extern "C" void orc_clean_slice(const kfx_point_3i* mn, const kfx_point_3i* mx, const kfx_round_buffer* rb, void* vol,
                                int elem_bytes) {
    for (int y = 0; y < rb->voxels_size.y; ++y)
        for (int x = 0; x < rb->voxels_size.x; ++x)
            for (int z = 0; z < rb->voxels_size.z; ++z)
                if (black_zone(x, y, z, mn, mx, rb))
                    memset((char*)vol + (size_t)orc_cyclic_address(x, y, z, rb) * elem_bytes, 0, elem_bytes);
}

// Extract_Slice (opencl_wrappers.cpp:856-964): for each chunk [Min, Max) of
// `step` slices (mem_utils.cpp:43-85 divisor 4 -> depth / 8):
//   scan_slice_of_both_volumes.cl:76-208 — 4 predicate planes
//     [plane][z - Min][y][x];
//   exclusive scan of the 4 * step * Ny * Nx predicates;
//   extract_slice_as_cloud.cl:72-202 — scatter each true predicate's point
//     and colour to its scanned index,
// with the z-neighbour bounded by the chunk (z + 1 < Max) in both passes.
// (The reference's extract pass tests z + 1 < Nz instead, writing points the
// scan did not reserve at chunk boundaries; the scan's bound is followed.)
// Points beyond `capacity` are counted but not written.  colors may be NULL.
extern "C" uint64_t orc_extract_slice(const kfx_point_3i* mn, const kfx_point_3i* mx, const kfx_round_buffer* rb,
                                      const kfx_tsdf* vol, const kfx_color* col, kfx_xyzi* pts, kfx_color* cols,
                                      uint64_t capacity) {
    const int W = rb->voxels_size.x, H = rb->voxels_size.y, D = rb->voxels_size.z;
    const int step = (D >= 8 && D % 8 == 0) ? D / 8 : D;
    const size_t plane = (size_t)W * H * step;
    std::vector<uint32_t> pred(4 * plane);
    uint64_t total = 0;
    for (int Min = 0; Min < D; Min += step) {
        const int Max = Min + step;
        std::fill(pred.begin(), pred.end(), 0u);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                for (int z = Min; z < Max; ++z) {
                    const size_t cur = (size_t)(z % step) * W * H + (size_t)y * W + x;
                    if (!black_zone(x, y, z, mn, mx, rb)) continue;
                    int w;
                    const float F = fetch(vol, rb, x, y, z, &w);
                    if (w != 0 && F != 1.f) {
                        int wn = 0;
                        float fn;
                        if (x + 1 < W) {
                            fn = fetch(vol, rb, x + 1, y, z, &wn);
                            if (sign_change(F, fn, wn)) pred[cur] = 1;
                        }
                        if (y + 1 < H) {
                            fn = fetch(vol, rb, x, y + 1, z, &wn);
                            if (sign_change(F, fn, wn)) pred[cur + plane] = 1;
                        }
                        if (z + 1 < Max) {
                            fn = fetch(vol, rb, x, y, z + 1, &wn);
                            if (sign_change(F, fn, wn)) pred[cur + 2 * plane] = 1;
                        }
                    }
                    if (col && color_set(col[orc_cyclic_address(x, y, z, rb)])) pred[cur + 3 * plane] = 1;
                }
        // exclusive scan
        std::vector<uint64_t> idx(4 * plane);
        uint64_t run = 0;
        for (size_t i = 0; i < 4 * plane; ++i) {
            idx[i] = run;
            run += pred[i];
        }
        for (size_t i = 0; i < 4 * plane; ++i) {
            if (!pred[i]) continue;
            const int p = (int)(i / plane);
            const size_t r = i % plane;
            const int z = Min + (int)(r / ((size_t)W * H));
            const int y = (int)((r / W) % H), x = (int)(r % W);
            const int px = x + (p == 0), py = y + (p == 1), pz = z + (p == 2);
            const uint64_t o = total + idx[i];
            if (o < capacity) {
                if (pts) pts[o] = kfx_xyzi{(float)px, (float)py, (float)pz, 1.0f};
                if (cols) cols[o] = col ? col[orc_cyclic_address(px, py, pz, rb)] : kfx_color{0, 0, 0, 0};
            }
        }
        total += run;
    }
    return total;
}
