// kfx_slice.hip — cyclic volume shift support: clearing the slab that leaves
// the volume and extracting its zero-crossing voxels as a point cloud.
//
// "Black zone" (clean_color_slice.cl:29-40): local voxel (x, y, z) belongs to
// it when, on any axis, the unwrapped coordinate x + origin_GRID.x lies in
// [min, max] (bounds inclusive, no modulo).
//
// Extraction reproduces the reference's three-pass compaction
// (scan_slice_of_both_volumes.cl:76-208, the exclusive scan, and
// extract_slice_as_cloud.cl:72-202) without materialising its 4 predicate
// planes: the volume is processed in z chunks of `step` slices
// (mem_utils.cpp:43-85); inside a chunk the output order is
// [plane][z][y][x] with planes
//   0: x-neighbour zero crossing -> point (x+1, y, z)
//   1: y-neighbour zero crossing -> point (x, y+1, z)
//   2: z-neighbour zero crossing (z+1 inside the chunk) -> point (x, y, z+1)
//   3: colour voxel set          -> point (x, y, z)
// and chunks follow each other.  Pass 1 counts the four predicates of every
// (y, z) voxel row with one wave per row; pass 2 scans the counts in output
// order; pass 3 recomputes the predicates and writes each point at its row
// offset plus its rank inside the row (ballot + popcount), so the result is
// the reference's order exactly.
#include "kfx_common.h"

namespace kfx {

int check_round_buffer(const kfx_round_buffer* rb, const char* who);

struct SliceParams {
    kfx_round_buffer rb;
    int minb[3], maxb[3];
    int step;       // chunk depth in z
};

__device__ __forceinline__ bool in_black_zone(const SliceParams& p, int x, int y, int z) {
    const int ix = x + p.rb.origin_GRID.x, iy = y + p.rb.origin_GRID.y, iz = z + p.rb.origin_GRID.z;
    return (ix >= p.minb[0] && ix <= p.maxb[0]) || (iy >= p.minb[1] && iy <= p.maxb[1]) ||
           (iz >= p.minb[2] && iz <= p.maxb[2]);
}

// Local x range of row (y, z) that lies in the black zone: the whole row when
// y or z matches, else the x-slab.  The zone test uses the unwrapped
// coordinate x + origin, which is monotonic in x, so the slab is one interval
// [lo, hi] of local x.  Returns the number of intervals (0 or 1).
__device__ __forceinline__ int row_ranges(const SliceParams& p, int y, int z, int lo[2], int hi[2]) {
    const int nx = p.rb.voxels_size.x;
    const int iy = y + p.rb.origin_GRID.y, iz = z + p.rb.origin_GRID.z;
    if ((iy >= p.minb[1] && iy <= p.maxb[1]) || (iz >= p.minb[2] && iz <= p.maxb[2])) {
        lo[0] = 0; hi[0] = nx - 1;
        return 1;
    }
    // x + ox in [minb, maxb]  <=>  x in [minb - ox, maxb - ox] within [0, nx)
    const int a = max(0, p.minb[0] - p.rb.origin_GRID.x), b = min(nx - 1, p.maxb[0] - p.rb.origin_GRID.x);
    if (a > b) return 0;
    lo[0] = a; hi[0] = b;
    return 1;
}

// Zero crossing between a voxel of value v and its neighbour n
// (scan_slice_of_both_volumes.cl:125-160).  With F = value / 32767 decoded,
// F != 1 is value != 32767 and the signs of F are those of the int16, so the
// test runs on the stored integers.
__device__ __forceinline__ bool crossing(int v, const kfx_tsdf n) {
    return (n.weight != 0 && n.value != KFX_MAX_DEPTH_VALUE) && ((v > 0 && n.value < 0) || (v < 0 && n.value > 0));
}

__device__ __forceinline__ bool color_set(kfx_color c) { return c.r != 0 || c.g != 0 || c.b != 0 || c.w != 0; }

// The four predicates of voxel (x, y, z) (in the black zone).
__device__ __forceinline__ unsigned predicates(const SliceParams& p, const kfx_tsdf* __restrict__ vol,
                                               const kfx_color* __restrict__ col, int x, int y, int z, int chunk_end) {
    unsigned m = 0;
    const kfx_tsdf v = vol[cyclic_address_in_range(x, y, z, p.rb)];
    if (v.weight != 0 && v.value != KFX_MAX_DEPTH_VALUE) {
        if (x + 1 < p.rb.voxels_size.x && crossing(v.value, vol[cyclic_address_in_range(x + 1, y, z, p.rb)])) m |= 1u;
        if (y + 1 < p.rb.voxels_size.y && crossing(v.value, vol[cyclic_address_in_range(x, y + 1, z, p.rb)])) m |= 2u;
        if (z + 1 < chunk_end && crossing(v.value, vol[cyclic_address_in_range(x, y, z + 1, p.rb)])) m |= 4u;
    }
    if (col != nullptr && color_set(col[cyclic_address_in_range(x, y, z, p.rb)])) m |= 8u;
    return m;
}

// counts[plane][z][y]: predicate counts of every voxel row; one wave per row,
// 4 rows per workgroup.
__global__ void __launch_bounds__(256) slice_count_kernel(const kfx_tsdf* __restrict__ vol,
                                                          const kfx_color* __restrict__ col, SliceParams p,
                                                          uint32_t* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ny = p.rb.voxels_size.y, nz = p.rb.voxels_size.z;
    if (row >= (long)ny * nz) return;
    const int z = (int)(row / ny), y = (int)(row % ny);
    const int chunk_end = min((z / p.step + 1) * p.step, nz);
    int lo[2], hi[2];
    const int nr = row_ranges(p, y, z, lo, hi);
    unsigned c[4] = {0, 0, 0, 0};
    for (int r = 0; r < nr; ++r)
        for (int x = lo[r] + lane; x <= hi[r]; x += 64) {
            const unsigned m = predicates(p, vol, col, x, y, z, chunk_end);
            c[0] += m & 1u; c[1] += (m >> 1) & 1u; c[2] += (m >> 2) & 1u; c[3] += (m >> 3) & 1u;
        }
    for (int k = 0; k < 4; ++k)
        for (int off = 32; off > 0; off >>= 1) c[k] += __shfl_xor(c[k], off);
    if (lane == 0) {
        const size_t plane = (size_t)ny * nz;
        for (int k = 0; k < 4; ++k) counts[k * plane + row] = c[k];
    }
}

// Exclusive scan of the counts in output order (chunk, plane, z, y), in
// place, with the grand total in *total.  Work is split into segments of
// 1024 rows: pass A sums each segment, pass B (one workgroup) scans the
// segment sums, pass C scans inside each segment.
constexpr int kScanSeg = 1024;

// Position of segment s of the output order -> (plane, first row).  Rows of
// one chunk are contiguous in [z][y] order; the output visits, per chunk,
// plane 0's rows, plane 1's rows, ...
struct ScanLayout {
    long rows_per_chunk;  // step * ny
    long rows_total;      // nz * ny
    int nchunks;
    long segs_per_chunk_plane;
};

__device__ __forceinline__ uint32_t* seg_base(uint32_t* counts, const ScanLayout& L, long seg, long& nrows) {
    const long per_chunk = L.segs_per_chunk_plane * 4;
    const long chunk = seg / per_chunk;
    const long rem = seg % per_chunk;
    const int plane = (int)(rem / L.segs_per_chunk_plane);
    const long s_in = rem % L.segs_per_chunk_plane;
    const long row0 = chunk * L.rows_per_chunk + s_in * kScanSeg;
    const long chunk_rows = min(L.rows_per_chunk, L.rows_total - chunk * L.rows_per_chunk);
    nrows = max(0L, min((long)kScanSeg, chunk_rows - s_in * kScanSeg));
    return counts + (size_t)plane * L.rows_total + row0;
}

__global__ void __launch_bounds__(256) scan_seg_sum_kernel(uint32_t* __restrict__ counts, ScanLayout L,
                                                           uint32_t* __restrict__ seg_sums, long nseg) {
    const long seg = blockIdx.x;
    if (seg >= nseg) return;
    long n;
    const uint32_t* base = seg_base(counts, L, seg, n);
    uint32_t s = 0;
    for (long i = threadIdx.x; i < n; i += 256) s += base[i];
    __shared__ uint32_t red[4];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) seg_sums[seg] = red[0] + red[1] + red[2] + red[3];
}

// Single-workgroup exclusive scan of n values in place (64-bit running sum
// to detect overflow of the 32-bit point index); writes the total.
__global__ void __launch_bounds__(1024) scan_single_kernel(uint32_t* __restrict__ v, long n,
                                                           unsigned long long* __restrict__ total) {
    __shared__ unsigned long long wsum[16];
    __shared__ unsigned long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (long base = 0; base < n; base += 1024) {
        const long i = base + threadIdx.x;
        const unsigned long long x = i < n ? v[i] : 0ull;
        unsigned long long incl = x;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned long long y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        unsigned long long before = carry;
        for (int k = 0; k < w; ++k) before += wsum[k];
        if (i < n) v[i] = (uint32_t)(before + incl - x);
        __syncthreads();
        if (threadIdx.x == 1023) carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) scan_seg_apply_kernel(uint32_t* __restrict__ counts, ScanLayout L,
                                                             const uint32_t* __restrict__ seg_offsets, long nseg) {
    const long seg = blockIdx.x;
    if (seg >= nseg) return;
    long n;
    uint32_t* base = seg_base(counts, L, seg, n);
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = seg_offsets[seg];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (long b = 0; b < n; b += 256) {
        const long i = b + threadIdx.x;
        const uint32_t x = i < n ? base[i] : 0u;
        uint32_t incl = x;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        uint32_t before = carry;
        for (int k = 0; k < w; ++k) before += wsum[k];
        if (i < n) base[i] = before + incl - x;
        __syncthreads();
        if (threadIdx.x == 255) carry = before + incl;
        __syncthreads();
    }
}

// Pass 3.  Writes XYZI points (out_pts) or records (out_rec, with the
// global grid offset g0), and colours when out_col is given.
__global__ void __launch_bounds__(256) slice_emit_kernel(const kfx_tsdf* __restrict__ vol,
                                                         const kfx_color* __restrict__ col, SliceParams p,
                                                         const uint32_t* __restrict__ offsets,
                                                         kfx_xyzi* __restrict__ out_pts,
                                                         kfx_slice_record* __restrict__ out_rec, int3 g0,
                                                         kfx_color* __restrict__ out_col) {
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ny = p.rb.voxels_size.y, nz = p.rb.voxels_size.z;
    if (row >= (long)ny * nz) return;
    const int z = (int)(row / ny), y = (int)(row % ny);
    const int chunk_end = min((z / p.step + 1) * p.step, nz);
    int lo[2], hi[2];
    const int nr = row_ranges(p, y, z, lo, hi);
    const size_t plane_stride = (size_t)ny * nz;
    uint32_t base[4];
    for (int k = 0; k < 4; ++k) base[k] = offsets[k * plane_stride + row];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = 0; r < nr; ++r)
        for (int x0 = lo[r]; x0 <= hi[r]; x0 += 64) {
            const int x = x0 + lane;
            const unsigned m = (x <= hi[r]) ? predicates(p, vol, col, x, y, z, chunk_end) : 0u;
            for (int k = 0; k < 4; ++k) {
                const bool bit = (m >> k) & 1u;
                const unsigned long long ballot = __ballot(bit);
                if (bit) {
                    const uint32_t idx = base[k] + (uint32_t)__popcll(ballot & below);
                    const int px = x + (k == 0), py = y + (k == 1), pz = z + (k == 2);
                    if (out_pts != nullptr) out_pts[idx] = kfx_xyzi{(float)px, (float)py, (float)pz, 1.0f};
                    if (out_rec != nullptr) {
                        const kfx_tsdf v = vol[cyclic_address_in_range(px, py, pz, p.rb)];
                        out_rec[idx] = kfx_slice_record{px + g0.x, py + g0.y, pz + g0.z, v.value, v.weight};
                    }
                    if (out_col != nullptr) {
                        kfx_color c{0, 0, 0, 0};
                        if (col != nullptr) c = col[cyclic_address_in_range(px, py, pz, p.rb)];
                        out_col[idx] = c;
                    }
                }
                base[k] += (uint32_t)__popcll(ballot);
            }
        }
}

// Colour of the voxel under each point (Extract_Color_Volume,
// common/src/opencl_wrappers.cpp:546-590; kernel extract_colors absent):
// voxel floor(p / cell) of the local grid, its stored colour; (0,0,0,0)
// when the point lies outside the cube or is not finite.
__global__ void __launch_bounds__(256) extract_colors_kernel(const kfx_color* __restrict__ col,
                                                             const kfx_xyzi* __restrict__ pts, uint64_t n,
                                                             kfx_point_3f cell, kfx_round_buffer rb,
                                                             kfx_color* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const kfx_xyzi p = pts[i];
    const float fx = floorf(p.x / cell.x), fy = floorf(p.y / cell.y), fz = floorf(p.z / cell.z);
    kfx_color c{0, 0, 0, 0};
    if (fx >= 0.0f && fx < (float)rb.voxels_size.x && fy >= 0.0f && fy < (float)rb.voxels_size.y && fz >= 0.0f &&
        fz < (float)rb.voxels_size.z)
        c = col[cyclic_address_in_range((int)fx, (int)fy, (int)fz, rb)];
    out[i] = c;
}

// clean_tsdf_slice (call site opencl_wrappers.cpp:835-854, kernel absent) and
// clean_color_slice.cl:41-67: reset every black-zone voxel, one lane per
// (x, y) column sweeping z, skipping the z range outside the zone for
// columns that are in it only through z.
template <typename T>
__global__ void __launch_bounds__(256) clean_slice_kernel(T* __restrict__ vol, SliceParams p) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= p.rb.voxels_size.x || y >= p.rb.voxels_size.y) return;
    const int nz = p.rb.voxels_size.z;
    const int ix = x + p.rb.origin_GRID.x, iy = y + p.rb.origin_GRID.y;
    int z0 = 0, z1 = nz - 1;
    if (!((ix >= p.minb[0] && ix <= p.maxb[0]) || (iy >= p.minb[1] && iy <= p.maxb[1]))) {
        z0 = max(0, p.minb[2] - p.rb.origin_GRID.z);
        z1 = min(nz - 1, p.maxb[2] - p.rb.origin_GRID.z);
    }
    const T zero{};
    for (int z = z0; z <= z1; ++z) vol[cyclic_address_in_range(x, y, z, p.rb)] = zero;
}

int determine_step(int depth) {
    // mem_utils.cpp:43-85 with divisor 4: 512 -> 64, 256 -> 32, 128 -> 16,
    // i.e. depth / 8; extended to any depth divisible by 8 (one chunk
    // otherwise), where the reference throws.
    if (depth >= 8 && depth % 8 == 0) return depth / 8;
    return depth;
}

static int slice_setup(const char* who, const kfx_point_3i* minb, const kfx_point_3i* maxb,
                       const kfx_round_buffer* rb, SliceParams& p) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(minb && maxb, "%s: null bounds", who);
    int rc = check_round_buffer(rb, who);
    if (rc) return rc;
    p.rb = *rb;
    p.minb[0] = minb->x; p.minb[1] = minb->y; p.minb[2] = minb->z;
    p.maxb[0] = maxb->x; p.maxb[1] = maxb->y; p.maxb[2] = maxb->z;
    p.step = determine_step(rb->voxels_size.z);
    return KFX_OK;
}

}  // namespace kfx

struct kfx_slicer {
    uint32_t* counts = nullptr;   // 4 * ny * nz
    uint32_t* seg_sums = nullptr;
    unsigned long long* total = nullptr;
    unsigned long long* host_total = nullptr;  // pinned
    long rows = 0, nseg = 0;
};

using namespace kfx;

extern "C" {

int kfx_determine_extract_step(int depth, int* step, int* num_of_steps) {
    KFX_REQUIRE(step && num_of_steps && depth > 0, "kfx_determine_extract_step: bad arguments");
    *step = determine_step(depth);
    *num_of_steps = depth / *step;
    return KFX_OK;
}

int kfx_clean_tsdf_slice(const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds, const kfx_round_buffer* rb,
                         kfx_tsdf* volume) {
    SliceParams p;
    int rc = slice_setup("kfx_clean_tsdf_slice", min_bounds, max_bounds, rb, p);
    if (rc) return rc;
    KFX_REQUIRE(volume != nullptr, "kfx_clean_tsdf_slice: null volume");
    ProfScope prof("clean_tsdf_slice");
    dim3 grid(divup(rb->voxels_size.x, 64), divup(rb->voxels_size.y, 4));
    clean_slice_kernel<kfx_tsdf><<<grid, 256, 0, stream()>>>(volume, p);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_clean_color_slice(const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds,
                          const kfx_round_buffer* rb, kfx_color* volume) {
    SliceParams p;
    int rc = slice_setup("kfx_clean_color_slice", min_bounds, max_bounds, rb, p);
    if (rc) return rc;
    KFX_REQUIRE(volume != nullptr, "kfx_clean_color_slice: null volume");
    ProfScope prof("clean_color_slice");
    dim3 grid(divup(rb->voxels_size.x, 64), divup(rb->voxels_size.y, 4));
    clean_slice_kernel<kfx_color><<<grid, 256, 0, stream()>>>(volume, p);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_slicer_create(const kfx_round_buffer* rb, kfx_slicer** out) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(rb && out, "kfx_slicer_create: null pointer");
    kfx_slicer* s = new kfx_slicer();
    s->rows = (long)rb->voxels_size.y * rb->voxels_size.z;
    const int step = determine_step(rb->voxels_size.z);
    const long rows_per_chunk = (long)step * rb->voxels_size.y;
    const long nchunks = (rb->voxels_size.z + step - 1) / step;
    s->nseg = nchunks * 4 * ((rows_per_chunk + kScanSeg - 1) / kScanSeg);
    hipError_t e = hipMalloc((void**)&s->counts, sizeof(uint32_t) * 4 * (size_t)s->rows);
    if (e == hipSuccess) e = hipMalloc((void**)&s->seg_sums, sizeof(uint32_t) * (size_t)s->nseg);
    if (e == hipSuccess) e = hipMalloc((void**)&s->total, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->host_total, sizeof(unsigned long long), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipFree(s->counts);
        (void)hipFree(s->seg_sums);
        (void)hipFree(s->total);
        delete s;
        return set_error(KFX_ERR_HIP, "kfx_slicer_create: %s", hipGetErrorString(e));
    }
    *out = s;
    return KFX_OK;
}

int kfx_slicer_destroy(kfx_slicer* s) {
    if (!s) return KFX_OK;
    KFX_REQUIRE_INIT();
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    (void)hipFree(s->counts);
    (void)hipFree(s->seg_sums);
    (void)hipFree(s->total);
    (void)hipHostFree(s->host_total);
    delete s;
    return KFX_OK;
}

static int extract(const char* who, kfx_slicer* s, const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds,
                   const kfx_round_buffer* rb, const kfx_tsdf* tsdf_volume, const kfx_color* color_volume,
                   kfx_xyzi* points, kfx_slice_record* records, kfx_color* colors, uint64_t capacity,
                   uint64_t* count) {
    SliceParams p;
    int rc = slice_setup(who, min_bounds, max_bounds, rb, p);
    if (rc) return rc;
    KFX_REQUIRE(s && tsdf_volume && count, "%s: null pointer", who);
    KFX_REQUIRE(s->rows == (long)rb->voxels_size.y * rb->voxels_size.z,
                "%s: slicer was created for another volume size", who);
    const long rows = s->rows;
    const unsigned row_blocks = (unsigned)((rows + 3) / 4);
    {
        ProfScope prof("slice_count");
        slice_count_kernel<<<row_blocks, 256, 0, stream()>>>(tsdf_volume, color_volume, p, s->counts);
        KFX_CHECK_LAUNCH();
    }
    ScanLayout L;
    L.rows_per_chunk = (long)p.step * rb->voxels_size.y;
    L.rows_total = rows;
    L.nchunks = (rb->voxels_size.z + p.step - 1) / p.step;
    L.segs_per_chunk_plane = (L.rows_per_chunk + kScanSeg - 1) / kScanSeg;
    {
        ProfScope prof("slice_scan");
        scan_seg_sum_kernel<<<(unsigned)s->nseg, 256, 0, stream()>>>(s->counts, L, s->seg_sums, s->nseg);
        KFX_CHECK_LAUNCH();
        scan_single_kernel<<<1, 1024, 0, stream()>>>(s->seg_sums, s->nseg, s->total);
        KFX_CHECK_LAUNCH();
        scan_seg_apply_kernel<<<(unsigned)s->nseg, 256, 0, stream()>>>(s->counts, L, s->seg_sums, s->nseg);
        KFX_CHECK_LAUNCH();
    }
    KFX_CHECK_HIP(hipMemcpyAsync(s->host_total, s->total, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream()));
    KFX_CHECK_HIP(hipStreamSynchronize(stream()));
    const unsigned long long n = *s->host_total;
    *count = n;
    KFX_REQUIRE(n < 0xFFFFFFFFull, "%s: %llu points overflow the 32-bit index", who, n);
    if (n == 0) return KFX_OK;
    if (n > capacity || (points == nullptr && records == nullptr) || (points != nullptr && colors == nullptr))
        return set_error(KFX_ERR_CAPACITY, "%s: %llu points, capacity %llu", who, n, (unsigned long long)capacity);
    const int3 g0 = make_int3((int)rintf(rb->origin_GRID_global.x), (int)rintf(rb->origin_GRID_global.y),
                              (int)rintf(rb->origin_GRID_global.z));
    ProfScope prof("slice_emit");
    slice_emit_kernel<<<row_blocks, 256, 0, stream()>>>(tsdf_volume, color_volume, p, s->counts, points, records, g0,
                                                        colors);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_extract_slice(kfx_slicer* s, const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds,
                      const kfx_round_buffer* rb, const kfx_tsdf* tsdf_volume, const kfx_color* color_volume,
                      kfx_xyzi* points, kfx_color* colors, uint64_t capacity, uint64_t* count) {
    return extract("kfx_extract_slice", s, min_bounds, max_bounds, rb, tsdf_volume, color_volume, points, nullptr,
                   colors, capacity, count);
}

int kfx_extract_slice_records(kfx_slicer* s, const kfx_point_3i* min_bounds, const kfx_point_3i* max_bounds,
                              const kfx_round_buffer* rb, const kfx_tsdf* tsdf_volume,
                              const kfx_color* color_volume, kfx_slice_record* records, kfx_color* colors,
                              uint64_t capacity, uint64_t* count) {
    return extract("kfx_extract_slice_records", s, min_bounds, max_bounds, rb, tsdf_volume, color_volume, nullptr,
                   records, colors, capacity, count);
}

int kfx_extract_colors(const kfx_color* color_volume, const kfx_round_buffer* rb, const kfx_xyzi* points,
                       uint64_t n, const kfx_point_3f* cell_size, kfx_color* colors) {
    KFX_REQUIRE_INIT();
    int rc = check_round_buffer(rb, "kfx_extract_colors");
    if (rc) return rc;
    KFX_REQUIRE(cell_size && cell_size->x > 0.0f && cell_size->y > 0.0f && cell_size->z > 0.0f,
                "kfx_extract_colors: cell size must be positive");
    if (n == 0) return KFX_OK;
    KFX_REQUIRE(color_volume && points && colors, "kfx_extract_colors: null pointer");
    ProfScope prof("extract_colors");
    extract_colors_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream()>>>(color_volume, points, n, *cell_size,
                                                                              *rb, colors);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

}  // extern "C"
