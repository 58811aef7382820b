// opencl_wrappers.cpp — the reference's host API (common/src/opencl_wrappers.cpp)
// implemented over the kfx C ABI.  Each wrapper validates its buffers,
// converts the POD arguments (layout-identical) and calls the matching
// kfx_* entry point; failures throw std::runtime_error.
#include "opencl_wrappers.h"

#include <math.h>

#include <stdexcept>
#include <string>

#include "kfx.h"

namespace {

void check(int rc, const char* what) {
    if (rc != KFX_OK) throw std::runtime_error(std::string(what) + ": " + kfx_last_error());
}

template <class T>
T* dev(gpu_array_base* b, size_t min_elems, const char* what) {
    if (!b || !b->get_device_ptr()) throw std::runtime_error(std::string(what) + ": unallocated buffer");
    if (b->get_total_size() < min_elems)
        throw std::runtime_error(std::string(what) + ": buffer holds " + std::to_string(b->get_total_size()) +
                                 " elements, needs " + std::to_string(min_elems));
    return static_cast<T*>(b->get_device_ptr());
}

const kfx_intrinsics* K(const intrinsics& i) { return reinterpret_cast<const kfx_intrinsics*>(&i); }
const kfx_matrix_3f* M(const matrix_3f& m) { return reinterpret_cast<const kfx_matrix_3f*>(&m); }
const kfx_point_3f* P(const point_3f& p) { return reinterpret_cast<const kfx_point_3f*>(&p); }
const kfx_point_3i* PI(const point_3i& p) { return reinterpret_cast<const kfx_point_3i*>(&p); }
const kfx_round_buffer* RB(const round_buffer* r) { return reinterpret_cast<const kfx_round_buffer*>(r); }

size_t npix(const intrinsics& i) { return (size_t)i.cols * (size_t)i.rows; }

}  // namespace

// ---- gpu_array_base / kernel handles ---------------------------------------

gpu_array_base::~gpu_array_base() {}

void gpu_array_base::dev_alloc() {
    if (device_ptr) return;
    check(kfx_init(0), "gpu_array_base::dev_alloc");
    check(kfx_malloc(&device_ptr, total_elems * elem_size()), "gpu_array_base::dev_alloc");
}

void gpu_array_base::dev_free() {
    if (!device_ptr) return;
    kfx_free(device_ptr);
    device_ptr = nullptr;
}

void gpu_array_base::memcpy_to_host(void* host_ptr) {
    check(kfx_memcpy_dtoh(host_ptr, device_ptr, total_elems * elem_size()), "gpu_array_base::memcpy_to_host");
}

void gpu_array_base::memcpy_from_host(void* host_ptr) {
    check(kfx_memcpy_htod(device_ptr, host_ptr, total_elems * elem_size()), "gpu_array_base::memcpy_from_host");
}

void gpu_array_base::copy_to(gpu_array_base* dst) {
    if (!dst || dst->get_total_size() * dst->elem_size() != total_elems * elem_size())
        throw std::runtime_error("gpu_array_base::copy_to: size mismatch");
    check(kfx_memcpy_dtod(dst->get_device_ptr(), device_ptr, total_elems * elem_size()), "gpu_array_base::copy_to");
}

void gpu_array_base::reset_gpu_part() {
    check(kfx_fill_u8(device_ptr, 0, total_elems * elem_size()), "gpu_array_base::reset_gpu_part");
}

opencl_kernel::opencl_kernel(const std::string& path_to_kernel_file) {
    // kernel name = file stem (opencl_helper.cpp:251-265)
    std::string s = path_to_kernel_file;
    const size_t slash = s.find_last_of('/');
    if (slash != std::string::npos) s = s.substr(slash + 1);
    const size_t dot = s.find_last_of('.');
    if (dot != std::string::npos) s = s.substr(0, dot);
    name = s;
}

kernel_manager::kernel_manager(const std::string& path_to_kernels, std::vector<std::string>& names)
    : path(path_to_kernels) {
    check(kfx_init(0), "kernel_manager");
    for (auto& n : names) {
        opencl_kernel* k = new opencl_kernel(n);
        kernels[k->get_name()] = k;
    }
}

kernel_manager::~kernel_manager() {
    for (auto& kv : kernels) delete kv.second;
}

opencl_kernel* kernel_manager::Get_Kernel(std::string& kernel_name) { return Get_Kernel(kernel_name.c_str()); }

opencl_kernel* kernel_manager::Get_Kernel(const char* kernel_name) {
    const std::string stem = opencl_kernel(kernel_name).get_name();
    auto it = kernels.find(stem);
    if (it == kernels.end()) throw std::runtime_error("kernel_manager::Get_Kernel: unknown kernel " + stem);
    return it->second;
}

// ---- per-step wrappers -------------------------------------------------------

void Integrate_Tsdf_Data(gpu_array_base* gDepth_Frame, gpu_array_base* gDepth_Frame_Scaled,
                         gpu_array_base* gTsdf_Volume, intrinsics& Intr, matrix_3f& Rotation_Inverse,
                         vector_3f& Translation, float Truncation_Distance, point_3f& Cell_Size,
                         round_buffer* Tsdf_Buffer, kernel_manager*, int) {
    const size_t n = npix(Intr);
    check(kfx_integrate_tsdf(dev<uint16_t>(gDepth_Frame, n, "Integrate_Tsdf_Data"),
                             dev<float>(gDepth_Frame_Scaled, n, "Integrate_Tsdf_Data"),
                             dev<kfx_tsdf>(gTsdf_Volume, (size_t)Tsdf_Buffer->buffer_length, "Integrate_Tsdf_Data"),
                             K(Intr), M(Rotation_Inverse), P(Translation), Truncation_Distance, P(Cell_Size),
                             RB(Tsdf_Buffer)),
          "Integrate_Tsdf_Data");
}

void Raycast(gpu_array_base* gTsdf_Volume, gpu_array_base* Vertex_Map, gpu_array_base* Normal_Map,
             round_buffer* Round_Buffer, vector_3f& Translation, matrix_3f& Rotation, float Truncation_Distance,
             intrinsics& Intr, point_3f& Cell_Size, opencl_kernel*) {
    const size_t n = npix(Intr);
    check(kfx_raycast(dev<kfx_tsdf>(gTsdf_Volume, (size_t)Round_Buffer->buffer_length, "Raycast"),
                      dev<kfx_xyz>(Vertex_Map, n, "Raycast"), dev<kfx_xyz>(Normal_Map, n, "Raycast"),
                      RB(Round_Buffer), P(Translation), M(Rotation), Truncation_Distance, K(Intr), P(Cell_Size)),
          "Raycast");
}

static void resize_map(gpu_array_base* in, gpu_array_base* out, intrinsics& Intr, int normalize, const char* who) {
    const size_t n = npix(Intr);
    check(kfx_resize_map(dev<kfx_xyz>(in, n, who), Intr.cols, Intr.rows, dev<kfx_xyz>(out, n / 4, who), normalize),
          who);
}

void Resize_Vertex_Map(gpu_array_base* Vertex_Map_In, gpu_array_base* Vertex_Map_Out, intrinsics& Intr,
                       opencl_kernel*) {
    resize_map(Vertex_Map_In, Vertex_Map_Out, Intr, 0, "Resize_Vertex_Map");
}

void Resize_Normal_Map(gpu_array_base* Normal_Map_In, gpu_array_base* Normal_Map_Out, intrinsics& Intr,
                       opencl_kernel*) {
    resize_map(Normal_Map_In, Normal_Map_Out, Intr, 1, "Resize_Normal_Map");
}

void Transform_Maps(gpu_array_base* Vertex_Map_Src, gpu_array_base* Normal_Map_Src, gpu_array_base* Vertex_Map_Dst,
                    gpu_array_base* Normal_Map_Dst, matrix_3f& Rotation, vector_3f& Translation, intrinsics& Intr,
                    opencl_kernel*) {
    const size_t n = npix(Intr);
    check(kfx_transform_maps(dev<kfx_xyz>(Vertex_Map_Src, n, "Transform_Maps"),
                             dev<kfx_xyz>(Normal_Map_Src, n, "Transform_Maps"), M(Rotation), P(Translation),
                             Intr.cols, Intr.rows, dev<kfx_xyz>(Vertex_Map_Dst, n, "Transform_Maps"),
                             dev<kfx_xyz>(Normal_Map_Dst, n, "Transform_Maps")),
          "Transform_Maps");
}

void Integrate_Colors(gpu_array_base* gTsdf_Color_Volume, gpu_array_base* gRGB_Frame, gpu_array_base* Vertex_Map,
                      intrinsics& Intr, matrix_3f& Rotation_Inverse, vector_3f& Translation,
                      float Truncation_Distance, int Max_Weight, point_3f Cell_Size, round_buffer* Round_Buffer,
                      opencl_kernel*) {
    const size_t n = npix(Intr);
    check(kfx_integrate_colors(dev<kfx_color>(gTsdf_Color_Volume, (size_t)Round_Buffer->buffer_length,
                                              "Integrate_Colors"),
                               dev<kfx_color>(gRGB_Frame, n, "Integrate_Colors"),
                               dev<kfx_xyz>(Vertex_Map, n, "Integrate_Colors"), K(Intr), M(Rotation_Inverse),
                               P(Translation), Truncation_Distance, Max_Weight, P(Cell_Size), RB(Round_Buffer)),
          "Integrate_Colors");
}

void Apply_Billateral_Filtration(gpu_array_base* gDepth, const intrinsics& Intr, float, opencl_kernel*,
                                 gpu_array_base* gDepth_Filtered) {
    const size_t n = npix(Intr);
    check(kfx_bilateral_filter(dev<uint16_t>(gDepth, n, "Apply_Billateral_Filtration"),
                               dev<uint16_t>(gDepth_Filtered, n, "Apply_Billateral_Filtration"), Intr.cols,
                               Intr.rows, DEFAULT_SIGMA_SPACE, DEFAULT_SIGMA_COLOR, 0xFFFFFFFFu),
          "Apply_Billateral_Filtration");
}

void Create_Pyramid(vector<gpu_array_base*>& depth_pyramids, int Num_Of_Pyramid_Level, intrinsics Intr,
                    kernel_manager*) {
    if ((int)depth_pyramids.size() < Num_Of_Pyramid_Level)
        throw std::runtime_error("Create_Pyramid: fewer buffers than levels");
    for (int i = 1; i < Num_Of_Pyramid_Level; ++i) {
        const intrinsics src = Intr(i - 1), dst = Intr(i);
        check(kfx_pyr_down(dev<uint16_t>(depth_pyramids[i - 1], npix(src), "Create_Pyramid"), src.cols, src.rows,
                           dev<uint16_t>(depth_pyramids[i], npix(dst), "Create_Pyramid"), DEFAULT_SIGMA_COLOR),
              "Create_Pyramid");
    }
}

void Create_Vertex_Maps(vector<gpu_array_base*>& depths, vector<gpu_array_base*>& vertexes, intrinsics Intr,
                        int Pyramid_Level_Num, kernel_manager*) {
    if ((int)depths.size() < Pyramid_Level_Num || (int)vertexes.size() < Pyramid_Level_Num)
        throw std::runtime_error("Create_Vertex_Maps: fewer buffers than levels");
    for (int i = 0; i < Pyramid_Level_Num; ++i) {
        const intrinsics li = Intr(i);
        check(kfx_vertex_map(dev<uint16_t>(depths[i], npix(li), "Create_Vertex_Maps"),
                             dev<kfx_xyz>(vertexes[i], npix(li), "Create_Vertex_Maps"), K(li)),
              "Create_Vertex_Maps");
    }
}

void Create_Normal_Maps(vector<gpu_array_base*>& vertexes, vector<gpu_array_base*>& normals, intrinsics Intr,
                        int Pyramid_Level_Num, kernel_manager*) {
    if ((int)vertexes.size() < Pyramid_Level_Num || (int)normals.size() < Pyramid_Level_Num)
        throw std::runtime_error("Create_Normal_Maps: fewer buffers than levels");
    for (int i = 0; i < Pyramid_Level_Num; ++i) {
        const intrinsics li = Intr(i);
        check(kfx_normal_map(dev<kfx_xyz>(vertexes[i], npix(li), "Create_Normal_Maps"),
                             dev<kfx_xyz>(normals[i], npix(li), "Create_Normal_Maps"), li.cols, li.rows),
              "Create_Normal_Maps");
    }
}

void Estimate_Combined(matrix_3f& Rotation_Current, vector_3f& Translation_Current, gpu_array_base* Vertex_Map_Current,
                       gpu_array_base* Normal_Map_Current, matrix_3f& Rotation_Previous_Inverse,
                       vector_3f& Translation_Previous, intrinsics& Intr, gpu_array_base* Vertex_Map_Previous,
                       gpu_array_base* Normal_Map_Previous, float Distance_Threshold, float Angle_Threshold,
                       gpu_array_base* gbuf, gpu_array_base* gMBuf, double* A, double* b, kernel_manager*) {
    const size_t n = npix(Intr);
    if (!gbuf || gbuf->get_total_size() < 27) throw std::runtime_error("Estimate_Combined: gbuf too small");
    const int stride = (int)(gbuf->get_total_size() / 27);
    check(kfx_estimate_combined(M(Rotation_Current), P(Translation_Current),
                                dev<kfx_xyz>(Vertex_Map_Current, n, "Estimate_Combined"),
                                dev<kfx_xyz>(Normal_Map_Current, n, "Estimate_Combined"), M(Rotation_Previous_Inverse),
                                P(Translation_Previous), K(Intr),
                                dev<kfx_xyz>(Vertex_Map_Previous, n, "Estimate_Combined"),
                                dev<kfx_xyz>(Normal_Map_Previous, n, "Estimate_Combined"), Distance_Threshold,
                                Angle_Threshold, dev<float>(gbuf, 27, "Estimate_Combined"), stride,
                                dev<float>(gMBuf, 27, "Estimate_Combined"), A, b),
          "Estimate_Combined");
}

void Clean_Buff(gpu_array_base* gPtr, opencl_kernel*) {
    if (!gPtr || gPtr->get_total_size() < 27) throw std::runtime_error("Clean_Buff: buffer too small");
    check(kfx_clean_buff(dev<float>(gPtr, 27, "Clean_Buff"), (int)(gPtr->get_total_size() / 27)), "Clean_Buff");
}

void Clean_Tsdf_Slice(point_3i Min_Bounds, point_3i Max_Bounds, round_buffer* Round_Buffer,
                      gpu_array_base* gTsdf_Volume, opencl_kernel*) {
    check(kfx_clean_tsdf_slice(PI(Min_Bounds), PI(Max_Bounds), RB(Round_Buffer),
                               dev<kfx_tsdf>(gTsdf_Volume, (size_t)Round_Buffer->buffer_length, "Clean_Tsdf_Slice")),
          "Clean_Tsdf_Slice");
}

void Clean_Color_Slice(point_3i Min_Bounds, point_3i Max_Bounds, round_buffer* Round_Buffer,
                       gpu_array_base* gColor_Volume, opencl_kernel*) {
    check(kfx_clean_color_slice(PI(Min_Bounds), PI(Max_Bounds), RB(Round_Buffer),
                                dev<kfx_color>(gColor_Volume, (size_t)Round_Buffer->buffer_length,
                                               "Clean_Color_Slice")),
          "Clean_Color_Slice");
}

void Extract_Slice(point_3i Min_Bounds, point_3i Max_Bounds, round_buffer* Round_Buffer, gpu_array_base* gTsdf_Volume,
                   gpu_array_base* gColor_Volume, point_3f, kernel_manager*, matrix<Point_XYZI>& Vertex_Slice,
                   matrix<Point_Color>& Color_Slice) {
    const size_t L = (size_t)Round_Buffer->buffer_length;
    kfx_slicer* s = nullptr;
    check(kfx_slicer_create(RB(Round_Buffer), &s), "Extract_Slice");
    const kfx_tsdf* vol = dev<kfx_tsdf>(gTsdf_Volume, L, "Extract_Slice");
    const kfx_color* col = gColor_Volume ? dev<kfx_color>(gColor_Volume, L, "Extract_Slice") : nullptr;
    uint64_t count = 0;
    int rc = kfx_extract_slice(s, PI(Min_Bounds), PI(Max_Bounds), RB(Round_Buffer), vol, col, nullptr, nullptr, 0,
                               &count);
    if (rc == KFX_OK && count == 0) {
        kfx_slicer_destroy(s);
        return;
    }
    if (rc != KFX_ERR_CAPACITY) {
        kfx_slicer_destroy(s);
        check(rc, "Extract_Slice");
    }
    void *dp = nullptr, *dc = nullptr;
    rc = kfx_malloc(&dp, sizeof(kfx_xyzi) * count);
    if (rc == KFX_OK) rc = kfx_malloc(&dc, sizeof(kfx_color) * count);
    if (rc == KFX_OK)
        rc = kfx_extract_slice(s, PI(Min_Bounds), PI(Max_Bounds), RB(Round_Buffer), vol, col,
                               static_cast<kfx_xyzi*>(dp), static_cast<kfx_color*>(dc), count, &count);
    matrix<Point_XYZI> part_v(count);
    matrix<Point_Color> part_c(count);
    if (rc == KFX_OK) rc = kfx_memcpy_dtoh(part_v.data, dp, sizeof(kfx_xyzi) * count);
    if (rc == KFX_OK) rc = kfx_memcpy_dtoh(part_c.data, dc, sizeof(kfx_color) * count);
    kfx_free(dp);
    kfx_free(dc);
    kfx_slicer_destroy(s);
    check(rc, "Extract_Slice");
    Vertex_Slice.extend_from(part_v);
    Color_Slice.extend_from(part_c);
}

void Extract_Tsdf_Volume(round_buffer* Round_Buffer, const point_3f& Cell_Size, gpu_array_base* gTsdf_Volume,
                         kernel_manager* Gpu_Methods, matrix<Point_XYZI>& Resulting_Cloud_Parallel) {
    // Every voxel is in the black zone when the x range spans the cube.
    const point_3i mn = make_point_3i(Round_Buffer->origin_GRID.x, 1, 1);
    const point_3i mx = make_point_3i(Round_Buffer->origin_GRID.x + Round_Buffer->voxels_size.x - 1, 0, 0);
    matrix<Point_XYZI> grid;
    matrix<Point_Color> unused;
    Extract_Slice(mn, mx, Round_Buffer, gTsdf_Volume, nullptr, Cell_Size, Gpu_Methods, grid, unused);
    const size_t n = grid.Get_Num_Of_Elem() * (grid.data != nullptr);
    if (n == 0) return;
    for (size_t i = 0; i < n; ++i) {
        XYZI& p = grid.data[i].data;
        p.x = (p.x + 0.5f) * Cell_Size.x;
        p.y = (p.y + 0.5f) * Cell_Size.y;
        p.z = (p.z + 0.5f) * Cell_Size.z;
    }
    Resulting_Cloud_Parallel.extend_from(grid);
}

void Extract_Color_Volume(gpu_array_base* gColor_Volume, round_buffer* Round_Buffer, matrix<Point_XYZI>& Point_Cloud,
                          point_3f Cell_Size, opencl_kernel*, matrix<Point_Color>& Resulting_Color_Point_Cloud) {
    const size_t n = Point_Cloud.Get_Num_Of_Elem() * (Point_Cloud.data != nullptr);
    matrix<Point_Color> out(n);
    if (n) {
        void *dp = nullptr, *dc = nullptr;
        int rc = kfx_malloc(&dp, sizeof(kfx_xyzi) * n);
        if (rc == KFX_OK) rc = kfx_malloc(&dc, sizeof(kfx_color) * n);
        if (rc == KFX_OK) rc = kfx_memcpy_htod(dp, Point_Cloud.data, sizeof(kfx_xyzi) * n);
        if (rc == KFX_OK)
            rc = kfx_extract_colors(dev<kfx_color>(gColor_Volume, (size_t)Round_Buffer->buffer_length,
                                                   "Extract_Color_Volume"),
                                    RB(Round_Buffer), static_cast<const kfx_xyzi*>(dp), n, P(Cell_Size),
                                    static_cast<kfx_color*>(dc));
        if (rc == KFX_OK) rc = kfx_memcpy_dtoh(out.data, dc, sizeof(kfx_color) * n);
        kfx_free(dp);
        kfx_free(dc);
        check(rc, "Extract_Color_Volume");
    }
    Resulting_Color_Point_Cloud.free_buffer();
    Resulting_Color_Point_Cloud.extend_from(out);
}
