// opencl_wrappers.h — the reference's per-step host API
// (common/inc/opencl_wrappers.h), same names, parameter order and meaning,
// implemented over the kfx C ABI (include/kfx.h) and its HIP kernels for
// gfx950.  Buffers are gpu_array<> handles in HIP memory; the kernel_manager
// / opencl_kernel arguments select nothing (the kernels are built into the
// library) and are accepted for source compatibility.  Work is ordered on the
// library stream; calls that hand results to the host (Estimate_Combined,
// Extract_Slice, gpu_array_base::memcpy_to_host) block.  Failures throw
// std::runtime_error.
//
// Only the V + C overload of Extract_Slice exists (the reference's
// Point_XYZI_Color overload throws unconditionally, opencl_wrappers.cpp:809-811).
#ifndef KFX_DROPIN_OPENCL_WRAPPERS
#define KFX_DROPIN_OPENCL_WRAPPERS

#include <vector>

#include "base_point_types.h"
#include "common_defines.h"
#include "elementary_points.h"
#include "gpu_array.h"
#include "intrinsics.h"
#include "kernel_manager.h"
#include "matrix.h"
#include "round_buffer.h"

using std::vector;

void Integrate_Tsdf_Data(gpu_array_base* gDepth_Frame, gpu_array_base* gDepth_Frame_Scaled,
                         gpu_array_base* gTsdf_Color_Volume, intrinsics& Intr, matrix_3f& Rotation_Inverse,
                         vector_3f& Translation, float Truncation_Distance, point_3f& Cell_Size,
                         round_buffer* Tsdf_Buffer, kernel_manager* Gpu_Methods, int DBG_Frame_Num);

void Raycast(gpu_array_base* gTsdf_Color_Volume, gpu_array_base* Vertex_Map, gpu_array_base* Normal_Map,
             round_buffer* Round_Buffer, vector_3f& Translation, matrix_3f& Rotation, float Truncation_Distance,
             intrinsics& Intr, point_3f& Cell_Size, opencl_kernel* raycast_kernel);

// Intr describes the input (finer) map; the output is half its size.
void Resize_Vertex_Map(gpu_array_base* Vertex_Map_In, gpu_array_base* Vertex_Map_Out, intrinsics& Intr,
                       opencl_kernel* resize_vertex_kernel);
void Resize_Normal_Map(gpu_array_base* Normal_Map_In, gpu_array_base* Normal_Map_Out, intrinsics& Intr,
                       opencl_kernel* resize_normal_kernel);

void Transform_Maps(gpu_array_base* Vertex_Map_Src, gpu_array_base* Normal_Map_Src, gpu_array_base* Vertex_Map_Dst,
                    gpu_array_base* Normal_Map_Dst, matrix_3f& Rotation, vector_3f& Translation, intrinsics& Intr,
                    opencl_kernel* transform_maps_kernel);

// Colour volume update (common/inc/opencl_wrappers.h:59-69); Vertex_Map is the
// raycast of the current pose in volume coordinates, gRGB_Frame cols x rows
// Color.  Max_Weight must lie in [1, 255].
void Integrate_Colors(gpu_array_base* gTsdf_Color_Volume, gpu_array_base* gRGB_Frame, gpu_array_base* Vertex_Map,
                      intrinsics& Intr, matrix_3f& Rotation_Inverse, vector_3f& Translation,
                      float Truncation_Distance, int Max_Weight, point_3f Cell_Size, round_buffer* Round_Buffer,
                      opencl_kernel* integrate_colors_kernel);

// Trunc_Dist is accepted and ignored: the reference binds its bits to the
// kernel's integer max-distance argument, where it never truncates
// (opencl_wrappers.cpp:231 vs compute_bilateral_filtering.cl:13).
void Apply_Billateral_Filtration(gpu_array_base* gDepth, const intrinsics& Intr, float Trunc_Dist,
                                 opencl_kernel* billateral_filtering_kernel, gpu_array_base* gDepth_Filtered);

void Create_Pyramid(vector<gpu_array_base*>& depth_pyramids, int Num_Of_Pyramid_Level, intrinsics Intr,
                    kernel_manager* Gpu_Methods);

void Create_Vertex_Maps(vector<gpu_array_base*>& depths, vector<gpu_array_base*>& vertexes, intrinsics Intr,
                        int Pyramid_Level_Num, kernel_manager* Gpu_Methods);

void Create_Normal_Maps(vector<gpu_array_base*>& vertexes, vector<gpu_array_base*>& normals, intrinsics Intr,
                        int Pyramid_Level_Num, kernel_manager* Gpu_Methods);

// gbuf holds 27 rows of gbuf->get_total_size() / 27 partials;
// gMBuf holds 27 floats.  A (6x6 row-major) and b (6) receive the normal
// equations (opencl_wrappers.cpp:426-437).
void Estimate_Combined(matrix_3f& Rotation_Current, vector_3f& Translation_Current, gpu_array_base* Vertex_Map_Current,
                       gpu_array_base* Normal_Map_Current, matrix_3f& Rotation_Previous_Inverse,
                       vector_3f& Translation_Previous, intrinsics& Intr, gpu_array_base* Vertex_Map_Previous,
                       gpu_array_base* Normal_Map_Previous, float Distance_Threshold, float Angle_Threshold,
                       gpu_array_base* gbuf, gpu_array_base* gMBuf, double* A, double* b,
                       kernel_manager* Gpu_Methods);

void Clean_Buff(gpu_array_base* gPtr, opencl_kernel* clean_buff_kernel);

void Clean_Tsdf_Slice(point_3i Min_Bounds, point_3i Max_Bounds, round_buffer* Round_Buffer,
                      gpu_array_base* gTsdf_Volume, opencl_kernel* clean_tsdf_kernel);
void Clean_Color_Slice(point_3i Min_Bounds, point_3i Max_Bounds, round_buffer* Round_Buffer,
                       gpu_array_base* gColor_Volume, opencl_kernel* clean_color_kernel);

// gColor_Volume may be NULL (no colour volume).
void Extract_Slice(point_3i Min_Bounds, point_3i Max_Bounds, round_buffer* Round_Buffer, gpu_array_base* gTsdf_Volume,
                   gpu_array_base* gColor_Volume, point_3f Cell_Size, kernel_manager* Gpu_Methods,
                   matrix<Point_XYZI>& Vertex_Slice, matrix<Point_Color>& Color_Slice);

// Whole-volume cloud (common/inc/opencl_wrappers.h:124-129; kernels
// scan_the_whole_array / extract_tsdf_cloud absent): Extract_Slice's zero
// crossings over the entire cube, in its order, as metric volume coordinates
// of the voxel centre ((g + 1/2) * cell), w = 1; appended to the cloud.
void Extract_Tsdf_Volume(round_buffer* Round_Buffer, const point_3f& Cell_Size, gpu_array_base* gTsdf_Volume,
                         kernel_manager* Gpu_Methods, matrix<Point_XYZI>& Resulting_Cloud_Parallel);
// Colours of a metric cloud (common/inc/opencl_wrappers.h:131-138): the
// colour of voxel floor(p / cell) for every point; the result has the
// cloud's length.
void Extract_Color_Volume(gpu_array_base* gColor_Volume, round_buffer* Round_Buffer, matrix<Point_XYZI>& Point_Cloud,
                          point_3f Cell_Size, opencl_kernel* extract_colors_kernel,
                          matrix<Point_Color>& Resulting_Color_Point_Cloud);

#endif
