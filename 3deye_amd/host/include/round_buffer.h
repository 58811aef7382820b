// round_buffer.h — cyclic TSDF buffer state of the reference host API
// (common/inc/types/round_buffer.h; constructors as common/src/round_buffer.cpp:5-74).
#ifndef KFX_DROPIN_ROUND_BUFFER
#define KFX_DROPIN_ROUND_BUFFER

#include "dim3.h"
#include "elementary_points.h"

typedef struct round_buffer {
    int start_index;
    int end_index;
    int buffer_length;
    point_3i origin_GRID;
    point_3f origin_GRID_global;
    point_3f origin_metric;
    point_3f volume_size;
    point_3i voxels_size;

    // Defaults: 512^3 voxels over a 3 m cube at the origin.
    round_buffer() { set_default_values(); }
    round_buffer(size_t w, size_t h, size_t d) {
        set_default_values();
        set_voxels_size(w, h, d);
        set_memory_index();
    }
    round_buffer(const dim3& dims) {
        set_default_values();
        set_voxels_size(dims.width, dims.height, dims.depth);
        set_memory_index();
    }
    round_buffer(point_3i voxels, point_3f metric) {
        set_default_values();
        voxels_size = voxels;
        volume_size = metric;
        set_memory_index();
    }
    void set_default_values() {
        origin_GRID = make_point_3i(0, 0, 0);
        origin_GRID_global = make_point_3f(0.f, 0.f, 0.f);
        origin_metric = origin_GRID_global;
        voxels_size = make_point_3i(512, 512, 512);
        volume_size = make_point_3f(3.f, 3.f, 3.f);
        end_index = 0;
        set_memory_index();
    }
    void set_voxels_size(size_t w, size_t h, size_t d) {
        voxels_size.x = (int)w;
        voxels_size.y = (int)h;
        voxels_size.z = (int)d;
    }
    void set_memory_index() {
        start_index = 0;
        buffer_length = voxels_size.x * voxels_size.y * voxels_size.z;
    }
} round_buffer;

static_assert(sizeof(round_buffer) == sizeof(kfx_round_buffer), "round_buffer layout");

#endif
