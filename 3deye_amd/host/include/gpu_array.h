// gpu_array.h — typed device buffer (reference: common/inc/gpu_array.h:21-156).
#ifndef KFX_DROPIN_GPU_ARRAY
#define KFX_DROPIN_GPU_ARRAY

#include "gpu_array_base.h"

template <class T>
class gpu_array : public gpu_array_base {
  public:
    // Device buffer of dims.x * dims.y * dims.z elements.
    explicit gpu_array(dim3 dims, bool should_allocate = true) {
        sizes = dims;
        total_elems = dims.x * dims.y * dims.z;
        if (should_allocate) dev_alloc();
    }
    // Device buffer initialised from host data.
    gpu_array(T* host_data, dim3 dims) : gpu_array(dims, true) {
        if (host_data) memcpy_from_host(host_data);
    }
    ~gpu_array() override { dev_free(); }
    size_t elem_size() const override { return sizeof(T); }
    T* device_data() const { return static_cast<T*>(device_ptr); }
};

#endif
