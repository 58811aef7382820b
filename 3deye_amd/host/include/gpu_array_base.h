// gpu_array_base.h — device buffer handle of the reference host API
// (common/inc/gpu_array_base.h:17-123), backed by HIP memory through the kfx
// C ABI.  Sizes are 64-bit (the reference's unsigned int total_size
// overflows at 4 GiB).  Errors throw std::runtime_error with kfx_last_error().
#ifndef KFX_DROPIN_GPU_ARRAY_BASE
#define KFX_DROPIN_GPU_ARRAY_BASE

#include <stddef.h>

#include "dim3.h"

class gpu_array_base {
  protected:
    void* device_ptr = nullptr;
    dim3 sizes;
    size_t total_elems = 0;

  public:
    gpu_array_base() {}
    virtual ~gpu_array_base();

    virtual size_t elem_size() const = 0;
    // number of elements
    virtual size_t get_total_size() const { return total_elems; }
    virtual unsigned int width() const { return (unsigned int)sizes.width; }
    virtual unsigned int height() const { return (unsigned int)sizes.height; }
    virtual unsigned int depth() const { return (unsigned int)sizes.depth; }
    virtual dim3 get_size() const { return sizes; }

    void* get_device_ptr() const { return device_ptr; }

    virtual void dev_alloc();
    virtual void dev_free();
    // Blocking transfers of the whole buffer.
    virtual void memcpy_to_host(void* host_ptr);
    virtual void memcpy_from_host(void* host_ptr);
    // Device-to-device copy of the whole buffer into dst (same size).
    virtual void copy_to(gpu_array_base* dst);
    // Zero the device memory.
    virtual void reset_gpu_part();
};

#endif
