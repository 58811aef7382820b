// opencl_kernel.h — kernel handle of the reference host API
// (common/inc/opencl_kernel.h).  Kernels are compiled into libkfx; the
// handle only carries the kernel's name so existing call sites keep their
// signatures.
#ifndef KFX_DROPIN_OPENCL_KERNEL
#define KFX_DROPIN_OPENCL_KERNEL

#include <string>

class opencl_kernel {
    std::string name;

  public:
    explicit opencl_kernel(const std::string& path_to_kernel_file);
    const std::string get_name() const { return name; }
};

#endif
