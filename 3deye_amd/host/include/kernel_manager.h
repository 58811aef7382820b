// kernel_manager.h — name -> kernel registry of the reference host API
// (common/inc/kernel_manager.h:7-21).  Construction initialises the GPU
// (kfx_init on device 0); unknown kernel names throw.
#ifndef KFX_DROPIN_KERNEL_MANAGER
#define KFX_DROPIN_KERNEL_MANAGER

#include <map>
#include <string>
#include <vector>

#include "opencl_kernel.h"

class kernel_manager {
    std::string path;
    std::map<std::string, opencl_kernel*> kernels;

  public:
    kernel_manager(const std::string& path_to_kernels, std::vector<std::string>& names);
    ~kernel_manager();
    opencl_kernel* Get_Kernel(std::string& kernel_name);
    opencl_kernel* Get_Kernel(const char* kernel_name);
    std::string Get_Path() { return path; }
};

#endif
