// kfx_runtime.hip — device selection, the library stream, memory and events.
//
// Replaces the reference's OpenCL runtime layer (common/src/opencl_helper.cpp,
// common/src/gpu_array_base.cpp): one HIP stream per process at default
// priority instead of a blocking OpenCL queue; device memory through
// hipMalloc with 64-bit sizes.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "kfx_common.h"

namespace kfx {

namespace {
thread_local char g_last_error[1024] = "";
hipStream_t g_stream = nullptr;
int g_device = -1;
}  // namespace

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

namespace {
struct ProfRecord {
    std::string name;
    hipEvent_t begin, end;
};
bool g_prof_on = false;
std::mutex g_prof_mu;
std::vector<ProfRecord> g_prof;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pooled_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

ProfScope::ProfScope(const char* name) : slot(-1) {
    if (!g_prof_on || name == nullptr) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(g_stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    hipEvent_t a = pooled_event(), b = pooled_event();
    if (!a || !b) return;
    if (hipEventRecord(a, g_stream) != hipSuccess) return;
    g_prof.push_back(ProfRecord{name, a, b});
    slot = (int)g_prof.size() - 1;
}

ProfScope::~ProfScope() {
    if (slot < 0) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    (void)hipEventRecord(g_prof[slot].end, g_stream);
}

hipStream_t stream() { return g_stream; }
int current_device() { return g_device; }
bool initialised() { return g_stream != nullptr; }

__global__ void fill_u8_kernel(uint8_t* __restrict__ dst, uint32_t word, size_t bytes) {
    // 16-byte stores for the aligned body, byte stores for the tail.
    const size_t n16 = bytes / 16;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    uint4 w4 = make_uint4(word, word, word, word);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        reinterpret_cast<uint4*>(dst)[i] = w4;
    for (size_t i = n16 * 16 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes; i += stride)
        dst[i] = (uint8_t)(word & 0xffu);
}

// Streaming copy (16-byte vector loads and stores, grid-stride): the
// reference point for achievable HBM bandwidth and for calibrating the
// profiler's FETCH_SIZE / WRITE_SIZE counters.
__global__ void __launch_bounds__(256) copy16_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace kfx

using namespace kfx;

namespace kfx {
// Every positive float with exponent in [e_lo, e_hi]: rcp_rn against the IEEE
// quotient 1/x.
__global__ void __launch_bounds__(256) check_reciprocal_kernel(int e_lo, unsigned long long total,
                                                               unsigned long long* __restrict__ bad) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float x = __uint_as_float(((uint32_t)(e_lo + 127 + (int)(i >> 23)) << 23) | (uint32_t)(i & 0x7FFFFFu));
    if (__float_as_uint(rcp_rn(x)) != __float_as_uint(1.0f / x)) atomicAdd(bad, 1ull);
}
}  // namespace kfx

extern "C" {

const char* kfx_last_error(void) { return g_last_error; }

int kfx_device_count(int* count) {
    KFX_REQUIRE(count != nullptr, "kfx_device_count: null output");
    KFX_CHECK_HIP(hipGetDeviceCount(count));
    return KFX_OK;
}

int kfx_init(int device) {
    int n = 0;
    KFX_CHECK_HIP(hipGetDeviceCount(&n));
    KFX_REQUIRE(device >= 0 && device < n, "kfx_init: device %d out of range (%d visible)", device, n);
    if (g_stream != nullptr && g_device == device) return KFX_OK;
    KFX_CHECK_HIP(hipSetDevice(device));
    if (g_stream != nullptr) {
        KFX_CHECK_HIP(hipStreamDestroy(g_stream));
        g_stream = nullptr;
    }
    KFX_CHECK_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_device = device;
    return noalloc_setup(device);
}

int kfx_check_reciprocal(int exp_lo, int exp_hi, uint64_t* mismatches) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(mismatches && exp_lo >= -126 && exp_hi <= 127 && exp_lo <= exp_hi,
                "kfx_check_reciprocal: exponents [%d, %d] outside [-126, 127]", exp_lo, exp_hi);
    unsigned long long* d = nullptr;
    KFX_CHECK_HIP(hipMallocAsync((void**)&d, sizeof(unsigned long long), stream()));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned long long), stream());
    const unsigned long long total = (unsigned long long)(exp_hi - exp_lo + 1) << 23;
    if (e == hipSuccess) {
        kfx::check_reciprocal_kernel<<<(unsigned)((total + 255) / 256), 256, 0, stream()>>>(exp_lo, total, d);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, stream());
    (void)hipFreeAsync(d, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    KFX_CHECK_HIP(e);
    *mismatches = h;
    return KFX_OK;
}

int kfx_synchronize(void) {
    KFX_REQUIRE_INIT();
    KFX_CHECK_HIP(hipStreamSynchronize(g_stream));
    return KFX_OK;
}

int kfx_get_stream(void** hip_stream) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(hip_stream != nullptr, "kfx_get_stream: null output");
    *hip_stream = (void*)g_stream;
    return KFX_OK;
}

int kfx_malloc(void** dev_ptr, size_t bytes) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(dev_ptr != nullptr, "kfx_malloc: null output");
    *dev_ptr = nullptr;
    if (bytes == 0) return KFX_OK;
    KFX_CHECK_HIP(hipMalloc(dev_ptr, bytes));
    return KFX_OK;
}

int kfx_free(void* dev_ptr) {
    KFX_REQUIRE_INIT();
    if (dev_ptr == nullptr) return KFX_OK;
    KFX_CHECK_HIP(hipStreamSynchronize(g_stream));
    KFX_CHECK_HIP(hipFree(dev_ptr));
    return KFX_OK;
}

int kfx_memcpy_htod(void* dev_dst, const void* host_src, size_t bytes) {
    KFX_REQUIRE_INIT();
    if (bytes == 0) return KFX_OK;
    KFX_REQUIRE(dev_dst && host_src, "kfx_memcpy_htod: null pointer");
    KFX_CHECK_HIP(hipMemcpyAsync(dev_dst, host_src, bytes, hipMemcpyHostToDevice, g_stream));
    KFX_CHECK_HIP(hipStreamSynchronize(g_stream));
    return KFX_OK;
}

int kfx_memcpy_dtoh(void* host_dst, const void* dev_src, size_t bytes) {
    KFX_REQUIRE_INIT();
    if (bytes == 0) return KFX_OK;
    KFX_REQUIRE(host_dst && dev_src, "kfx_memcpy_dtoh: null pointer");
    KFX_CHECK_HIP(hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, g_stream));
    KFX_CHECK_HIP(hipStreamSynchronize(g_stream));
    return KFX_OK;
}

int kfx_memcpy_dtod(void* dev_dst, const void* dev_src, size_t bytes) {
    KFX_REQUIRE_INIT();
    if (bytes == 0) return KFX_OK;
    KFX_REQUIRE(dev_dst && dev_src, "kfx_memcpy_dtod: null pointer");
    KFX_CHECK_HIP(hipMemcpyAsync(dev_dst, dev_src, bytes, hipMemcpyDeviceToDevice, g_stream));
    return KFX_OK;
}

int kfx_fill_u8(void* dev_dst, uint8_t value, size_t bytes) {
    KFX_REQUIRE_INIT();
    if (bytes == 0) return KFX_OK;
    KFX_REQUIRE(dev_dst != nullptr, "kfx_fill_u8: null pointer");
    KFX_REQUIRE(((uintptr_t)dev_dst & 15u) == 0, "kfx_fill_u8: destination must be 16-byte aligned");
    const uint32_t word = 0x01010101u * value;
    size_t blocks = (bytes / 16 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    fill_u8_kernel<<<(unsigned)blocks, 256, 0, g_stream>>>((uint8_t*)dev_dst, word, bytes);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_profile_enable(int on) {
    KFX_REQUIRE_INIT();
    g_prof_on = on != 0;
    return KFX_OK;
}

int kfx_profile_reset(void) {
    KFX_REQUIRE_INIT();
    KFX_CHECK_HIP(hipStreamSynchronize(g_stream));
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto& r : g_prof) {
        g_event_pool.push_back(r.begin);
        g_event_pool.push_back(r.end);
    }
    g_prof.clear();
    return KFX_OK;
}

int kfx_profile_query(const char* kernel, double* total_ms, uint64_t* launches) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(kernel && total_ms && launches, "kfx_profile_query: null argument");
    KFX_CHECK_HIP(hipStreamSynchronize(g_stream));
    std::lock_guard<std::mutex> lk(g_prof_mu);
    double sum = 0.0;
    uint64_t n = 0;
    for (auto& r : g_prof) {
        if (r.name != kernel) continue;
        float ms = 0.0f;
        KFX_CHECK_HIP(hipEventElapsedTime(&ms, r.begin, r.end));
        sum += ms;
        ++n;
    }
    *total_ms = sum;
    *launches = n;
    return KFX_OK;
}

int kfx_copy_kernel(void* dev_dst, const void* dev_src, size_t bytes) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(dev_dst && dev_src, "kfx_copy_kernel: null pointer");
    KFX_REQUIRE(((uintptr_t)dev_dst & 15u) == 0 && ((uintptr_t)dev_src & 15u) == 0 && bytes % 16 == 0,
                "kfx_copy_kernel: pointers and size must be multiples of 16 bytes");
    if (bytes == 0) return KFX_OK;
    ProfScope prof("copy");
    const size_t n16 = bytes / 16;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    size_t blocks = std::min((n16 + 255) / 256, (size_t)cus * 8);
    copy16_kernel<<<(unsigned)blocks, 256, 0, g_stream>>>((uint4*)dev_dst, (const uint4*)dev_src, n16);
    KFX_CHECK_LAUNCH();
    return KFX_OK;
}

int kfx_event_create(void** event) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(event != nullptr, "kfx_event_create: null output");
    hipEvent_t e;
    KFX_CHECK_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return KFX_OK;
}

int kfx_event_destroy(void* event) {
    KFX_REQUIRE_INIT();
    if (event == nullptr) return KFX_OK;
    KFX_CHECK_HIP(hipEventDestroy((hipEvent_t)event));
    return KFX_OK;
}

int kfx_event_record(void* event) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(event != nullptr, "kfx_event_record: null event");
    KFX_CHECK_HIP(hipEventRecord((hipEvent_t)event, g_stream));
    return KFX_OK;
}

int kfx_event_elapsed_ms(void* start, void* stop, float* ms) {
    KFX_REQUIRE_INIT();
    KFX_REQUIRE(start && stop && ms, "kfx_event_elapsed_ms: null argument");
    KFX_CHECK_HIP(hipEventSynchronize((hipEvent_t)stop));
    KFX_CHECK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return KFX_OK;
}

}  // extern "C"
