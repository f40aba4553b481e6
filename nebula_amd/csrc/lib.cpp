// lib.cpp -- library-wide C-ABI entry points: error text, version, device.
#include "nbl_host.h"

namespace nbl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

void ensure_device()
{
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        throw no_device("no HIP device visible (hipGetDeviceCount: " + std::string(hipGetErrorString(e)) + ")");
}

}  // namespace nbl

namespace {
// Plain streaming copy: four independent 16-byte loads per lane in flight,
// then their stores, with non-temporal hints (the data is touched once).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const u32x4 a = __builtin_nontemporal_load(src + i);
        const u32x4 b = __builtin_nontemporal_load(src + i + stride);
        const u32x4 c = __builtin_nontemporal_load(src + i + 2 * stride);
        const u32x4 d = __builtin_nontemporal_load(src + i + 3 * stride);
        __builtin_nontemporal_store(a, dst + i);
        __builtin_nontemporal_store(b, dst + i + stride);
        __builtin_nontemporal_store(c, dst + i + 2 * stride);
        __builtin_nontemporal_store(d, dst + i + 3 * stride);
    }
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
}  // namespace

extern "C" {

int nbl_bench_copy(uint64_t bytes, int32_t reps, double* ms_per_copy)
{
    return nbl::guarded([&] {
        nbl::require(bytes >= 16 && bytes % 16 == 0 && reps >= 1, "nbl_bench_copy: bytes must be a multiple of 16, reps >= 1");
        nbl::ensure_device();
        const size_t n = bytes / 16;
        nbl::dbuf<u32x4> a(n), b(n);
        NBL_HIP(hipMemset(a.p, 1, bytes));
        hipEvent_t e0, e1;
        NBL_HIP(hipEventCreate(&e0)); NBL_HIP(hipEventCreate(&e1));
        NBL_HIP(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy, dim3(256 * 8 * 4), dim3(256), 0, 0, a.p, b.p, n);
        NBL_HIP(hipGetLastError());
        NBL_HIP(hipEventRecord(e1, 0));
        NBL_HIP(hipEventSynchronize(e1));
        float ms = 0;
        NBL_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
        if (ms_per_copy) *ms_per_copy = ms / reps;
    });
}

const char* nbl_last_error(void) { return nbl::g_last_error.c_str(); }

int nbl_version(int* major, int* minor, int* patch)
{
    // Feature level of the reference this library stands in for (config/config.h:35-37).
    if (major) *major = 1;
    if (minor) *minor = 0;
    if (patch) *patch = 2;
    return NBL_OK;
}

int nbl_device_info(char* name, size_t name_len, int* compute_units)
{
    return nbl::guarded([&] {
        nbl::ensure_device();
        int dev = 0;
        NBL_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        NBL_HIP(hipGetDeviceProperties(&p, dev));
        if (name && name_len) {
            std::strncpy(name, p.gcnArchName[0] ? p.gcnArchName : p.name, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (compute_units) *compute_units = p.multiProcessorCount;
    });
}

}  // extern "C"
