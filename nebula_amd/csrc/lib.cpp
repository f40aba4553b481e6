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

extern "C" {

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
