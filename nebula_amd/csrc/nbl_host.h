// nbl_host.h -- host-side plumbing shared by the C-ABI translation units:
// thread-local error text, HIP status checks, exception-to-status guard and
// an owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbl_hip.h"

namespace nbl {

void set_error(const std::string& msg);

struct hip_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct invalid_arg : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct no_device : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw std::bad_alloc();
    throw hip_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define NBL_HIP(call) ::nbl::hip_check((call), #call)

inline void require(bool cond, const char* msg)
{
    if (!cond) throw invalid_arg(msg);
}

// Runs f and maps exceptions to NBL_ERR_* codes (no exception crosses the ABI).
template<typename F>
int guarded(F&& f)
{
    try {
        f();
        return NBL_OK;
    } catch (const invalid_arg& e) {
        set_error(e.what());
        return NBL_ERR_INVALID;
    } catch (const no_device& e) {
        set_error(e.what());
        return NBL_ERR_NODEVICE;
    } catch (const std::bad_alloc&) {
        set_error("allocation failed (bad_alloc)");
        return NBL_ERR_ALLOC;
    } catch (const hip_error& e) {
        set_error(e.what());
        return NBL_ERR_HIP;
    } catch (const std::exception& e) {
        set_error(e.what());
        return NBL_ERR_RUNTIME;
    } catch (...) {
        set_error("unknown error");
        return NBL_ERR_RUNTIME;
    }
}

// Device allocation owned by the library.
template<typename T>
struct dbuf {
    T* p = nullptr;
    size_t n = 0;
    dbuf() = default;
    explicit dbuf(size_t count) { alloc(count); }
    dbuf(const dbuf&) = delete;
    dbuf& operator=(const dbuf&) = delete;
    dbuf(dbuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    dbuf& operator=(dbuf&& o) noexcept { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; return *this; }
    ~dbuf() { reset(); }
    void alloc(size_t count)
    {
        reset();
        if (count == 0) return;
        NBL_HIP(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
    }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
    }
    void upload(const T* host, size_t count, hipStream_t s = 0)
    {
        if (count) NBL_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* host, size_t count, hipStream_t s = 0) const
    {
        if (count) NBL_HIP(hipMemcpyAsync(host, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Fails with NBL_ERR_NODEVICE semantics if no HIP device is visible.
void ensure_device();

// Grid size for a grid-stride kernel over n items.
inline unsigned grid_for(size_t n, unsigned block, unsigned cap = 65535u * 4)
{
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace nbl
