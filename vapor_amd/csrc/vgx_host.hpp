// vgx_host.hpp - host-side error handling shared by the C ABI sources: C++
// failures inside the library become status codes and a thread-local message.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/vapor_mac3d.h"

namespace vgx_host {

// The calling thread's last error message (vgx_last_error).
std::string& error_slot();

inline int fail(int code, const std::string& message) {
    error_slot() = message;
    return code;
}

struct Failure {
    int code;
    std::string message;
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Failure{VGX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}
inline void invalid(const std::string& message) { throw Failure{VGX_EINVAL, message}; }

template <class F> int guarded(F&& f) {
    try {
        f();
        return VGX_OK;
    } catch (const Failure& failure) {
        return fail(failure.code, failure.message);
    } catch (const std::bad_alloc&) {
        return fail(VGX_EDEVICE, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(VGX_EDEVICE, e.what());
    }
}

template <class T> T* device_alloc(size_t count, const char* what) {
    void* p = nullptr;
    check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), what);
    return static_cast<T*>(p);
}

}  // namespace vgx_host
