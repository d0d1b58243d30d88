// Host-side pieces shared by the library's translation units (runtime.hip,
// vmf.hip): the device handle, error reporting and allocation helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/xitils_amd.h"

#include <mutex>

struct MultiState;   // unified-batch buffers (runtime.hip)
struct RenderPool;   // xt_render's path buffers, shared by the device's scenes (runtime.hip)

struct xt_device {
    int ordinal = 0;
    int num_cu = 0;
    std::string arch;
    MultiState* multi = nullptr;        // created by the first unified xt_render_batch
    RenderPool* pool = nullptr;         // created by the first xt_render
    std::mutex pool_mu;                 // guards creating `pool`
    unsigned char* scratch = nullptr;   // this device's address of the static scratch array (xt_render_static)
};

// Records msg as this thread's last error (xt_last_error) and returns s.
xt_status xt_fail(xt_status s, const std::string& msg);

#define XT_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return xt_fail(XT_ERR_DEVICE, std::string(#call) + " failed: " + hipGetErrorString(e_));    \
    } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}
