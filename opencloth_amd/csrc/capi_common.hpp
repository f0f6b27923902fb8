// Helpers shared by the C ABI sources (capi*.cpp): the per-thread error
// message behind oc_last_error, HIP error propagation and arena alignment.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "opencloth.h"

namespace oc {

// Records msg as this thread's oc_last_error text and returns code
// (defined in capi.cpp).
int set_error(int code, const std::string& msg);

// Sub-allocations of a device arena start on 256-byte boundaries.
inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace oc

// Returns OC_ERR_DEVICE with the failing call and HIP's message.
#define OC_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return oc::set_error(OC_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
