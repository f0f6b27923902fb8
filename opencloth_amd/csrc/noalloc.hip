// oc_cloth_step_noalloc: oc_cloth_step whose CG partials and scalars live in
// a fixed device array instead of a buffer the library allocates.  No
// counterpart in the reference.  It is for callers that cannot allocate
// device memory around the step (HIP graph capture, serving loops,
// harnesses that forbid hipMalloc): the call makes no device or host
// allocation.  The handle's own state (x, v, the CG vectors, the blocks and
// the status words) is allocated once by oc_cloth_create, before such a loop.
//
// The array (`scratch`) is a file-scope __device__ array: every GPU the
// library runs on holds its own copy until the process ends, at its own
// address, which hipGetSymbolAddress returns with that device current.  Host
// code uses only that address: the array's name is a host-side stand-in,
// not a device pointer.
//
// Every call on one GPU shares that GPU's array, and the array keeps the
// last call's contents: calls that overlap on one GPU (two streams, two host
// threads) race and must not be issued.  Nothing in it needs to start at
// zero: every solve writes each partial and scalar before it reads it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_common.hpp"
#include "oc_internal.hpp"

namespace {

// 4 MiB: the CG partials and scalars of one solve take
// 2 (tiles + 1)(cg_max_iters + 1) + 2 floats, e.g. 31,808 for a 512x512
// cloth (512 tiles) at 30 iterations.
constexpr size_t kScratchFloats = size_t(1) << 20;

static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device
std::mutex g_mu;

// Address of this device's copy of `scratch`, looked up once per device.
int scratch_address(int dev, float** out) {
    if (dev < 0 || dev >= kMaxDevices) return oc::set_error(OC_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        OC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return oc::set_error(OC_ERR_DEVICE, "hipGetSymbolAddress returned NULL");
        g_scratch[dev] = static_cast<float*>(p);
    }
    *out = g_scratch[dev];
    return OC_OK;
}

}  // namespace

extern "C" int oc_cloth_step_noalloc(oc_cloth cloth, int32_t nsteps) {
    if (!cloth) return oc::set_error(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (nsteps < 0) return oc::set_error(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    const size_t need = oc::cloth_partial_floats(cloth);
    if (need > kScratchFloats)
        return oc::set_error(OC_ERR_INVALID_ARGUMENT,
                             "the CG partials of this cloth need " + std::to_string(need) +
                                 " floats, more than the fixed scratch array holds (" +
                                 std::to_string(kScratchFloats) +
                                 "): 2 (tiles + 1)(cg_max_iters + 1) + 2 must fit");
    int dev = 0;
    OC_HIP(hipGetDevice(&dev));
    if (dev != oc::cloth_device(cloth))
        return oc::set_error(OC_ERR_INVALID_ARGUMENT, "the cloth's device is not the current device");
    float* part = nullptr;
    const int rc = scratch_address(dev, &part);
    if (rc != OC_OK) return rc;
    return oc::cloth_step(cloth, nsteps, part);
}
