// noalloc.hip -- the device memory of euchip_render_noalloc.
//
// Every per-frame scratch buffer of a render (setup records, vertex data, tile
// bins and their overflow pool, tile lists, counters) is carved from one array
// in the code object's own device memory, so a render through it allocates
// nothing.  Each GPU the library runs on holds its own copy of the array until
// the process ends.  The array's name is a host-side stand-in: host code uses
// only the address looked up here, once per device, with that device current.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

namespace euchip {

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices] = {};
}  // namespace

// The device address of `scratch` on `device`, which must be the current
// device; looked up on the first call for the device and kept.
hipError_t noalloc_scratch(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!ptr) return hipErrorInvalidSymbol;
        g_addr[device] = ptr;
    }
    *out = g_addr[device];
    return hipSuccess;
}

}  // namespace euchip
