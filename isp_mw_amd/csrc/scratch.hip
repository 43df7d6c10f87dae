// Device scratch memory of the allocation-free entry point
// (isp_bm3d_plane_f32_noalloc).  No reference counterpart.
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy, at its own address, from the time its code object is loaded
// until the process ends.  Host code never uses the array's name as a device
// address (on the host it is only a stand-in symbol): scratch_address() asks
// the runtime for the current device's address once per device, at context
// creation, and callers keep that pointer.
#include "bm3d_internal.h"

#include <mutex>

namespace {

__device__ __align__(256) unsigned char scratch[ispbm3d::kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_addr[kMaxDevices] = {};

}  // namespace

namespace ispbm3d {

hipError_t scratch_address(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the device must be current
        void *p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
        g_addr[device] = p;
    }
    *out = g_addr[device];
    return hipSuccess;
}

}  // namespace ispbm3d
