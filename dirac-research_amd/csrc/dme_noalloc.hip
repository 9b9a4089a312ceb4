// dme_noalloc.hip -- the device scratch array of the no-allocation entry
// point (dme_motion_estimate_noalloc, include/dme.h).
//
// `scratch` is a file-scope device array: every GPU the library runs on
// holds its own copy, at its own address, until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host the name is
// only a stand-in, not a device address); it passes the address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device when a context is created on it (dme_create).

#include <hip/hip_runtime.h>

#include <mutex>

#include "dme_internal.h"

namespace dme {

static __device__ __align__(256) unsigned char scratch[DME_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_lock;
void *g_addr[kMaxDevices] = {};
}  // namespace

// Address of this device's scratch array (device current on entry).
int noalloc_scratch(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return set_error(DME_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(DME_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_addr[device] = p;
    }
    *out = g_addr[device];
    return DME_OK;
}

}  // namespace dme
