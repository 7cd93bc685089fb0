// mi_noalloc.hip -- the fixed device scratch of the entry point that allocates nothing
// (mi_proj_find_projections_noalloc_device): a file-scope device array, one copy per GPU, held until the
// process ends.  Host code never uses the array's name as an address: the per-device address comes from
// hipGetSymbolAddress, taken once per device with that device current.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mi_noalloc.hpp"

namespace {

__device__ __align__(256) unsigned char scratch[mi::kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

} // namespace

namespace mi {

hipError_t noAllocScratch(int device, void** ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) {
            e = hipSetDevice(device);
            if (e != hipSuccess) return e;
        }
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) {
            const hipError_t e2 = hipSetDevice(cur);
            if (e == hipSuccess) e = e2;
        }
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return hipSuccess;
}

} // namespace mi
