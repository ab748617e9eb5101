// Fixed device scratch for the entry point that allocates nothing
// (mqw_clenshaw_noalloc).
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy, at its own address, from the first use of the library's code
// object on that device until the process ends.  In host code the name
// `scratch` is only a host-side stand-in; the device address of each copy is
// looked up once per device, with that device current, by
// hipGetSymbolAddress, and only that address is handed to kernels and copies.
//
// Every call on one GPU uses that GPU's one array, and the array keeps the
// last call's contents: calls that overlap on one GPU race.  Nothing in it
// needs to start at zero: each word a call reads, that call wrote first.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mqw_internal.h"

static __device__ __align__(256) unsigned char scratch[mqw::kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];      // device address of each device's copy
hipError_t g_err[kMaxDevices]; // result of the lookup
bool g_done[kMaxDevices];

} // namespace

namespace mqw {

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[device] = e == hipSuccess ? p : nullptr;
        g_err[device] = e;
        g_done[device] = true;
        if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
    }
    return g_err[device];
}

void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_done[device] ? g_ptr[device] : nullptr;
}

} // namespace mqw
