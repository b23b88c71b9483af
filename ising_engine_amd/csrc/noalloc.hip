// noalloc.hip -- the fixed device workspace of the entry points that allocate
// nothing (ige_steepest_descent_noalloc).
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy for the life of the process.  Its device address is looked up
// once per device with that device current (ige_context_create calls
// noalloc_prepare), kept per device, and only that address is ever handed to
// kernels or copies: in host code the array's name is a host-side stand-in,
// not a device address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.hpp"

static __device__ __align__(256) unsigned char scratch[ige::kNoAllocScratchBytes];

namespace ige {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu[kMaxDevices];
void* g_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_prepare(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu[device]);
    if (g_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    g_ptr[device] = p;
    return hipSuccess;
}

void* noalloc_region(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu[device]);
    return g_ptr[device];
}

std::mutex& noalloc_mutex(int device) { return g_mu[device < 0 || device >= kMaxDevices ? 0 : device]; }

}  // namespace ige
