// The fixed scratch array behind grrt_ext_render_device_noalloc and the
// per-device lookup of its address.  In host code `scratch` names a
// host-side stand-in: only the address hipGetSymbolAddress returns on each
// GPU is used for kernel arguments and copies.
#include "noalloc.hpp"

#include <mutex>

namespace grrt_amd {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mutex);
    if (g_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = p;
    return hipSuccess;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_mutex);
    return g_ptr[device];
}

}  // namespace grrt_amd
