// Device-resident scratch of the allocation-free render entry point
// (grrt_ext_render_device_noalloc): a fixed array in the library's code
// object, one per GPU, held from the library's first use on that GPU until
// the process ends.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace grrt_amd {

// 128 MiB: the launch block, the counters and the per-sample colour and
// class buffers of one frame (25 bytes per sample, so up to ~5.3 M samples:
// a 1024^2 frame at 4 samples per pixel fits).
constexpr size_t kNoAllocScratchBytes = size_t{128} << 20;

// Resolves the array's address on `device` (which must be current); once
// per device, at context creation.
hipError_t noalloc_scratch_init(int device);
// The array's device address on `device`, or nullptr before the init call.
void* noalloc_scratch(int device);

}  // namespace grrt_amd
