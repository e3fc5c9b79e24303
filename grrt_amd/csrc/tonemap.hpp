// Device tone mapping of an RGBA float32 frame (src/tonemapper.cpp:7-71).
#pragma once
#include <hip/hip_runtime.h>

namespace grrt_amd {

// Per-block partial of the luminance statistics.
struct TonePartial {
    double log_sum;               // sum of ln(1e-6 + L) over lit pixels
    double lmax;                  // largest L
    unsigned long long lit;       // pixels with L > 1e-10
};

// Exposure and squared white point derived from the partials.
struct ToneParams {
    double exposure, lw2;
    int active;                   // 0: all-black frame, left unchanged
};

constexpr int kToneMaxBlocks = 1024;

// Blocks of the two per-pixel passes (and partials needed) for n pixels.
int tonemap_blocks(long long n_pixels);
// Tone-maps d_rgba (n_pixels RGBA float32, device memory) in place on `st`;
// part holds tonemap_blocks(n_pixels) partials, tp one ToneParams.
hipError_t launch_tonemap(float* d_rgba, long long n_pixels, TonePartial* part, ToneParams* tp, hipStream_t st);

}  // namespace grrt_amd
