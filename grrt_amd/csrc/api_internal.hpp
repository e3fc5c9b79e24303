// Library-internal helpers shared by the C ABI implementation files.
#pragma once
#include <hip/hip_runtime.h>

#include <exception>
#include <string>
#include <vector>

#include "context.hpp"
#include "grrt_ext.h"

namespace grrt_api {

extern thread_local std::string g_last_error;

struct HipFailure : std::exception {
    std::string msg;
    explicit HipFailure(std::string m) : msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipFailure(std::string(what) + ": " + hipGetErrorString(e));
}

// Set-up time upload of a host table (synchronous: the source may be a temporary).
template <class T>
void upload(grrt_amd::DevBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
    (void)st;
    check(b.ensure(v.size()), "hipMalloc(table)");
    check(hipMemcpy(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy(table)");
}

// Runs f() and turns any exception into the context's error string and -1.
template <class F>
int guarded(GRRTContext* ctx, F&& f) {
    if (!ctx) { g_last_error = "null context"; return -1; }
    try {
        return f();
    } catch (const std::exception& e) {
        ctx->error_msg = e.what();
        return -1;
    } catch (...) {
        ctx->error_msg = "unknown error";
        return -1;
    }
}

// One view of a launch: metric, camera, window and sample placement, disk.
void fill_view(const GRRTContext* ctx, grrt_amd::ViewParams& V, double a, double r_obs, double th_obs,
               double ph_obs, double fov, int W, int H, const grrt_amd::CameraFrame& cam, const double* flux,
               double flux_max, double r_in, int x0, int y0, int tw, int th, int tile);
int samples_per_pixel(int mode, int spp, int& grid);

// Progress of the call in flight: the claimed fraction of the current
// launch's ray queue, mapped into [progress_lo, progress_hi] of the call and
// never below the last value reported.
float live_progress(const GRRTContext* ctx);

// Start of a rendering API call: clears the cancel request and the progress.
void begin_call(GRRTContext* ctx);

// The texture background of the context's scene into a launch block (or
// none).
void set_background(const GRRTContext* ctx, grrt_amd::TraceParams& P);

// Copies a device frame of n floats to the caller's host buffer through the
// context's pinned staging buffer (one DMA copy, then a threaded host copy).
void frame_to_host(GRRTContext* ctx, float* host, const float* dev, size_t n);

// Batched views: record (views -> graph) and replay.  seed_stride != 0
// gives view v the seeded-jitter key seed + v * seed_stride (default: all
// views share the context's key, as a single-view render would use).
void prepare_batch(GRRTContext* ctx, const GRRTView* views, int n, uint64_t seed_stride = 0);
int run_batch(GRRTContext* ctx);

}  // namespace grrt_api
