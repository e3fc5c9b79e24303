// GRRTContext: the state behind the opaque handle of the C ABI.
#pragma once
#include "tonemap.hpp"
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "grrt/types.h"
#include "grrt_ext.h"
#include "scene_host.hpp"
#include "volumetric_host.hpp"
#include "trace_params.hpp"
#include "copy_pool.hpp"

namespace grrt_amd {

// Device buffer that grows on demand and is released with the context.
template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n) {
        if (n <= cap) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T));
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

hipError_t launch_reset(DevCounters* c, uint8_t* cls, long long n_rays, hipStream_t st);
// Lanes per ray of the spectral volumetric kernel for nbins bins (a power of
// two, kVolBinsPerLane bins per lane in registers), or 0 when nbins exceeds
// 64 * kVolBinsPerLane and the bins live in device scratch, one lane per ray.
int vol_group_size(int nbins);
// Persistent one-wavefront blocks: waves_per_simd (1 or 2) per SIMD.
hipError_t launch_trace_vol(const TraceParams* P, long long n_rays, int nbins, bool spectral, bool fp32_rt,
                            int num_cus, int waves_per_simd, hipStream_t st);
hipError_t launch_finalize_spectral(const double* spec, const uint8_t* cls, void* out, long long npix, int nS,
                                    int nb, bool f32, hipStream_t st);
// mode: 0 one view, 1 batch of views, 2 listed rays of one view.
// integrator: 0 adaptive Dormand-Prince, 1 step-doubling RK4.
hipError_t launch_trace_thin(const TraceParams* P, int mode, int blocks, int threads, hipStream_t st,
                             int integrator);
hipError_t launch_finalize(const double* rgb, const uint8_t* cls, float* out, long long npix, int nS,
                           hipStream_t st);
double flops_per_attempt();

namespace sg {
struct SurrogateState;
void destroy_state(SurrogateState* s);
}  // namespace sg

}  // namespace grrt_amd

struct GRRTContext {
    GRRTParams params{};
    grrt_amd::Resolved res{};
    std::string error_msg;
    std::vector<double> frequency_bins;

    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;           // second stream (default priority): surrogate fallback traces
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    hipEvent_t band_ev[4] = {nullptr, nullptr, nullptr, nullptr};   // cross-stream ordering, no timing
    hipEvent_t copy_ev[4] = {nullptr, nullptr, nullptr, nullptr};   // chunks of a frame copy, no timing

    // Scene tables in device memory.
    grrt_amd::CameraFrame cam{};
    double flux_max = 0.0;
    grrt_amd::DevBuf<double> flux, spec, star_tpb;
    grrt_amd::DevBuf<int> star_start;
    grrt_amd::DevBuf<float> bg_tex;          // equirectangular background, RGB, row 0 at theta = 0
    int bg_w = 0, bg_h = 0;
    std::string bg_path;                     // the file bg_tex was read from
    bool bg_on = false;                      // the scene samples bg_tex for escaped rays

    // Volumetric disk tables in device memory.
    grrt_amd::VolDev vol{};
    double vol_build_ms = 0.0;
    grrt_amd::DevBuf<double> vol_H, vol_zmax, vol_rho_mid, vol_rho_prof, vol_T_prof, vol_lnk, vol_kes;
    grrt_amd::DevBuf<float> vol_lnk_f, vol_kes_f;
    grrt_amd::DevBuf<int> vol_perm;
    grrt_amd::DevBuf<double> nu_obs, lognu_obs;
    grrt_amd::DevBuf<unsigned char> rt_scratch;
    grrt_amd::DevBuf<double> sample_spec;
    grrt_amd::DevBuf<unsigned char> spec_out;
    bool mixed_precision = false;        // FP32 radiative transfer in the volumetric path
    bool far_field_f32 = false;          // and FP32 geodesic steps outside r = 15 M
    int integrator = 0;                  // 0: adaptive Dormand-Prince, 1: step-doubling RK4

    // Per-render buffers.
    grrt_amd::DevBuf<double> sample_rgb;
    grrt_amd::DevBuf<uint8_t> sample_cls;
    grrt_amd::DevBuf<GRRTRayRecord> records;
    grrt_amd::DevBuf<float> frame;
    grrt_amd::DevBuf<grrt_amd::TonePartial> tone_part;   // device tone mapping: per-block partials
    grrt_amd::DevBuf<grrt_amd::ToneParams> tone_params;
    grrt_amd::DevBuf<grrt_amd::DevCounters> counters;
    grrt_amd::DevBuf<grrt_amd::TraceParams> params_dev;   // launch parameters
    grrt_amd::TraceParams* params_host = nullptr;          // pinned staging copy
    grrt_amd::DevCounters* counters_host = nullptr;        // pinned: counters read back after a launch
    // Host-mapped block: the cancel flag the kernels poll at each queue
    // claim, and the claim position they report back.
    volatile grrt_amd::HostShared* shared = nullptr;
    grrt_amd::HostShared* shared_dev = nullptr;      // device alias of shared
    volatile int* cancel_flag = nullptr;             // &shared->cancel
    // A render in flight: grrt_progress reads shared->claimed against
    // queue_len (progress_lo..progress_hi of the whole call).
    std::atomic<int> in_flight{0};
    std::atomic<long long> queue_len{1};
    float progress_lo = 0.0f, progress_hi = 1.0f;

    // Sampling and tuning.
    int sampling_mode = GRRT_SAMPLING_HASH;
    int sampling_spp = 1;
    uint64_t sampling_seed = 0;
    bool keep_records = false;
    bool records_valid = false;
    // Persistent launch shape of the frame tracer: 3 blocks of 256 per CU
    // (its 168 VGPRs fit 3 waves per SIMD), refill at 8 idle lanes, 64-ray
    // claims; on the Kerr a=0.998 1024^2 x 4 frame 17.1 ms against 17.4 ms
    // at 2 blocks and 18.1 ms at 2 blocks with 128-ray claims.
    int tune_block = 256;
    int tune_blocks_per_cu = 3;
    int tune_refill = 8;
    int tune_chunk = 64;
    int tune_tile = 0;
    int vol_waves_per_simd = 2;        // volumetric tracer: register budget and persistent waves per SIMD

    // Batched views: per-view tables and parameter blocks, and the captured
    // launch sequence, reused while the same views are rendered again.
    std::vector<GRRTView> batch_views;
    grrt_amd::DevBuf<grrt_amd::ViewParams> batch_params;
    grrt_amd::DevBuf<double> batch_flux;
    grrt_amd::DevBuf<float> batch_out;
    // The batch owns its launch block, samples and counters, so renders of
    // the context's own frame between replays leave the recorded graph valid.
    grrt_amd::DevBuf<grrt_amd::TraceParams> batch_tp;
    grrt_amd::DevBuf<double> batch_rgb;
    grrt_amd::DevBuf<uint8_t> batch_cls;
    grrt_amd::DevBuf<grrt_amd::DevCounters> batch_counters;
    grrt_amd::DevBuf<GRRTRayRecord> batch_records;   // with keep_records
    bool last_batch = false;            // the last render was a batch (classes come from batch_cls)
    hipGraph_t batch_graph = nullptr;
    hipGraphExec_t batch_exec = nullptr;
    long long batch_rays = 0, batch_pixels = 0;
    int batch_tuning[4] = {0, 0, 0, 0};

    // Learned surrogate (network, optimiser state, datasets), created on demand.
    grrt_amd::sg::SurrogateState* sg = nullptr;

    // Pinned host staging for frames returned to a caller's host buffer:
    // one DMA copy from the device frame, then a threaded host copy.
    float* host_stage = nullptr;
    size_t host_stage_cap = 0;          // floats
    std::unique_ptr<grrt_amd::CopyPool> copy_pool;

    GRRTStats last{};
    long long last_rays = 0;
    // Per-sample buffers of the last single-view render (the context's own,
    // or the allocation-free entry point's shared scratch).
    const uint8_t* last_cls = nullptr;
    const double* last_rgb = nullptr;
    mutable std::atomic<float> progress{1.0f};   // highest progress reported in the current call
};
