// Launch parameters of the thin-disk trace kernel, shared by host and device.
// Passed by value as the kernel argument (well under the 4 KiB limit).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "grrt_ext.h"

namespace grrt_amd {

constexpr int kFluxN = 500;       // include/grrt/scene/accretion_disk.h:14 default LUT size
constexpr int kSpecN = 1000;      // include/grrt/color/spectrum.h:11 default entries
constexpr double kSpecTmin = 1000.0, kSpecTmax = 100000.0;
constexpr int kStarGT = 180, kStarGP = 360;   // include/grrt/scene/celestial_sphere.h:31-32
constexpr int kStarCells = kStarGT * kStarGP;
constexpr int kVolBinsPerLane = 4;  // spectral bins held in registers per lane of a ray's group
constexpr int kVolMaxR = 512;       // radial tables staged in LDS by the volumetric tracer (n_r = 500)

// Queue and statistics counters in device memory (reset by a kernel before
// every launch, so the launch sequence is graph-capturable).
struct DevCounters {
    unsigned long long queue_head;    // next unclaimed ray index
    unsigned long long steps;
    unsigned long long attempts;
    unsigned long long derivs;
    unsigned long long rays_done;
    unsigned long long march_steps;   // volumetric ray-march RK4 steps
    unsigned long long pad[2];
};

// Host-mapped, coherent block per context: written by the host (cancel) or
// the device (claimed), read by the other side while a render runs.
struct HostShared {
    int cancel;
    int pad;
    unsigned long long claimed;
};

// Pointers into device memory plus the scalars of one volumetric disk.
struct VolDev {
    double M, a;                         // the disk's own mass and spin (src/api.cpp:122-126)
    double r_horizon, r_min, r_outer, r_isco, taper_w, E_isco, L_isco;
    double rho_scale, noise_scale, turbulence;
    int octaves, n_r, n_z;
    const double *H, *z_max, *rho_mid, *rho_prof, *T_prof;   // n_r, n_r*n_z
    // Opacity tables: ln(max(kappa_abs, 1e-100)) [nu][rho][T] and kappa_es [rho][T],
    // in FP64 and rounded to FP32 (the FP32-transfer path reads those: half
    // the bytes, so the absorption table stays in the L2 cache).
    const double *ln_kabs, *kes;
    const float *ln_kabs_f, *kes_f;
    int n_nu, n_rho, n_T;
    double lnu0, lnu1, lrho0, lrho1, lT0, lT1;
    const int* perm;                     // 512 entries
    // Reciprocal scales, so the per-sample lookups multiply instead of
    // dividing: (n_r - 1)/(r_outer - r_min), 1/noise_scale, and (n - 1)/(hi - lo)
    // of the three log axes of the opacity tables.
    double r_scale, inv_noise, nu_scale, rho_scale_ax, T_scale;
};

// Everything that belongs to one camera view: the metric (spin varies per
// view in a batch), the camera, the pixel window and sample placement, and
// the thin disk (its inner edge follows the view's ISCO).
struct ViewParams {
    double M, a;
    double r_term;       // horizon radius + 0.01 (include/grrt/geodesic/geodesic_tracer.h:64)

    // Camera: static observer tetrad and the covariant metric there.
    int W, H;
    double r_obs, th_obs, ph_obs, fov;
    double e0t;                         // e0 = (e0t, 0, 0, 0)
    double e1[4], e2[4], e3[4];
    double g00, g03, g11, g22, g33;     // g_lower at the observer

    // Sampling: ray index -> (pixel in window, sample).
    int mode, grid, nS;
    double cell;
    uint64_t seed;
    int x0, y0, tw, th;                 // pixel window (full frame or tile)
    int tile;                           // queue order: 0 row-major, else tile x tile pixel blocks

    // Thin disk (src/accretion_disk.cpp).
    int disk;
    double r_in, r_out, t_peak, flux_max, ut_obs;
    const double* flux_lut;             // kFluxN

    // Place of this view in a launch: its first queue slot (queue order may
    // pad to whole tiles), its first sample in the pixel-major sample
    // buffers, its first output pixel, and its sample count.
    long long q_offset, ray_offset, pix_offset, n_rays;
};

// Rays of one view in queue order: row-major samples (tile == 0) or
// tile x tile pixel blocks, each block's samples together, so a wavefront's
// lanes trace neighbouring rays.  The tiled order pads the window to whole
// blocks; padded slots decode as invalid.
__host__ __device__ inline long long view_queue_length(const ViewParams& V) {
    const long long px = static_cast<long long>(V.tw) * V.th;
    if (V.tile <= 0) return px * V.nS;
    const long long nx = (V.tw + V.tile - 1) / V.tile, ny = (V.th + V.tile - 1) / V.tile;
    return nx * ny * V.tile * V.tile * V.nS;
}

struct TraceParams {
    ViewParams view;                    // the single view of a frame, tile or spectral render

    // Integrator and termination shared by all views.
    double r_escape;     // 1000 M (src/api.cpp:132)
    double tol;
    int max_steps;
    long long n_rays;    // rays in the launch (all views of a batch)

    // Batch of views (trace_batch): n_views entries sorted by ray_offset.
    const ViewParams* views;
    int n_views;

    const double* spec_lut;             // kSpecN * 3

    // Star background (src/celestial_sphere.cpp), CSR over the 180x360 grid.
    int stars;
    const int* star_start;              // kStarCells + 1
    const double* star_tpb;             // (theta, phi, brightness) in cell order
    // Texture background (GRRT_BG_TEXTURE): equirectangular linear RGB,
    // bg_w x bg_h, row 0 at theta = 0, column 0 at phi = -pi; nullptr: none.
    const float* bg_tex;
    int bg_w, bg_h;

    // Outputs.
    double* sample_rgb;                 // n_rays * 3
    uint8_t* sample_cls;                // n_rays
    GRRTRayRecord* records;             // n_rays or nullptr
    DevCounters* counters;
    const volatile int* cancel;         // host-mapped flag, may be nullptr
    // Host-mapped: the first ray of the latest claimed chunk, stored by the
    // claiming lane (a progress reading for the host; may be nullptr).
    unsigned long long* claimed;

    // Listed rays (trace mode 2): fallback ids, then judged ids whose results
    // are written after the view's n_rays samples; counts in device memory.
    const long long* fb_list;
    const long long* judge_list;
    const unsigned long long* list_counts;   // [0] fallbacks, [1] judged
    long long judge_cap;
    long long judge_out;                // sample index of the first judged result

    int refill_min;                     // idle lanes that trigger a refill
    int chunk;                          // rays claimed per queue atomic (>= 64)

    // Volumetric disk (the volumetric tracer only).
    VolDev vol;
    int nbins;                          // 3 for RGB, else spectral bins
    // Claim order within a row: position p of the row's rays is ray
    // (p * row_stride) mod row_rays, row_stride coprime to row_rays.
    long long row_stride;
    int far_f32;                        // FP32 transfer path: FP32 geodesic steps outside r = 15 M
    int integrator;                     // 0: adaptive Dormand-Prince (the reference tracer's); 1: step-doubling RK4
    const double* nu_obs;               // observer-frame frequencies [Hz]
    const double* lognu_obs;            // log10 of the same
    void* rt_scratch;                   // spectral accumulators, 4 * nbins * n_rays of RT (bins > 64 * kVolBinsPerLane)
    double* sample_spec;                // n_rays * nbins spectral samples
};

}  // namespace grrt_amd
