// The GRRT C ABI (include/grrt/api.h) and its extensions (include/grrt_ext.h)
// over the HIP tracer.  Every entry point catches exceptions at the boundary
// and reports failures through grrt_error / grrt_last_error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <thread>

#include "api_internal.hpp"
#include "noalloc.hpp"
#include "writers_host.hpp"
#include "context.hpp"
#include "grrt/api.h"
#include "grrt_ext.h"

static_assert(sizeof(GRRTParams) == 240, "GRRTParams must keep the reference's LP64 layout");
static_assert(offsetof(GRRTParams, mass) == 16 && offsetof(GRRTParams, cam_yaw) == 64 &&
                  offsetof(GRRTParams, disk_enabled) == 88 && offsetof(GRRTParams, disk_volumetric) == 120 &&
                  offsetof(GRRTParams, disk_seed) == 144 && offsetof(GRRTParams, mass_solar) == 168 &&
                  offsetof(GRRTParams, background_type) == 184 &&
                  offsetof(GRRTParams, integrator_tolerance) == 200 &&
                  offsetof(GRRTParams, samples_per_pixel) == 212 && offsetof(GRRTParams, backend) == 220 &&
                  offsetof(GRRTParams, frequency_bins_hz) == 232,
              "GRRTParams field offsets must match the reference");
static_assert(sizeof(GRRTRayRecord) == 144, "GRRTRayRecord layout");

namespace grrt_api {

thread_local std::string g_last_error;


// Volumetric disk parameters exactly as the reference wires them
// (src/api.cpp:79-127): the disk takes the raw mass and spin fields, its
// own outer radius default (30 M), a peak temperature that may be derived
// from the black-hole mass and Eddington fraction, and an opacity frequency
// range widened to [0.1 nu_min, 10 nu_max] of the create-time bins.
struct VolSetup {
    double M, a, r_outer, t_peak;
    grrt_amd::VolParams vp;
};

VolSetup resolve_volumetric(const GRRTParams& p) {
    VolSetup v{};
    double T = p.disk_temperature > 0 ? p.disk_temperature : 1e7;
    if (p.mass_solar > 0.0 && p.eddington_fraction > 0.0 && p.disk_temperature <= 0.0) {
        // radiative efficiency 1 - E_isco (Bardeen, Press & Teukolsky 1972)
        const double as = std::abs(p.spin);
        const double z1 = 1.0 + std::cbrt(1.0 - as * as) * (std::cbrt(1.0 + as) + std::cbrt(1.0 - as));
        const double z2 = std::sqrt(3.0 * as * as + z1 * z1);
        const double risco = 3.0 + z2 - std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
        const double vi = 1.0 / std::sqrt(risco);
        const double E = (1.0 - 2.0 * vi * vi + as * vi * vi * vi) / std::sqrt(1.0 - 3.0 * vi * vi + 2.0 * as * vi * vi * vi);
        const double eta = 1.0 - E;
        T = 5.0e7 * std::pow(eta, 0.25) * std::pow(p.mass_solar, -0.25) * std::pow(p.eddington_fraction, 0.25);
    }
    v.M = p.mass;
    v.a = p.spin;
    v.r_outer = p.disk_outer > 0 ? p.disk_outer : 30.0;
    v.t_peak = T;
    v.vp.alpha = p.disk_alpha;
    v.vp.turbulence = p.disk_turbulence;
    v.vp.seed = static_cast<uint32_t>(p.disk_seed);
    v.vp.noise_scale = p.disk_noise_scale;
    v.vp.noise_octaves = p.disk_noise_octaves;
    if (p.num_frequency_bins > 0 && p.frequency_bins_hz) {
        double lo = p.frequency_bins_hz[0], hi = lo;
        for (int i = 1; i < p.num_frequency_bins; ++i) {
            lo = std::min(lo, p.frequency_bins_hz[i]);
            hi = std::max(hi, p.frequency_bins_hz[i]);
        }
        v.vp.opacity_nu_min = lo * 0.1;
        v.vp.opacity_nu_max = hi * 10.0;
    }
    return v;
}

void build_volumetric_tables(GRRTContext* ctx) {
    using namespace grrt_amd;
    const auto t0 = std::chrono::steady_clock::now();
    const VolSetup vs = resolve_volumetric(ctx->params);
    VolTables T;
    build_volumetric(vs.M, vs.a, vs.r_outer, vs.t_peak, vs.vp, T, &ctx->stream);
    std::vector<double> lnk(T.op.kappa_abs.size());
    for (size_t i = 0; i < lnk.size(); ++i) lnk[i] = std::log(std::max(T.op.kappa_abs[i], 1e-100));
    upload(ctx->vol_H, T.H, ctx->stream);
    upload(ctx->vol_zmax, T.z_max, ctx->stream);
    upload(ctx->vol_rho_mid, T.rho_mid, ctx->stream);
    upload(ctx->vol_rho_prof, T.rho_profile, ctx->stream);
    upload(ctx->vol_T_prof, T.T_profile, ctx->stream);
    upload(ctx->vol_lnk, lnk, ctx->stream);
    upload(ctx->vol_kes, T.op.kappa_es, ctx->stream);
    upload(ctx->vol_lnk_f, std::vector<float>(lnk.begin(), lnk.end()), ctx->stream);
    upload(ctx->vol_kes_f, std::vector<float>(T.op.kappa_es.begin(), T.op.kappa_es.end()), ctx->stream);
    if (T.n_r > kVolMaxR) throw HipFailure("volumetric radial tables larger than the tracer's LDS stage");
    upload(ctx->vol_perm, std::vector<int>(T.perm, T.perm + 512), ctx->stream);
    check(hipStreamSynchronize(ctx->stream), "upload volumetric tables");
    VolDev& V = ctx->vol;
    V.M = T.M; V.a = T.a;
    V.r_horizon = T.r_horizon; V.r_min = T.r_min; V.r_outer = T.r_outer; V.r_isco = T.r_isco;
    V.taper_w = T.taper_width; V.E_isco = T.E_isco; V.L_isco = T.L_isco;
    V.rho_scale = T.rho_scale; V.noise_scale = T.noise_scale; V.turbulence = T.prm.turbulence;
    V.octaves = T.prm.noise_octaves; V.n_r = T.n_r; V.n_z = T.n_z;
    V.H = ctx->vol_H.ptr; V.z_max = ctx->vol_zmax.ptr; V.rho_mid = ctx->vol_rho_mid.ptr;
    V.rho_prof = ctx->vol_rho_prof.ptr; V.T_prof = ctx->vol_T_prof.ptr;
    V.ln_kabs = ctx->vol_lnk.ptr; V.kes = ctx->vol_kes.ptr;
    V.ln_kabs_f = ctx->vol_lnk_f.ptr; V.kes_f = ctx->vol_kes_f.ptr;
    V.n_nu = T.op.n_nu; V.n_rho = T.op.n_rho; V.n_T = T.op.n_T;
    V.lnu0 = T.op.log_nu_min; V.lnu1 = T.op.log_nu_max; V.lrho0 = T.op.log_rho_min;
    V.lrho1 = T.op.log_rho_max; V.lT0 = T.op.log_T_min; V.lT1 = T.op.log_T_max;
    V.perm = ctx->vol_perm.ptr;
    V.r_scale = (V.n_r - 1) / (V.r_outer - V.r_min);
    V.inv_noise = 1.0 / V.noise_scale;
    V.nu_scale = (V.n_nu - 1) / (V.lnu1 - V.lnu0);
    V.rho_scale_ax = (V.n_rho - 1) / (V.lrho1 - V.lrho0);
    V.T_scale = (V.n_T - 1) / (V.lT1 - V.lT0);
    ctx->vol_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// (Re)build the scene tables from ctx->params.
void build_scene(GRRTContext* ctx) {
    using namespace grrt_amd;
    ctx->res = resolve_params(ctx->params);
    const Resolved& r = ctx->res;
    if (r.W <= 0 || r.H <= 0) throw HipFailure("width and height must be positive");
    ctx->cam = build_camera(r.M, r.a, r.r_obs, r.th_obs, r.ph_obs);
    std::vector<double> flux, spec;
    ctx->flux_max = 0.0;
    if (r.disk) ctx->flux_max = build_flux_table(r.M, r.a, r.r_in, r.r_out, flux);
    else flux.assign(kFluxN, 0.0);
    build_chromaticity_table(spec);
    upload(ctx->flux, flux, ctx->stream);
    upload(ctx->spec, spec, ctx->stream);
    if (r.stars && ctx->star_start.ptr == nullptr) {
        StarTable t;
        build_star_table(t);
        upload(ctx->star_start, t.start, ctx->stream);
        upload(ctx->star_tpb, t.tpb, ctx->stream);
    }
    const char* tex_path = ctx->params.background_texture_path;
    ctx->bg_on = false;
    if (r.texture && tex_path && *tex_path) {
        // The reference declares this mode (GRRT_BG_TEXTURE with
        // background_texture_path, include/grrt/types.h:15-17) and renders
        // it black (src/api.cpp:136 builds the sphere for GRRT_BG_STARS
        // only); with no path that stays so.  With a path the image is read
        // once per path and kept in device memory.
        const char* path = tex_path;
        if (ctx->bg_tex.ptr == nullptr || ctx->bg_path != path) {
            Texture tex;
            try {
                tex = read_texture(path);
            } catch (const std::exception& e) {
                throw HipFailure(e.what());
            }
            upload(ctx->bg_tex, tex.rgb, ctx->stream);
            ctx->bg_w = tex.width;
            ctx->bg_h = tex.height;
            ctx->bg_path = path;
        }
        ctx->bg_on = true;
    }
    if (r.volumetric) build_volumetric_tables(ctx);
    ctx->sampling_spp = r.spp;
    check(hipStreamSynchronize(ctx->stream), "upload tables");
}

int samples_per_pixel(int mode, int spp, int& grid);

// One view of the launch: metric, camera, window and sample placement, disk.
void fill_view(const GRRTContext* ctx, grrt_amd::ViewParams& V, double a, double r_obs, double th_obs,
               double ph_obs, double fov, int W, int H, const grrt_amd::CameraFrame& cam,
               const double* flux, double flux_max, double r_in, int x0, int y0, int tw, int th, int tile) {
    const grrt_amd::Resolved& r = ctx->res;
    V = grrt_amd::ViewParams{};
    V.M = r.M;
    V.a = a;
    V.r_term = grrt_amd::kerr_horizon(r.M, a) + 0.01;
    V.W = W;
    V.H = H;
    V.r_obs = r_obs;
    V.th_obs = th_obs;
    V.ph_obs = ph_obs;
    V.fov = fov;
    V.e0t = cam.e0[0];
    for (int m = 0; m < 4; ++m) {
        V.e1[m] = cam.e1[m];
        V.e2[m] = cam.e2[m];
        V.e3[m] = cam.e3[m];
    }
    V.g00 = cam.g[0][0];
    V.g03 = cam.g[0][3];
    V.g11 = cam.g[1][1];
    V.g22 = cam.g[2][2];
    V.g33 = cam.g[3][3];
    int grid = 1;
    V.nS = samples_per_pixel(ctx->sampling_mode, ctx->sampling_spp, grid);
    V.mode = ctx->sampling_mode;
    V.grid = grid;
    V.cell = grid > 0 ? 1.0 / grid : 1.0;
    V.seed = ctx->sampling_seed;
    V.x0 = x0;
    V.y0 = y0;
    V.tw = tw;
    V.th = th;
    V.tile = tile;
    V.disk = r.disk;
    V.r_in = r_in;
    V.r_out = r.r_out;
    V.t_peak = r.t_disk;
    V.flux_max = flux_max;
    V.ut_obs = 1.0 / std::sqrt(1.0 - 2.0 * r.M / r_obs);
    V.flux_lut = flux;
    V.n_rays = static_cast<long long>(tw) * th * V.nS;
}

int samples_per_pixel(int mode, int spp, int& grid) {
    if (mode == GRRT_SAMPLING_CENTER) { grid = 1; return 1; }
    if (mode == GRRT_SAMPLING_HASH) {
        grid = std::max(1, static_cast<int>(std::sqrt(static_cast<double>(std::max(spp, 1)))));
        return grid * grid;
    }
    grid = 0;
    return std::max(spp, 1);
}

float live_progress(const GRRTContext* ctx) {
    const double claimed = static_cast<double>(ctx->shared->claimed);
    const double f = std::clamp(claimed / static_cast<double>(ctx->queue_len.load()), 0.0, 1.0);
    const float p = ctx->progress_lo + (ctx->progress_hi - ctx->progress_lo) * static_cast<float>(f);
    // Claim positions from different wavefronts can land out of order:
    // report the largest value seen in this call (ctx->progress).
    float seen = ctx->progress.load();
    while (p > seen && !ctx->progress.compare_exchange_weak(seen, p)) {
    }
    return std::max(p, seen);
}

void set_background(const GRRTContext* ctx, grrt_amd::TraceParams& P) {
    const bool on = ctx->bg_on && ctx->bg_tex.ptr;
    P.bg_tex = on ? ctx->bg_tex.ptr : nullptr;
    P.bg_w = on ? ctx->bg_w : 0;
    P.bg_h = on ? ctx->bg_h : 0;
}

void begin_call(GRRTContext* ctx) {
    *ctx->cancel_flag = 0;
    ctx->progress.store(0.0f);
    ctx->progress_lo = 0.0f;
    ctx->progress_hi = 1.0f;
}

// Claim stride of the volumetric tracers within a row of L rays: the
// smallest integer above L/8 that is coprime to L, so that the row's claim
// order is a permutation and consecutive claims land about L/8 apart.
long long spread_stride(long long L) {
    long long m = L / 8 + 1;
    while (std::gcd(m, L) != 1) ++m;
    return m;
}

void frame_to_host(GRRTContext* ctx, float* host, const float* dev, size_t n) {
    if (ctx->host_stage_cap < n) {
        if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
        ctx->host_stage = nullptr;
        ctx->host_stage_cap = 0;
        void* p = nullptr;
        check(hipHostMalloc(&p, n * sizeof(float), hipHostMallocDefault), "hipHostMalloc(staging)");
        ctx->host_stage = static_cast<float*>(p);
        ctx->host_stage_cap = n;
    }
    if (!ctx->copy_pool) {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        ctx->copy_pool = std::make_unique<grrt_amd::CopyPool>(static_cast<int>(std::clamp(hw / 2, 1u, 8u)) - 1);
    }
    // The DMA copy into pinned staging runs in chunks; the host threads copy
    // chunk k into the caller's buffer while chunk k+1 is in flight.
    constexpr int kChunks = sizeof(ctx->copy_ev) / sizeof(ctx->copy_ev[0]);
    const size_t bytes = n * sizeof(float);
    const size_t chunk = ((bytes + kChunks - 1) / kChunks + 4095) & ~static_cast<size_t>(4095);
    char* const stage = reinterpret_cast<char*>(ctx->host_stage);
    const char* const src = reinterpret_cast<const char*>(dev);
    int issued = 0;
    for (size_t off = 0; off < bytes; off += chunk, ++issued) {
        check(hipMemcpyAsync(stage + off, src + off, std::min(chunk, bytes - off), hipMemcpyDeviceToHost, ctx->stream),
              "copy frame");
        check(hipEventRecord(ctx->copy_ev[issued], ctx->stream), "hipEventRecord");
    }
    for (int k = 0; k < issued; ++k) {
        const size_t off = static_cast<size_t>(k) * chunk;
        check(hipEventSynchronize(ctx->copy_ev[k]), "copy frame");
        ctx->copy_pool->copy(reinterpret_cast<char*>(host) + off, stage + off, std::min(chunk, bytes - off));
    }
}

enum class Output { kRGB, kSpectralF64, kSpectralF32 };

// Caller-provided device buffers for an RGB render (the allocation-free
// entry point): the launch block, counters and per-sample buffers.
struct FrameScratch {
    grrt_amd::TraceParams* params;
    grrt_amd::DevCounters* counters;
    uint8_t* cls;
    double* rgb;
};

// Render the pixel window [x0,x0+tw) x [y0,y0+th) into device memory `out`:
// RGBA float32 (tw*th*4), or tw*th*nbins spectral values (double or float).
// With fs (RGB only, no records) the render uses those buffers and
// allocates nothing.  Returns 0, or 1 if cancelled.
int render_window(GRRTContext* ctx, int x0, int y0, int tw, int th, void* d_out,
                  grrt_progress_fn progress, void* ud, Output kind = Output::kRGB,
                  const FrameScratch* fs = nullptr) {
    using namespace grrt_amd;
    const auto t_start = std::chrono::steady_clock::now();
    const Resolved& r = ctx->res;
    const bool spectral = kind != Output::kRGB;
    const bool vol = r.volumetric != 0;
    int grid = 1;
    const int nS = samples_per_pixel(ctx->sampling_mode, ctx->sampling_spp, grid);
    const long long npix = static_cast<long long>(tw) * th;
    const long long n_rays = npix * nS;
    const int nb = spectral ? static_cast<int>(ctx->frequency_bins.size()) : 3;
    if (fs && (spectral || ctx->keep_records)) throw HipFailure("caller scratch covers RGB renders without records");
    if (!fs) {
        check(ctx->sample_rgb.ensure(static_cast<size_t>(3 * n_rays)), "hipMalloc(samples)");
        check(ctx->sample_cls.ensure(static_cast<size_t>(n_rays)), "hipMalloc(classes)");
        if (ctx->keep_records) check(ctx->records.ensure(static_cast<size_t>(n_rays)), "hipMalloc(records)");
        check(ctx->counters.ensure(1), "hipMalloc(counters)");
        check(ctx->params_dev.ensure(1), "hipMalloc(params)");
    }
    TraceParams* const params_dev = fs ? fs->params : ctx->params_dev.ptr;

    TraceParams P{};
    fill_view(ctx, P.view, r.a, r.r_obs, r.th_obs, r.ph_obs, r.fov, r.W, r.H, ctx->cam, ctx->flux.ptr,
              ctx->flux_max, r.r_in, x0, y0, tw, th, spectral || vol ? 0 : ctx->tune_tile);
    P.r_escape = 1000.0;
    P.tol = r.tol;
    P.max_steps = r.max_steps;
    P.n_rays = view_queue_length(P.view);
    P.view.n_rays = n_rays;
    P.spec_lut = ctx->spec.ptr;
    P.stars = r.stars;
    P.star_start = ctx->star_start.ptr;
    P.star_tpb = ctx->star_tpb.ptr;
    set_background(ctx, P);
    P.sample_rgb = fs ? fs->rgb : ctx->sample_rgb.ptr;
    P.sample_cls = fs ? fs->cls : ctx->sample_cls.ptr;
    P.records = ctx->keep_records ? ctx->records.ptr : nullptr;
    P.counters = fs ? fs->counters : ctx->counters.ptr;
    P.cancel = &ctx->shared_dev->cancel;
    P.claimed = &ctx->shared_dev->claimed;
    P.refill_min = std::clamp(ctx->tune_refill, 1, 64);
    P.chunk = std::max(ctx->tune_chunk, 64);
    P.integrator = ctx->integrator;
    if (vol) {
        P.vol = ctx->vol;
        P.nbins = nb;
        P.row_stride = spread_stride(static_cast<long long>(tw) * nS);
        P.far_f32 = ctx->far_field_f32 ? 1 : 0;
        if (spectral) {
            std::vector<double> lg(nb);
            for (int k = 0; k < nb; ++k) lg[k] = std::log10(ctx->frequency_bins[k]);
            upload(ctx->nu_obs, ctx->frequency_bins, ctx->stream);
            upload(ctx->lognu_obs, lg, ctx->stream);
            const size_t rt = ctx->mixed_precision ? sizeof(float) : sizeof(double);
            if (vol_group_size(nb) == 0)
                check(ctx->rt_scratch.ensure(4 * static_cast<size_t>(nb) * n_rays * rt), "hipMalloc(rt scratch)");
            check(ctx->sample_spec.ensure(static_cast<size_t>(nb) * n_rays), "hipMalloc(spectra)");
            P.nu_obs = ctx->nu_obs.ptr;
            P.lognu_obs = ctx->lognu_obs.ptr;
            P.rt_scratch = ctx->rt_scratch.ptr;
            P.sample_spec = ctx->sample_spec.ptr;
        }
    }

    // The cancel flag is cleared once per API call (begin_call), not here:
    // a request made between the bands of one call must stop the call.
    ctx->shared->claimed = 0;
    ctx->queue_len.store(std::max<long long>(P.n_rays, 1));
    ctx->in_flight.store(1);
    const int threads = ctx->tune_block;
    const int blocks = std::max(1, ctx->num_cus * ctx->tune_blocks_per_cu);
    // The stream is idle here (every render ends with a synchronise), so the
    // pinned staging copy can be rewritten.
    *ctx->params_host = P;
    check(hipMemcpyAsync(params_dev, ctx->params_host, sizeof(TraceParams), hipMemcpyHostToDevice,
                         ctx->stream), "upload params");
    check(launch_reset(P.counters, P.sample_cls, n_rays, ctx->stream), "reset kernel");
    check(hipEventRecord(ctx->ev[0], ctx->stream), "hipEventRecord");
    if (vol) check(launch_trace_vol(params_dev, n_rays, nb, spectral, ctx->mixed_precision, ctx->num_cus,
                                    ctx->vol_waves_per_simd, ctx->stream),
                   "volumetric trace launch");
    else check(launch_trace_thin(params_dev, 0, blocks, threads, ctx->stream, ctx->integrator),
               "trace kernel launch");
    check(hipEventRecord(ctx->ev[1], ctx->stream), "hipEventRecord");
    if (spectral)
        check(launch_finalize_spectral(P.sample_spec, P.sample_cls, d_out, npix, nS, nb, kind == Output::kSpectralF32,
                                       ctx->stream),
              "finalize launch");
    else
        check(launch_finalize(P.sample_rgb, P.sample_cls, static_cast<float*>(d_out), npix, nS, ctx->stream),
              "finalize launch");
    check(hipEventRecord(ctx->ev[2], ctx->stream), "hipEventRecord");

    // Wait on the caller's thread.  With a callback, report the fraction of
    // the ray queue claimed so far (the persistent kernels store each claim
    // position into host-mapped memory), rising monotonically, as the
    // reference's renderer reports finished rows (src/renderer.cpp:74-81).
    if (progress) {
        float last = -1.0f;
        while (hipEventQuery(ctx->ev[2]) == hipErrorNotReady) {
            const float f = grrt_api::live_progress(ctx);
            if (f > last) {
                last = f;
                progress(f, ud);
            }
            std::this_thread::sleep_for(std::chrono::microseconds(500));
        }
    }
    const hipError_t sync = hipEventSynchronize(ctx->ev[2]);
    ctx->in_flight.store(0);
    check(sync, "render");
    check(hipMemcpyAsync(ctx->counters_host, P.counters, sizeof(DevCounters), hipMemcpyDeviceToHost, ctx->stream),
          "read counters");
    check(hipStreamSynchronize(ctx->stream), "read counters");
    const DevCounters c = *ctx->counters_host;
    float t_trace = 0.0f, t_fin = 0.0f;
    check(hipEventElapsedTime(&t_trace, ctx->ev[0], ctx->ev[1]), "event time");
    check(hipEventElapsedTime(&t_fin, ctx->ev[1], ctx->ev[2]), "event time");
    const bool cancelled = c.rays_done < static_cast<unsigned long long>(n_rays);
    GRRTStats& s = ctx->last;
    s = GRRTStats{};
    s.trace_ms = t_trace;
    s.finalize_ms = t_fin;
    s.rays = static_cast<int64_t>(c.rays_done);
    s.steps = static_cast<int64_t>(c.steps);
    s.attempts = static_cast<int64_t>(c.attempts);
    // one evaluation at each ray start plus six per attempt (k1 reused)
    s.deriv_evals = static_cast<int64_t>(c.rays_done + 6 * c.attempts);
    s.algo_flops = flops_per_attempt() * static_cast<double>(c.attempts);
    s.cancelled = cancelled ? 1 : 0;
    s.march_steps = static_cast<int64_t>(c.march_steps);
    s.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    ctx->last_rays = n_rays;
    ctx->last_batch = false;
    ctx->last_cls = P.sample_cls;
    ctx->last_rgb = P.sample_rgb;
    ctx->records_valid = ctx->keep_records && !cancelled;
    ctx->progress.store(ctx->progress_hi);
    if (progress) progress(ctx->progress_hi, ud);
    return cancelled ? 1 : 0;
}

// Spectral render of a width x height window into host memory.
int render_spectral_host(GRRTContext* ctx, void* host, int width, int height, bool f32,
                         grrt_progress_fn progress, void* ud) {
    if (ctx->frequency_bins.empty()) {
        ctx->error_msg = "No frequency bins set -- call grrt_set_frequency_bins first";
        return -1;
    }
    if (!host || width <= 0 || height <= 0) throw HipFailure("invalid spectral buffer");
    begin_call(ctx);
    const size_t n = static_cast<size_t>(width) * height * ctx->frequency_bins.size();
    const size_t bytes = n * (f32 ? sizeof(float) : sizeof(double));
    if (!ctx->res.volumetric) {
        // Without a volumetric disk every spectrum is zero
        // (src/geodesic_tracer.cpp:503-505).
        std::memset(host, 0, bytes);
        if (progress) progress(1.0f, ud);
        return 0;
    }
    check(ctx->spec_out.ensure(bytes), "hipMalloc(spectral out)");
    const int rc = render_window(ctx, 0, 0, width, height, ctx->spec_out.ptr, progress, ud,
                                 f32 ? Output::kSpectralF32 : Output::kSpectralF64);
    check(hipMemcpy(host, ctx->spec_out.ptr, bytes, hipMemcpyDeviceToHost), "copy spectra");
    return rc;
}

void tonemap_host(float* fb, int w, int h) {
    // Log-average auto-exposure with extended Reinhard and gamma 2.2
    // (src/tonemapper.cpp:7-71).
    const long long n = static_cast<long long>(w) * h;
    const double key = 0.18, delta = 1e-6;
    double log_sum = 0.0, lmax = 0.0;
    long long lit = 0;
    auto lum = [&](long long i) {
        const double R = std::max(0.0, static_cast<double>(fb[4 * i]));
        const double G = std::max(0.0, static_cast<double>(fb[4 * i + 1]));
        const double B = std::max(0.0, static_cast<double>(fb[4 * i + 2]));
        return 0.2126 * R + 0.7152 * G + 0.0722 * B;
    };
    for (long long i = 0; i < n; ++i) {
        const double L = lum(i);
        if (L > 1e-10) { log_sum += std::log(delta + L); ++lit; }
        if (L > lmax) lmax = L;
    }
    if (lmax <= 0.0 || lit == 0) return;
    const double exposure = key / std::exp(log_sum / lit);
    const double lw = lmax * exposure;
    const double lw2 = lw * lw;
    auto gam = [](double v) { return std::pow(std::clamp(v, 0.0, 1.0), 1.0 / 2.2); };
    for (long long i = 0; i < n; ++i) {
        const double L = lum(i);
        if (L <= 0.0) { fb[4 * i] = fb[4 * i + 1] = fb[4 * i + 2] = 0.0f; continue; }
        const double ls = L * exposure;
        const double lm = std::clamp(ls * (1.0 + ls / lw2) / (1.0 + ls), 0.0, 1.0);
        const double sc = lm / L;
        for (int c = 0; c < 3; ++c)
            fb[4 * i + c] = static_cast<float>(gam(std::max(0.0, static_cast<double>(fb[4 * i + c])) * sc));
    }
}

}  // namespace grrt_api

using namespace grrt_api;

namespace grrt_api {

// FITS metadata with the ABI's zero-means-default rules
// (src/api.cpp:296-302 of the reference).
grrt_amd::FitsMeta fits_meta(const GRRTParams& p) {
    grrt_amd::FitsMeta m;
    m.spin = p.spin;
    m.mass = p.mass > 0.0 ? p.mass : 1.0;
    m.observer_r = p.observer_r > 0.0 ? p.observer_r : 50.0;
    m.observer_theta = p.observer_theta > 0.0 ? p.observer_theta : 1.396;
    m.fov = p.fov > 0.0 ? p.fov : 1.047;
    m.samples_per_pixel = p.samples_per_pixel > 0 ? p.samples_per_pixel : 1;
    return m;
}

// Spectral cube streamed to a FITS file in bands of rows: each band is
// rendered on the GPU into device memory, copied back and written row by
// row at its place in the pre-extended file (BITPIX -64, or -32 for the
// float32 cube).  Returns 0, 1 if cancelled (unrendered rows stay zero), or
// -1 with the reference's argument errors.
int spectral_to_fits(GRRTContext* ctx, const char* path, int width, int height, bool f32,
                     grrt_progress_fn progress, void* ud) {
    if (ctx->frequency_bins.empty()) {
        ctx->error_msg = "No frequency bins set -- call grrt_set_frequency_bins first";
        return -1;
    }
    if (!path || path[0] == '\0') {
        ctx->error_msg = "grrt_render_spectral_to_fits_cb: output_path is null or empty";
        return -1;
    }
    if (width <= 0 || height <= 0) throw HipFailure("grrt_render_spectral_to_fits_cb: empty image");
    begin_call(ctx);
    const int nb = static_cast<int>(ctx->frequency_bins.size());
    grrt_amd::FitsCubeWriter out(path, width, height, nb, ctx->frequency_bins, fits_meta(ctx->params), f32 ? -32 : -64);
    int rc = 0;
    if (ctx->res.volumetric) {
        const size_t es = f32 ? sizeof(float) : sizeof(double);
        // About 2^18 spectral values per band.
        const int band = std::clamp((1 << 18) / std::max(width * nb, 1), 1, height);
        std::vector<unsigned char> host(static_cast<size_t>(width) * band * nb * es);
        check(ctx->spec_out.ensure(host.size()), "hipMalloc(spectral band)");
        for (int y0 = 0; y0 < height && rc == 0; y0 += band) {
            // A cancel request (from another thread or from the progress
            // callback between bands) ends the stream; the rows not yet
            // written stay zero in the pre-extended file.
            if (*ctx->cancel_flag) { rc = 1; break; }
            const int bh = std::min(band, height - y0);
            ctx->progress_lo = static_cast<float>(y0) / static_cast<float>(height);
            ctx->progress_hi = static_cast<float>(y0 + bh) / static_cast<float>(height);
            rc = render_window(ctx, 0, y0, width, bh, ctx->spec_out.ptr, nullptr, nullptr,
                               f32 ? Output::kSpectralF32 : Output::kSpectralF64);
            if (rc != 0) break;
            const size_t row_bytes = static_cast<size_t>(width) * nb * es;
            check(hipMemcpy(host.data(), ctx->spec_out.ptr, row_bytes * bh, hipMemcpyDeviceToHost), "copy spectra");
            for (int j = 0; j < bh; ++j) {
                if (f32) out.write_row(y0 + j, reinterpret_cast<const float*>(host.data() + row_bytes * j));
                else out.write_row(y0 + j, reinterpret_cast<const double*>(host.data() + row_bytes * j));
            }
            const float f = static_cast<float>(y0 + bh) / static_cast<float>(height);
            ctx->progress.store(f);
            if (progress) progress(f, ud);
        }
    } else if (progress) {
        // No volumetric disk: every spectrum is zero (src/geodesic_tracer.cpp:503-505),
        // which the pre-extended file already holds.
        progress(1.0f, ud);
    }
    out.finish();
    ctx->progress_lo = 0.0f;
    ctx->progress_hi = 1.0f;
    if (rc == 0) ctx->progress.store(1.0f);
    return rc;
}

// The CLI's two HDR outputs (cli/main.cpp:398-434) with the linear one
// streamed: the frame is rendered in bands of rows, and while the GPU renders
// band k+1 a writer thread encodes band k and appends its rows to the linear
// file and to the exposure statistics (in pixel order, so the scale equals a
// pass over the whole frame).  The exposure-scaled file is written from the
// finished frame.  A cancel request ends the stream after the current band:
// the linear file is completed with black rows, the rows not rendered are
// zero in the framebuffer, no scaled file is written, and the call returns 1.
int render_hdr(GRRTContext* ctx, const char* path_linear, const char* path_scaled, float* framebuffer,
               grrt_progress_fn progress, void* ud) {
    if (!path_linear || path_linear[0] == '\0') throw HipFailure("grrt_ext_render_hdr: linear output path is empty");
    const auto& r = ctx->res;
    const int W = r.W, H = r.H;
    begin_call(ctx);
    std::vector<float> own;
    float* fb = framebuffer;
    if (!fb) {
        own.assign(static_cast<size_t>(W) * H * 4, 0.0f);
        fb = own.data();
    }
    // About 2^18 pixels per band.
    const int band = std::clamp((1 << 18) / std::max(W, 1), 1, H);
    check(ctx->frame.ensure(4 * static_cast<size_t>(W) * band), "hipMalloc(frame band)");
    grrt_amd::RgbeWriter lin(path_linear, W, H);
    grrt_amd::ExposureSum exposure;
    std::thread writer;
    std::exception_ptr writer_error;
    // The writer thread is joined on every way out of this function.
    struct Join {
        std::thread& t;
        ~Join() {
            if (t.joinable()) t.join();
        }
    } join{writer};
    int rc = 0;
    int y = 0;
    for (; y < H; y += band) {
        if (*ctx->cancel_flag) { rc = 1; break; }
        const int bh = std::min(band, H - y);
        ctx->progress_lo = static_cast<float>(y) / static_cast<float>(H);
        ctx->progress_hi = static_cast<float>(y + bh) / static_cast<float>(H);
        rc = render_window(ctx, 0, y, W, bh, ctx->frame.ptr, nullptr, nullptr);
        if (rc != 0) break;
        if (writer.joinable()) writer.join();
        if (writer_error) std::rethrow_exception(writer_error);
        float* rows = fb + static_cast<size_t>(y) * W * 4;
        frame_to_host(ctx, rows, ctx->frame.ptr, 4 * static_cast<size_t>(W) * bh);
        writer = std::thread([&lin, &exposure, &writer_error, rows, W, bh] {
            try {
                exposure.add(rows, static_cast<long long>(W) * bh);
                lin.write_rows(rows, bh, 1.0f);
            } catch (...) {
                writer_error = std::current_exception();
            }
        });
        const float f = static_cast<float>(y + bh) / static_cast<float>(H);
        ctx->progress.store(f);
        if (progress) progress(f, ud);
    }
    if (writer.joinable()) writer.join();
    if (writer_error) std::rethrow_exception(writer_error);
    if (rc != 0 && y < H) std::fill(fb + static_cast<size_t>(y) * W * 4, fb + static_cast<size_t>(H) * W * 4, 0.0f);
    lin.finish();
    if (rc == 0 && path_scaled && path_scaled[0] != '\0') grrt_amd::write_rgbe(path_scaled, fb, W, H, exposure.scale());
    ctx->progress_lo = 0.0f;
    ctx->progress_hi = 1.0f;
    if (rc == 0) ctx->progress.store(1.0f);
    return rc;
}

}  // namespace grrt_api


extern "C" {

GRRTContext* grrt_create(const GRRTParams* params) {
    if (!params) { g_last_error = "grrt_create: params is null"; return nullptr; }
    GRRTContext* ctx = nullptr;
    try {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            g_last_error = "no HIP device available (this library has only the GPU backend)";
            return nullptr;
        }
        ctx = new GRRTContext();
        ctx->params = *params;
        check(hipGetDevice(&ctx->device), "hipGetDevice");
        check(hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
              "hipDeviceGetAttribute");
        check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& e : ctx->ev) check(hipEventCreate(&e), "hipEventCreate");
        for (auto& e : ctx->band_ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        for (auto& e : ctx->copy_ev) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        void* shared = nullptr;
        check(hipHostMalloc(&shared, sizeof(grrt_amd::HostShared), hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc");
        std::memset(shared, 0, sizeof(grrt_amd::HostShared));
        ctx->shared = static_cast<volatile grrt_amd::HostShared*>(shared);
        ctx->cancel_flag = &ctx->shared->cancel;
        void* dshared = nullptr;
        check(hipHostGetDevicePointer(&dshared, shared, 0), "hipHostGetDevicePointer");
        ctx->shared_dev = static_cast<grrt_amd::HostShared*>(dshared);
        void* staging = nullptr;
        check(hipHostMalloc(&staging, sizeof(grrt_amd::TraceParams), hipHostMallocDefault), "hipHostMalloc");
        ctx->params_host = static_cast<grrt_amd::TraceParams*>(staging);
        void* ch = nullptr;
        check(hipHostMalloc(&ch, sizeof(grrt_amd::DevCounters), hipHostMallocDefault), "hipHostMalloc");
        ctx->counters_host = static_cast<grrt_amd::DevCounters*>(ch);
        // The allocation-free entry point's scratch array on this GPU.
        check(grrt_amd::noalloc_scratch_init(ctx->device), "hipGetSymbolAddress(scratch)");
        build_scene(ctx);
        return ctx;
    } catch (const std::exception& e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "grrt_create: unknown error";
    }
    grrt_destroy(ctx);
    return nullptr;
}

void grrt_destroy(GRRTContext* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->flux.release();
    ctx->spec.release();
    ctx->star_tpb.release();
    ctx->star_start.release();
    ctx->sample_rgb.release();
    ctx->sample_cls.release();
    ctx->records.release();
    ctx->frame.release();
    ctx->counters.release();
    ctx->params_dev.release();
    if (ctx->batch_exec) (void)hipGraphExecDestroy(ctx->batch_exec);
    if (ctx->batch_graph) (void)hipGraphDestroy(ctx->batch_graph);
    grrt_amd::sg::destroy_state(ctx->sg);
    ctx->sg = nullptr;
    ctx->batch_records.release();
    ctx->batch_params.release();
    ctx->batch_tp.release();
    ctx->batch_rgb.release();
    ctx->batch_cls.release();
    ctx->batch_counters.release();
    ctx->batch_flux.release();
    ctx->batch_out.release();
    ctx->vol_H.release();
    ctx->vol_zmax.release();
    ctx->vol_rho_mid.release();
    ctx->vol_rho_prof.release();
    ctx->vol_T_prof.release();
    ctx->vol_lnk.release();
    ctx->vol_kes.release();
    ctx->vol_lnk_f.release();
    ctx->vol_kes_f.release();
    ctx->vol_perm.release();
    ctx->nu_obs.release();
    ctx->lognu_obs.release();
    ctx->rt_scratch.release();
    ctx->sample_spec.release();
    ctx->spec_out.release();
    if (ctx->params_host) (void)hipHostFree(ctx->params_host);
    if (ctx->counters_host) (void)hipHostFree(ctx->counters_host);
    if (ctx->shared) (void)hipHostFree(const_cast<grrt_amd::HostShared*>(ctx->shared));
    if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->band_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->copy_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void grrt_update_params(GRRTContext* ctx, const GRRTParams* params) {
    if (!ctx || !params) return;
    guarded(ctx, [&] {
        ctx->params = *params;
        ctx->batch_views.clear();      // tables may move: re-record the batch
        build_scene(ctx);
        return 0;
    });
}

int grrt_render(GRRTContext* ctx, float* framebuffer) {
    return grrt_render_cb(ctx, framebuffer, nullptr, nullptr);
}

int grrt_render_cb(GRRTContext* ctx, float* framebuffer, grrt_progress_fn progress, void* user_data) {
    return guarded(ctx, [&] {
        if (!framebuffer) throw HipFailure("framebuffer is null");
        const auto& r = ctx->res;
        const size_t npix = static_cast<size_t>(r.W) * r.H;
        const auto t0 = std::chrono::steady_clock::now();
        begin_call(ctx);
        check(ctx->frame.ensure(4 * npix), "hipMalloc(frame)");
        const int rc = render_window(ctx, 0, 0, r.W, r.H, ctx->frame.ptr, progress, user_data);
        const auto t1 = std::chrono::steady_clock::now();
        frame_to_host(ctx, framebuffer, ctx->frame.ptr, 4 * npix);
        const auto t2 = std::chrono::steady_clock::now();
        ctx->last.copy_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        ctx->last.total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
        return rc;
    });
}

int grrt_render_tile(GRRTContext* ctx, float* buffer, int x, int y, int tile_width, int tile_height) {
    return guarded(ctx, [&] {
        const auto& r = ctx->res;
        if (!buffer || tile_width <= 0 || tile_height <= 0 || x < 0 || y < 0 || x + tile_width > r.W ||
            y + tile_height > r.H)
            throw HipFailure("grrt_render_tile: tile outside the frame");
        const size_t npix = static_cast<size_t>(tile_width) * tile_height;
        begin_call(ctx);
        check(ctx->frame.ensure(4 * npix), "hipMalloc(frame)");
        const int rc = render_window(ctx, x, y, tile_width, tile_height, ctx->frame.ptr, nullptr, nullptr);
        frame_to_host(ctx, buffer, ctx->frame.ptr, 4 * npix);
        return rc;
    });
}

void grrt_cancel(GRRTContext* ctx) {
    if (ctx && ctx->cancel_flag) *ctx->cancel_flag = 1;
}

float grrt_progress(const GRRTContext* ctx) {
    if (!ctx) return 0.0f;
    return ctx->in_flight.load() ? grrt_api::live_progress(ctx) : ctx->progress.load();
}

const char* grrt_error(const GRRTContext* ctx) {
    return (!ctx || ctx->error_msg.empty()) ? nullptr : ctx->error_msg.c_str();
}

int grrt_cuda_available(void) {
    int n = 0;
    return (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
}

const char* grrt_last_error(void) { return g_last_error.empty() ? nullptr : g_last_error.c_str(); }

void grrt_tonemap(float* framebuffer, int width, int height) {
    if (framebuffer && width > 0 && height > 0) tonemap_host(framebuffer, width, height);
}

void grrt_debug_pixel(GRRTContext* ctx, int px, int py) {
    if (!ctx) { std::printf("grrt_debug_pixel: context not ready\n"); return; }
    guarded(ctx, [&] {
        const auto& r = ctx->res;
        if (px < 0 || py < 0 || px >= r.W || py >= r.H) throw HipFailure("pixel outside the frame");
        const int mode = ctx->sampling_mode;
        const bool keep = ctx->keep_records;
        ctx->sampling_mode = GRRT_SAMPLING_CENTER;
        ctx->keep_records = true;
        begin_call(ctx);
        check(ctx->frame.ensure(4), "hipMalloc(frame)");
        render_window(ctx, px, py, 1, 1, ctx->frame.ptr, nullptr, nullptr);
        ctx->sampling_mode = mode;
        ctx->keep_records = keep;
        GRRTRayRecord rec{};
        check(hipMemcpy(&rec, ctx->records.ptr, sizeof(rec), hipMemcpyDeviceToHost), "copy record");
        static const char* names[] = {"HORIZON", "ESCAPED", "MAX STEPS"};
        std::printf("Debugging pixel (%d, %d) in %dx%d image\n", px, py, r.W, r.H);
        std::printf("  -> %s after %d steps (%d attempts)  r=%.6f theta=%.6f phi=%.6f\n",
                    names[rec.term < 0 || rec.term > 2 ? 2 : rec.term], rec.steps, rec.attempts,
                    rec.x_final[1], rec.x_final[2], rec.x_final[3]);
        std::printf("     color=(%.4e %.4e %.4e) crossings=%d r_hit=%.6f g_hit=%.6f |H|max=%.3e\n",
                    rec.color[0], rec.color[1], rec.color[2], rec.n_cross, rec.r_hit, rec.g_hit, rec.h_max);
        return 0;
    });
}

void grrt_set_frequency_bins(GRRTContext* ctx, const double* frequencies_hz, int num_bins) {
    if (!ctx) return;
    if (num_bins > 0 && frequencies_hz) ctx->frequency_bins.assign(frequencies_hz, frequencies_hz + num_bins);
    else ctx->frequency_bins.clear();
}

int grrt_render_spectral(GRRTContext* ctx, double* spectral_buffer, int width, int height) {
    return grrt_render_spectral_cb(ctx, spectral_buffer, width, height, nullptr, nullptr);
}

int grrt_render_spectral_cb(GRRTContext* ctx, double* spectral_buffer, int width, int height,
                            grrt_progress_fn progress, void* user_data) {
    return guarded(ctx, [&] {
        return render_spectral_host(ctx, spectral_buffer, width, height, false, progress, user_data);
    });
}

int grrt_render_spectral_to_fits_cb(GRRTContext* ctx, const char* output_path, int width, int height,
                                    grrt_progress_fn progress, void* user_data) {
    return guarded(ctx, [&] { return spectral_to_fits(ctx, output_path, width, height, false, progress, user_data); });
}

// ------------------------------------------------------------ extensions ---

int grrt_ext_set_sampling(GRRTContext* ctx, int mode, int spp, uint64_t seed) {
    return guarded(ctx, [&] {
        if (mode < GRRT_SAMPLING_HASH || mode > GRRT_SAMPLING_SEEDED) throw HipFailure("unknown sampling mode");
        ctx->sampling_mode = mode;
        ctx->sampling_spp = std::max(spp, 1);
        ctx->sampling_seed = seed;
        ctx->batch_views.clear();      // sample placement is recorded with the batch
        return 0;
    });
}

int grrt_ext_set_records(GRRTContext* ctx, int enable) {
    return guarded(ctx, [&] {
        ctx->keep_records = enable != 0;
        ctx->batch_views.clear();      // the records buffer is recorded with the batch
        ctx->records_valid = false;
        return 0;
    });
}

int64_t grrt_ext_num_samples(const GRRTContext* ctx) {
    if (!ctx) return -1;
    if (ctx->last_batch) return ctx->last_rays;
    int grid = 1;
    const int nS = samples_per_pixel(ctx->sampling_mode, ctx->sampling_spp, grid);
    return static_cast<int64_t>(ctx->res.W) * ctx->res.H * nS;
}

int grrt_ext_get_records(GRRTContext* ctx, GRRTRayRecord* out, int64_t n) {
    return guarded(ctx, [&] {
        if (!ctx->records_valid) throw HipFailure("no records: enable them before rendering");
        if (!out || n < 0 || n > ctx->last_rays) throw HipFailure("record count out of range");
        const GRRTRayRecord* src = ctx->last_batch ? ctx->batch_records.ptr : ctx->records.ptr;
        check(hipMemcpy(out, src, static_cast<size_t>(n) * sizeof(GRRTRayRecord), hipMemcpyDeviceToHost),
              "copy records");
        return 0;
    });
}

int grrt_ext_get_classes(GRRTContext* ctx, uint8_t* out, int64_t n) {
    return guarded(ctx, [&] {
        if (!out || n < 0 || n > ctx->last_rays) throw HipFailure("class count out of range");
        const uint8_t* src = ctx->last_batch ? ctx->batch_cls.ptr : (ctx->last_cls ? ctx->last_cls : ctx->sample_cls.ptr);
        check(hipMemcpy(out, src, static_cast<size_t>(n), hipMemcpyDeviceToHost), "copy classes");
        return 0;
    });
}

int grrt_ext_get_samples(GRRTContext* ctx, double* rgb, int64_t n) {
    return guarded(ctx, [&] {
        if (!rgb || n < 0 || n > ctx->last_rays) throw HipFailure("sample count out of range");
        const double* src = ctx->last_batch ? ctx->batch_rgb.ptr : (ctx->last_rgb ? ctx->last_rgb : ctx->sample_rgb.ptr);
        check(hipMemcpy(rgb, src, static_cast<size_t>(n) * 3 * sizeof(double), hipMemcpyDeviceToHost), "copy samples");
        return 0;
    });
}

int grrt_ext_last_stats(const GRRTContext* ctx, GRRTStats* out) {
    if (!ctx || !out) return -1;
    *out = ctx->last;
    return 0;
}

int grrt_ext_render_device(GRRTContext* ctx, float* d_rgba) {
    return guarded(ctx, [&] {
        const auto& r = ctx->res;
        float* dst = d_rgba;
        if (!dst) {
            check(ctx->frame.ensure(4 * static_cast<size_t>(r.W) * r.H), "hipMalloc(frame)");
            dst = ctx->frame.ptr;
        }
        begin_call(ctx);
        return render_window(ctx, 0, 0, r.W, r.H, dst, nullptr, nullptr);
    });
}

int grrt_ext_render_device_noalloc(GRRTContext* ctx, float* d_rgba) {
    return guarded(ctx, [&] {
        if (!d_rgba) throw HipFailure("grrt_ext_render_device_noalloc: d_rgba must be a device buffer of width*height*4 floats");
        if (ctx->keep_records) throw HipFailure("grrt_ext_render_device_noalloc: per-ray records are not kept by this entry point");
        const auto& r = ctx->res;
        unsigned char* base = static_cast<unsigned char*>(grrt_amd::noalloc_scratch(ctx->device));
        if (!base) throw HipFailure("grrt_ext_render_device_noalloc: no scratch on this device");
        int grid = 1;
        const long long n = static_cast<long long>(r.W) * r.H * samples_per_pixel(ctx->sampling_mode, ctx->sampling_spp, grid);
        auto up = [](size_t x) { return (x + 255) & ~size_t{255}; };
        const size_t o_cnt = up(sizeof(grrt_amd::TraceParams));
        const size_t o_cls = o_cnt + up(sizeof(grrt_amd::DevCounters));
        const size_t o_rgb = o_cls + up(static_cast<size_t>(n));
        const size_t need = o_rgb + static_cast<size_t>(n) * 3 * sizeof(double);
        if (need > grrt_amd::kNoAllocScratchBytes)
            throw HipFailure("grrt_ext_render_device_noalloc: " + std::to_string(n) + " samples need " +
                             std::to_string(need >> 20) + " MiB of scratch; this entry point holds at most " +
                             std::to_string(grrt_amd::kNoAllocScratchBytes >> 20) +
                             " MiB (use grrt_ext_render_device)");
        const FrameScratch fs{reinterpret_cast<grrt_amd::TraceParams*>(base),
                              reinterpret_cast<grrt_amd::DevCounters*>(base + o_cnt), base + o_cls,
                              reinterpret_cast<double*>(base + o_rgb)};
        begin_call(ctx);
        return render_window(ctx, 0, 0, r.W, r.H, d_rgba, nullptr, nullptr, Output::kRGB, &fs);
    });
}

int grrt_ext_tonemap_device(GRRTContext* ctx, float* d_rgba, double* gpu_ms) {
    return guarded(ctx, [&] {
        const auto& r = ctx->res;
        const long long npix = static_cast<long long>(r.W) * r.H;
        float* dst = d_rgba;
        if (!dst) {
            if (!ctx->frame.ptr || ctx->frame.cap < static_cast<size_t>(4 * npix))
                throw HipFailure("no frame in the context: render it with grrt_ext_render_device(ctx, NULL)");
            dst = ctx->frame.ptr;
        }
        check(ctx->tone_part.ensure(static_cast<size_t>(grrt_amd::tonemap_blocks(npix))), "hipMalloc(tone partials)");
        check(ctx->tone_params.ensure(1), "hipMalloc(tone params)");
        check(hipEventRecord(ctx->ev[0], ctx->stream), "hipEventRecord");
        check(grrt_amd::launch_tonemap(dst, npix, ctx->tone_part.ptr, ctx->tone_params.ptr, ctx->stream), "tone map launch");
        check(hipEventRecord(ctx->ev[1], ctx->stream), "hipEventRecord");
        check(hipEventSynchronize(ctx->ev[1]), "tone map");
        if (gpu_ms) {
            float ms = 0.0f;
            check(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "event time");
            *gpu_ms = ms;
        }
        return 0;
    });
}

int grrt_ext_get_frame(GRRTContext* ctx, float* out) {
    return guarded(ctx, [&] {
        const auto& r = ctx->res;
        const size_t n = 4 * static_cast<size_t>(r.W) * r.H;
        if (!out) throw HipFailure("null output");
        if (!ctx->frame.ptr || ctx->frame.cap < n) throw HipFailure("no frame in the context");
        check(hipMemcpy(out, ctx->frame.ptr, n * sizeof(float), hipMemcpyDeviceToHost), "copy frame");
        return 0;
    });
}

int grrt_ext_set_tuning(GRRTContext* ctx, int block, int blocks_per_cu, int refill_min, int chunk, int tile) {
    return guarded(ctx, [&] {
        if (block > 0) {
            if (block % 64 != 0 || block > 256) throw HipFailure("block must be a multiple of 64, at most 256");
            ctx->tune_block = block;
        }
        if (blocks_per_cu > 0) ctx->tune_blocks_per_cu = blocks_per_cu;
        if (refill_min > 0) ctx->tune_refill = std::min(refill_min, 64);
        if (chunk > 0) {
            if (chunk < 64) throw HipFailure("chunk must be at least 64 rays");
            ctx->tune_chunk = chunk;
        }
        if (tile >= 0) {
            if (tile == 1 || tile > 32) throw HipFailure("tile must be 0 (row order) or 2..32 pixels");
            ctx->tune_tile = tile;
        }
        return 0;
    });
}

int grrt_ext_set_vol_tuning(GRRTContext* ctx, int waves_per_simd) {
    return guarded(ctx, [&] {
        if (waves_per_simd < 1 || waves_per_simd > 2) throw HipFailure("waves per SIMD must be 1 or 2");
        ctx->vol_waves_per_simd = waves_per_simd;
        return 0;
    });
}

int grrt_ext_set_precision(GRRTContext* ctx, int mode) {
    return guarded(ctx, [&] {
        if (mode < 0 || mode > 2) throw HipFailure("precision mode must be 0 (FP64), 1 or 2 (mixed)");
        ctx->mixed_precision = mode >= 1;
        ctx->far_field_f32 = mode == 2;
        return 0;
    });
}

int grrt_ext_set_integrator(GRRTContext* ctx, int mode) {
    return guarded(ctx, [&] {
        if (mode < 0 || mode > 1) throw HipFailure("integrator must be 0 (Dormand-Prince) or 1 (step-doubling RK4)");
        ctx->integrator = mode;
        ctx->batch_views.clear();      // the recorded batch carries its parameter block
        return 0;
    });
}

int grrt_ext_render_spectral_tile(GRRTContext* ctx, double* out, int x, int y, int tile_width, int tile_height) {
    return guarded(ctx, [&] {
        const auto& r = ctx->res;
        if (!out || tile_width <= 0 || tile_height <= 0 || x < 0 || y < 0 || x + tile_width > r.W ||
            y + tile_height > r.H)
            throw HipFailure("grrt_ext_render_spectral_tile: tile outside the frame");
        if (ctx->frequency_bins.empty()) throw HipFailure("No frequency bins set -- call grrt_set_frequency_bins first");
        const size_t n = static_cast<size_t>(tile_width) * tile_height * ctx->frequency_bins.size();
        if (!r.volumetric) {
            std::memset(out, 0, n * sizeof(double));
            return 0;
        }
        begin_call(ctx);
        check(ctx->spec_out.ensure(n * sizeof(double)), "hipMalloc(spectral out)");
        const int rc = render_window(ctx, x, y, tile_width, tile_height, ctx->spec_out.ptr, nullptr, nullptr,
                                     Output::kSpectralF64);
        check(hipMemcpy(out, ctx->spec_out.ptr, n * sizeof(double), hipMemcpyDeviceToHost), "copy spectra");
        return rc;
    });
}

int grrt_ext_render_spectral_f32(GRRTContext* ctx, float* out, int width, int height) {
    return guarded(ctx, [&] { return render_spectral_host(ctx, out, width, height, true, nullptr, nullptr); });
}

}  // extern "C"

struct GRRTVolTables {
    grrt_amd::VolTables t;
};

extern "C" {

GRRTVolTables* grrt_ext_vol_build(const GRRTParams* params) {
    if (!params) { g_last_error = "grrt_ext_vol_build: params is null"; return nullptr; }
    try {
        const VolSetup vs = resolve_volumetric(*params);
        auto* h = new GRRTVolTables();
        grrt_amd::build_volumetric(vs.M, vs.a, vs.r_outer, vs.t_peak, vs.vp, h->t);
        return h;
    } catch (const std::exception& e) {
        g_last_error = e.what();
    }
    return nullptr;
}

GRRTVolTables* grrt_ext_vol_build_ex(const GRRTParams* params, int flags) {
    if (!params) { g_last_error = "grrt_ext_vol_build_ex: params is null"; return nullptr; }
    if ((flags & ~3) != 0 || flags == 2) {
        g_last_error = "grrt_ext_vol_build_ex: flags must be 0, 1 or 3";
        return nullptr;
    }
    if (flags == 0) return grrt_ext_vol_build(params);
    hipStream_t st = nullptr;
    GRRTVolTables* h = nullptr;
    try {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) throw std::runtime_error("no GPU");
        if (hipStreamCreate(&st) != hipSuccess) throw std::runtime_error("hipStreamCreate failed");
        const VolSetup vs = resolve_volumetric(*params);
        h = new GRRTVolTables();
        grrt_amd::build_volumetric(vs.M, vs.a, vs.r_outer, vs.t_peak, vs.vp, h->t, &st, (flags & 2) != 0);
    } catch (const std::exception& e) {
        g_last_error = e.what();
        delete h;
        h = nullptr;
    }
    if (st) (void)hipStreamDestroy(st);
    return h;
}

GRRTVolTables* grrt_ext_vol_build_device(const GRRTParams* params) { return grrt_ext_vol_build_ex(params, 3); }

int grrt_ext_vol_scalars(const GRRTVolTables* h, double* out, int cap) {
    if (!h || !out) return -1;
    const auto& t = h->t;
    const double v[] = {t.r_isco, t.r_horizon, t.r_min, t.r_outer, t.taper_width, t.E_isco, t.L_isco,
                        t.rho_scale, t.noise_scale, t.prm.turbulence, static_cast<double>(t.prm.noise_octaves),
                        static_cast<double>(t.n_r), static_cast<double>(t.n_z), static_cast<double>(t.op.n_nu),
                        t.op.log_nu_min, t.op.log_nu_max, t.M, t.a, t.t_peak, t.build_ms_opacity,
                        t.build_ms_total, t.build_ms_columns, t.build_ms_columns_kernel};
    const int n = static_cast<int>(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < cap; ++i) out[i] = v[i];
    return n;
}

int64_t grrt_ext_vol_table(const GRRTVolTables* h, const char* name, double* out, int64_t cap) {
    if (!h || !name) return -1;
    const auto& t = h->t;
    const std::string n(name);
    std::vector<double> perm;
    const std::vector<double>* v = nullptr;
    if (n == "H") v = &t.H;
    else if (n == "z_max") v = &t.z_max;
    else if (n == "rho_mid") v = &t.rho_mid;
    else if (n == "T_eff") v = &t.T_eff;
    else if (n == "rho_profile") v = &t.rho_profile;
    else if (n == "T_profile") v = &t.T_profile;
    else if (n == "kappa_abs") v = &t.op.kappa_abs;
    else if (n == "kappa_es") v = &t.op.kappa_es;
    else if (n == "kappa_ross") v = &t.op.kappa_ross;
    else if (n == "mu") v = &t.op.mu;
    else if (n == "perm") {
        perm.assign(t.perm, t.perm + 512);
        v = &perm;
    } else {
        return -1;
    }
    const int64_t len = static_cast<int64_t>(v->size());
    if (out && cap >= len) std::memcpy(out, v->data(), v->size() * sizeof(double));
    return len;
}

void grrt_ext_vol_free(GRRTVolTables* h) { delete h; }

int grrt_ext_resolve_params(const GRRTParams* params, double* out, int cap) {
    if (!params || !out) return -1;
    try {
        const grrt_amd::Resolved r = grrt_amd::resolve_params(*params);
        const VolSetup vs = resolve_volumetric(*params);
        const double v[] = {static_cast<double>(r.W), static_cast<double>(r.H), r.M, r.a, r.r_obs, r.th_obs,
                            r.ph_obs, r.fov, static_cast<double>(r.max_steps), r.tol, static_cast<double>(r.disk),
                            static_cast<double>(r.volumetric), r.r_in, r.r_out, r.t_disk,
                            static_cast<double>(r.stars), static_cast<double>(r.spp), r.r_h, r.r_isco, vs.M, vs.a,
                            vs.r_outer, vs.t_peak, vs.vp.opacity_nu_min, vs.vp.opacity_nu_max};
        const int n = static_cast<int>(sizeof(v) / sizeof(v[0]));
        for (int i = 0; i < n && i < cap; ++i) out[i] = v[i];
        return n;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}

int grrt_ext_saha(double rho, double T, double out[5]) {
    if (!out) return -1;
    const grrt_amd::Ionization s = grrt_amd::saha(rho, T);
    out[0] = s.n_e;
    out[1] = s.mu;
    out[2] = s.n_ion_eff;
    out[3] = s.n_HI;
    out[4] = s.n_Hminus;
    return 0;
}

int grrt_ext_opacity(double nu, double rho, double T, double out[6]) {
    if (!out) return -1;
    const grrt_amd::Ionization s = grrt_amd::saha(rho, T);
    out[0] = grrt_amd::kappa_abs_point(nu, rho, T, s);
    out[1] = grrt_amd::kappa_es_point(rho, s);
    out[2] = grrt_amd::alpha_ff_point(nu, T, s);
    out[3] = grrt_amd::alpha_hminus_point(nu, T, s);
    out[4] = grrt_amd::alpha_bf_point(nu, T, s);
    out[5] = grrt_amd::planck_nu_host(nu, T);
    return 0;
}

}  // extern "C"

namespace grrt_api {

// (Re)build the batch's per-view tables and parameter blocks, and record
// the launch sequence as a graph.
void prepare_batch(GRRTContext* ctx, const GRRTView* views, int n, uint64_t seed_stride) {
    using namespace grrt_amd;
    const Resolved& r = ctx->res;
    if (r.volumetric) throw HipFailure("batched rendering covers the thin disk only");
    // Every view is checked before anything is freed or re-recorded, so a
    // rejected batch leaves the previous recording (graph, views, tables
    // and buffers) intact and replayable.
    for (int v = 0; v < n; ++v) {
        const GRRTView& w = views[v];
        if (w.width <= 0 || w.height <= 0) throw HipFailure("batch view with an empty frame");
        if (!(std::abs(w.spin) < 1.0)) throw HipFailure("batch view spin must satisfy |a/M| < 1");
    }
    // The recorded graph holds pointers into the buffers grown below: drop
    // it (and the views it was recorded for) before any of them moves.
    if (ctx->batch_exec) (void)hipGraphExecDestroy(ctx->batch_exec);
    if (ctx->batch_graph) (void)hipGraphDestroy(ctx->batch_graph);
    ctx->batch_exec = nullptr;
    ctx->batch_graph = nullptr;
    ctx->batch_views.clear();
    std::vector<ViewParams> vp(n);
    std::vector<double> flux(static_cast<size_t>(n) * kFluxN, 0.0);
    check(ctx->batch_flux.ensure(flux.size()), "hipMalloc(batch flux)");
    long long q = 0, rays = 0, pix = 0;
    for (int v = 0; v < n; ++v) {
        const GRRTView& w = views[v];
        const double a = w.spin * r.M;
        const double r_obs = w.observer_r > 0.0 ? w.observer_r : r.r_obs;
        const double th = w.observer_theta > 0.0 ? w.observer_theta : r.th_obs;
        const double fov = w.fov > 0.0 ? w.fov : r.fov;
        const double r_in = ctx->params.disk_inner > 0.0 ? ctx->params.disk_inner : kerr_isco(r.M, a);
        double fmax = 0.0;
        std::vector<double> f;
        if (r.disk) {
            fmax = build_flux_table(r.M, a, r_in, r.r_out, f);
            std::copy(f.begin(), f.end(), flux.begin() + static_cast<size_t>(v) * kFluxN);
        }
        const CameraFrame cam = build_camera(r.M, a, r_obs, th, w.observer_phi);
        fill_view(ctx, vp[v], a, r_obs, th, w.observer_phi, fov, w.width, w.height, cam,
                  ctx->batch_flux.ptr + static_cast<size_t>(v) * kFluxN, fmax, r_in, 0, 0, w.width, w.height,
                  ctx->tune_tile);
        vp[v].seed = ctx->sampling_seed + static_cast<uint64_t>(v) * seed_stride;
        vp[v].q_offset = q;
        vp[v].ray_offset = rays;
        vp[v].pix_offset = pix;
        q += view_queue_length(vp[v]);
        rays += vp[v].n_rays;
        pix += static_cast<long long>(w.width) * w.height;
    }
    upload(ctx->batch_flux, flux, ctx->stream);
    upload(ctx->batch_params, vp, ctx->stream);
    const int nS = vp[0].nS;
    check(ctx->batch_rgb.ensure(static_cast<size_t>(3 * rays)), "hipMalloc(batch samples)");
    check(ctx->batch_cls.ensure(static_cast<size_t>(rays)), "hipMalloc(batch classes)");
    check(ctx->batch_out.ensure(static_cast<size_t>(4 * pix)), "hipMalloc(batch frames)");
    check(ctx->batch_counters.ensure(1), "hipMalloc(batch counters)");
    check(ctx->batch_tp.ensure(1), "hipMalloc(batch params)");
    TraceParams P{};
    P.view = vp[0];
    P.views = ctx->batch_params.ptr;
    P.n_views = n;
    P.r_escape = 1000.0;
    P.tol = r.tol;
    P.max_steps = r.max_steps;
    P.n_rays = q;
    P.spec_lut = ctx->spec.ptr;
    P.stars = r.stars;
    P.star_start = ctx->star_start.ptr;
    P.star_tpb = ctx->star_tpb.ptr;
    set_background(ctx, P);
    P.sample_rgb = ctx->batch_rgb.ptr;
    P.sample_cls = ctx->batch_cls.ptr;
    if (ctx->keep_records) check(ctx->batch_records.ensure(static_cast<size_t>(rays)), "hipMalloc(batch records)");
    P.records = ctx->keep_records ? ctx->batch_records.ptr : nullptr;
    P.counters = ctx->batch_counters.ptr;
    P.cancel = nullptr;
    P.refill_min = std::clamp(ctx->tune_refill, 1, 64);
    P.chunk = std::max(ctx->tune_chunk, 64);
    P.integrator = ctx->integrator;
    check(hipMemcpy(ctx->batch_tp.ptr, &P, sizeof(TraceParams), hipMemcpyHostToDevice), "hipMemcpy(batch params)");
    // Record: reset, the persistent trace over all views, finalize.
    check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    hipError_t e = launch_reset(P.counters, P.sample_cls, rays, ctx->stream);
    if (e == hipSuccess)
        e = launch_trace_thin(ctx->batch_tp.ptr, 1, std::max(1, ctx->num_cus * ctx->tune_blocks_per_cu),
                              ctx->tune_block, ctx->stream, ctx->integrator);
    if (e == hipSuccess) e = launch_finalize(P.sample_rgb, P.sample_cls, ctx->batch_out.ptr, pix, nS, ctx->stream);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess && g) (void)hipGraphDestroy(g);
    check(e, "record batch launch");
    check(e2, "end capture");
    ctx->batch_graph = g;
    check(hipGraphInstantiate(&ctx->batch_exec, g, nullptr, nullptr, 0), "instantiate batch graph");
    ctx->batch_views.assign(views, views + n);
    ctx->batch_rays = rays;
    ctx->batch_pixels = pix;
    ctx->batch_tuning[0] = ctx->tune_tile;
    ctx->batch_tuning[1] = ctx->tune_blocks_per_cu;
    ctx->batch_tuning[2] = ctx->tune_refill;
    ctx->batch_tuning[3] = ctx->tune_chunk;
}

bool same_batch(const GRRTContext* ctx, const GRRTView* views, int n) {
    if (static_cast<int>(ctx->batch_views.size()) != n || !ctx->batch_exec) return false;
    if (ctx->batch_tuning[0] != ctx->tune_tile || ctx->batch_tuning[1] != ctx->tune_blocks_per_cu ||
        ctx->batch_tuning[2] != ctx->tune_refill || ctx->batch_tuning[3] != ctx->tune_chunk)
        return false;
    return std::memcmp(ctx->batch_views.data(), views, sizeof(GRRTView) * n) == 0;
}

// Replay the recorded batch; fills ctx->last.
int run_batch(GRRTContext* ctx) {
    using namespace grrt_amd;
    const auto t0 = std::chrono::steady_clock::now();
    check(hipEventRecord(ctx->ev[0], ctx->stream), "hipEventRecord");
    check(hipGraphLaunch(ctx->batch_exec, ctx->stream), "graph launch");
    check(hipEventRecord(ctx->ev[1], ctx->stream), "hipEventRecord");
    check(hipEventSynchronize(ctx->ev[1]), "batch render");
    DevCounters c{};
    check(hipMemcpy(&c, ctx->batch_counters.ptr, sizeof(c), hipMemcpyDeviceToHost), "read counters");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "event time");
    GRRTStats& s = ctx->last;
    s = GRRTStats{};
    s.trace_ms = ms;
    s.rays = static_cast<int64_t>(c.rays_done);
    s.steps = static_cast<int64_t>(c.steps);
    s.attempts = static_cast<int64_t>(c.attempts);
    s.deriv_evals = static_cast<int64_t>(c.rays_done + 6 * c.attempts);
    s.algo_flops = flops_per_attempt() * static_cast<double>(c.attempts);
    s.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->last_rays = ctx->batch_rays;
    ctx->records_valid = ctx->keep_records && ctx->batch_records.ptr != nullptr;
    ctx->last_batch = true;
    return 0;
}

}  // namespace grrt_api

extern "C" {

int grrt_ext_render_batch(GRRTContext* ctx, const GRRTView* views, int n, float* out, int device_out) {
    return guarded(ctx, [&] {
        if (!views || n <= 0 || (!out && !device_out))
            throw HipFailure("grrt_ext_render_batch: empty batch or null output");
        if (!same_batch(ctx, views, n)) prepare_batch(ctx, views, n);
        run_batch(ctx);
        const size_t bytes = static_cast<size_t>(4 * ctx->batch_pixels) * sizeof(float);
        if (out)
            check(hipMemcpy(out, ctx->batch_out.ptr, bytes,
                            device_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost),
                  "copy batch frames");
        return 0;
    });
}

int grrt_ext_autotune(GRRTContext* ctx, const GRRTView* views, int n, int best[4], double* best_ms) {
    return guarded(ctx, [&] {
        const int tiles[] = {0, 8};
        const int bpcs[] = {1, 2, 3};
        const int refills[] = {2, 4, 8, 16};
        const int chunks[] = {64, 128};
        double best_t = 1e300;
        int bt[4] = {ctx->tune_tile, ctx->tune_blocks_per_cu, ctx->tune_refill, ctx->tune_chunk};
        const auto& r = ctx->res;
        if (r.volumetric) throw HipFailure("autotune covers the thin-disk tracer");
        if (!views) check(ctx->frame.ensure(4 * static_cast<size_t>(r.W) * r.H), "hipMalloc(frame)");
        begin_call(ctx);
        for (int tile : tiles)
            for (int bpc : bpcs)
                for (int rf : refills)
                    for (int ch : chunks) {
                        ctx->tune_tile = tile;
                        ctx->tune_blocks_per_cu = bpc;
                        ctx->tune_refill = rf;
                        ctx->tune_chunk = ch;
                        double t = 0.0;
                        if (views) {
                            prepare_batch(ctx, views, n);
                            run_batch(ctx);      // warm
                            run_batch(ctx);
                            t = ctx->last.trace_ms;
                        } else {
                            render_window(ctx, 0, 0, r.W, r.H, ctx->frame.ptr, nullptr, nullptr);
                            t = ctx->last.trace_ms;
                        }
                        if (t < best_t) {
                            best_t = t;
                            bt[0] = tile; bt[1] = bpc; bt[2] = rf; bt[3] = ch;
                        }
                    }
        ctx->tune_tile = bt[0];
        ctx->tune_blocks_per_cu = bt[1];
        ctx->tune_refill = bt[2];
        ctx->tune_chunk = bt[3];
        if (views) prepare_batch(ctx, views, n);
        if (best) for (int i = 0; i < 4; ++i) best[i] = bt[i];
        if (best_ms) *best_ms = best_t;
        return 0;
    });
}

}  // extern "C"

extern "C" {

int grrt_ext_render_spectral_to_fits_f32(GRRTContext* ctx, const char* path, int width, int height,
                                         grrt_progress_fn progress, void* user_data) {
    return guarded(ctx, [&] { return spectral_to_fits(ctx, path, width, height, true, progress, user_data); });
}

int grrt_ext_write_fits_cube(const char* path, const void* cube, int bitpix, int width, int height, int num_bins,
                             const double* frequencies_hz, const GRRTParams* meta) {
    try {
        if (!path || !cube || (bitpix != -64 && bitpix != -32)) throw std::runtime_error("invalid FITS cube arguments");
        std::vector<double> f;
        if (frequencies_hz) f.assign(frequencies_hz, frequencies_hz + std::max(num_bins, 0));
        GRRTParams zero{};
        grrt_amd::FitsCubeWriter out(path, width, height, num_bins, f, fits_meta(meta ? *meta : zero), bitpix);
        const size_t row = static_cast<size_t>(width) * num_bins;
        for (int j = 0; j < height; ++j) {
            if (bitpix == -64) out.write_row(j, static_cast<const double*>(cube) + row * j);
            else out.write_row(j, static_cast<const float*>(cube) + row * j);
        }
        out.finish();
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}

int grrt_ext_render_hdr(GRRTContext* ctx, const char* path_linear, const char* path_scaled, float* framebuffer,
                        grrt_progress_fn progress, void* user_data) {
    return guarded(ctx, [&] { return render_hdr(ctx, path_linear, path_scaled, framebuffer, progress, user_data); });
}

int grrt_ext_read_texture(const char* path, int* width, int* height, float* rgb, int64_t capacity) {
    try {
        if (!path || !width || !height) throw std::runtime_error("null argument");
        const grrt_amd::Texture t = grrt_amd::read_texture(path);
        *width = t.width;
        *height = t.height;
        if (rgb) {
            if (capacity < static_cast<int64_t>(t.rgb.size())) throw std::runtime_error("texture buffer too small");
            std::memcpy(rgb, t.rgb.data(), t.rgb.size() * sizeof(float));
        }
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}

int grrt_ext_write_hdr(const char* path, const float* rgba, int width, int height, int normalised) {
    try {
        if (!path) throw std::runtime_error("null path");
        const float scale = normalised ? grrt_amd::hdr_exposure_scale(rgba, width, height) : 1.0f;
        grrt_amd::write_rgbe(path, rgba, width, height, scale);
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}

}  // extern "C"
