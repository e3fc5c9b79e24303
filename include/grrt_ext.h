/*
 * Additive extension entry points of libgrrt.so.  None of these exists in
 * the reference ABI; the 16 reference symbols in grrt/api.h are unchanged.
 * They expose what parity checks and measurement need: the sample
 * placement, per-ray records and outcome classes, device-resident output
 * and per-render counters.
 */
#ifndef GRRT_EXT_H
#define GRRT_EXT_H

#include <stdint.h>
#include "grrt/api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sample placement.
 *  HASH   — the reference renderer's stratified jitter (src/renderer.cpp:14-20,
 *           :48-55): grid = floor(sqrt(spp)), grid^2 samples.  The default.
 *  CENTER — one ray through each pixel centre (src/camera.cpp:65-67).
 *  SEEDED — spp samples at (i + jx, j + jy) with (jx, jy) uniform in [0,1)^2
 *           from a counter-based generator keyed by (seed, sample index):
 *           SplitMix64 finaliser over a Weyl sequence, jx from draw 2k+1,
 *           jy from draw 2k+2, 53-bit mantissas. */
enum { GRRT_SAMPLING_HASH = 0, GRRT_SAMPLING_CENTER = 1, GRRT_SAMPLING_SEEDED = 2 };

/* Outcome class of a ray: a disk hit is any equatorial crossing with
 * r_in <= r <= r_out; otherwise the termination decides. */
enum { GRRT_CLASS_CAPTURED = 0, GRRT_CLASS_ESCAPED = 1, GRRT_CLASS_DISK = 2,
       GRRT_CLASS_MAXSTEPS = 3 };

/* Per-ray record (144 bytes).  x_final[0] (coordinate time) is not
 * integrated and reads 0.  h_max / h_final are |H| = |g^{ab}p_a p_b|/2 over
 * accepted states and at the end. */
typedef struct {
    double color[3];
    double x_final[4];
    double p_final[4];
    double r_hit, g_hit;   /* first in-range disk crossing, 0 if none */
    double h_max, h_final;
    int32_t term;          /* 0 horizon, 1 escaped, 2 max steps */
    int32_t cls;
    int32_t steps, attempts, n_cross, pad;
} GRRTRayRecord;

typedef struct {
    double trace_ms;       /* trace kernel(s), HIP events on the render stream */
    double finalize_ms;    /* per-pixel sample reduction kernel */
    double total_ms;       /* whole call, host wall clock, incl. copies */
    int64_t rays;
    int64_t steps;         /* accepted Dormand–Prince steps */
    int64_t attempts;      /* Dormand–Prince attempts (accepted + rejected) */
    int64_t deriv_evals;   /* Kerr right-hand-side evaluations */
    double algo_flops;     /* algorithmic FP64 flops of the trace kernel */
    int32_t cancelled;
    int32_t pad;
    int64_t march_steps;   /* volumetric ray-march RK4 steps */
    double copy_ms;        /* device-to-host copy of the frame (grrt_render*), host wall clock */
} GRRTStats;

GRRT_EXPORT int grrt_ext_set_sampling(GRRTContext* ctx, int mode, int spp, uint64_t seed);
/* Keep per-ray records of subsequent renders (enable=1) for grrt_ext_get_records. */
GRRT_EXPORT int grrt_ext_set_records(GRRTContext* ctx, int enable);
/* Samples of the context's frame (width * height * samples per pixel), or
 * of all views when the last render was a batch (grrt_ext_render_batch). */
GRRT_EXPORT int64_t grrt_ext_num_samples(const GRRTContext* ctx);
GRRT_EXPORT int grrt_ext_get_records(GRRTContext* ctx, GRRTRayRecord* out, int64_t n);
GRRT_EXPORT int grrt_ext_get_classes(GRRTContext* ctx, uint8_t* out, int64_t n);
/* Per-sample linear RGB of the last frame (n <= grrt_ext_num_samples, three
 * doubles per sample, pixel-major sample order). */
GRRT_EXPORT int grrt_ext_get_samples(GRRTContext* ctx, double* rgb, int64_t n);
GRRT_EXPORT int grrt_ext_last_stats(const GRRTContext* ctx, GRRTStats* out);
/* Render the full frame into device memory (d_rgba: width*height*4 floats,
 * or NULL to keep it in the context).  No host copy is made. */
GRRT_EXPORT int grrt_ext_render_device(GRRTContext* ctx, float* d_rgba);
/* The same frame as grrt_ext_render_device(ctx, d_rgba) without any device
 * allocation: no reference counterpart; for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  d_rgba (width*height*4 floats of device memory) is required.
 * The launch block, counters and per-sample buffers live in a fixed 128 MiB
 * device array of the library, one per GPU, held from grrt_create on that GPU
 * until the process ends; frames above its size (25 bytes per sample, about
 * 5.3 M samples) are rejected with -1 and an error, and per-ray records are
 * not kept.  Every call on one GPU uses that GPU's array: calls that overlap
 * on one GPU (two threads, two streams) race.  Volumetric RGB frames are
 * covered too; spectral renders are not. */
GRRT_EXPORT int grrt_ext_render_device_noalloc(GRRTContext* ctx, float* d_rgba);
/* Tone-maps a frame in device memory in place with the grrt_tonemap curve
 * (log-average exposure, extended Reinhard with the frame's white point,
 * gamma 2.2; src/tonemapper.cpp:7-71, whose grrt_tonemap works on a host
 * buffer): d_rgba is width*height*4 floats, or NULL for the context's own
 * frame (grrt_ext_render_device(ctx, NULL), or a surrogate or batch-free
 * frame kept on the device).  gpu_ms (may be NULL) receives the kernels'
 * GPU time. */
GRRT_EXPORT int grrt_ext_tonemap_device(GRRTContext* ctx, float* d_rgba, double* gpu_ms);
/* Copies the context's own device frame (width*height*4 floats) to host
 * memory. */
GRRT_EXPORT int grrt_ext_get_frame(GRRTContext* ctx, float* out);
/* Persistent-kernel launch shape: threads per block (multiple of 64),
 * blocks per CU, the number of idle lanes (1..64) that triggers a
 * wavefront refill, and the rays a wavefront claims from the device queue
 * per atomic (>= 64); 0 keeps the current value.  tile: queue order of the
 * RGB frame, 0 = row-major samples, 2..32 = tile x tile pixel blocks;
 * negative keeps the current order.  Defaults: 256 threads, 3 blocks per
 * CU, refill at 8 idle lanes, 64-ray claims, row-major. */
GRRT_EXPORT int grrt_ext_set_tuning(GRRTContext* ctx, int block, int blocks_per_cu, int refill_min,
                                    int chunk, int tile);

/* One camera view of a batch: its own Kerr spin (dimensionless, a/M, used
 * as given: 0 is Schwarzschild, negative is a retrograde hole, unlike
 * GRRTParams.spin where <= 0 selects 0.998), observer position, field of
 * view and resolution (non-positive r, theta, fov take the context's
 * values).  Everything else (mass,
 * thin disk settings with r_in = ISCO of this spin unless disk_inner is set,
 * stars, tolerance, step limit, sample placement) comes from the context. */
typedef struct {
    double spin;
    double observer_r, observer_theta, observer_phi, fov;
    int width, height;
} GRRTView;

/* Renders n views of uneven size in ONE persistent kernel launch over the
 * concatenated ray queue (one lane per ray, per-lane view constants).  The
 * launch sequence (parameter upload, reset, trace, finalize) is recorded as
 * a HIP graph and replayed while the views stay the same.  out: the RGBA
 * float32 frames in view order, sum(width*height*4) floats; host memory,
 * or device memory when device_out != 0 (NULL with device_out != 0 keeps
 * the frames in the context's device memory). */
GRRT_EXPORT int grrt_ext_render_batch(GRRTContext* ctx, const GRRTView* views, int n, float* out,
                                      int device_out);

/* Picks the persistent kernel's launch shape by timing candidates on the
 * real workload: the context's frame (views == NULL) or the given batch.
 * Candidates: pixel tile (0 = row order, 8), blocks per CU (1, 2, 3), refill
 * threshold (2, 4, 8, 16 idle lanes), queue chunk (64, 128).  The best is
 * kept (for the batch: recorded with its graph) and written to
 * best[4] = {tile, blocks_per_cu, refill_min, chunk}; best_ms its time. */
GRRT_EXPORT int grrt_ext_autotune(GRRTContext* ctx, const GRRTView* views, int n, int best[4],
                                  double* best_ms);

/* Volumetric path precision: 0 = all FP64 (default), 1 = FP32 radiative
 * transfer (opacities, source, extinction) with compensated summation of
 * each bin's intensity, the geodesic in FP64; 2 = as 1, and the geodesic's
 * outer adaptive steps in FP32 (tolerance 1e-6) while r > 15 M.  Mode 2
 * moves each ray's disk entry by ~1e-5 M, which shifts the coarse
 * ray-march's step sequence: per-pixel intensities then differ from the
 * all-FP64 result by the ray-march's own discretisation error. */
GRRT_EXPORT int grrt_ext_set_precision(GRRTContext* ctx, int mode);
/* Adaptive integrator of the geodesic: 0 (default) the Dormand-Prince 5(4)
 * controller the reference tracer calls (src/rk4.cpp:87-136); 1 the
 * reference's step-doubling RK4 controller (src/rk4.cpp:33-80: a step of dl
 * against two of dl/2, accept and double below 1% of tol, else halve).
 * Applies to the thin-disk, batch, surrogate-fallback and volumetric
 * tracers. */
GRRT_EXPORT int grrt_ext_set_integrator(GRRTContext* ctx, int mode);
/* grrt_render_spectral into a float32 cube (width*height*nbins). */
/* The spectra of the pixel window [x, x+w) x [y, y+h) of the context's
 * frame (w*h*bins doubles, the same values as those pixels of
 * grrt_render_spectral): the spectral counterpart of grrt_render_tile. */
GRRT_EXPORT int grrt_ext_render_spectral_tile(GRRTContext* ctx, double* out, int x, int y, int tile_width,
                                              int tile_height);
/* Launch shape of the volumetric tracer: 1 or 2 persistent wavefronts per
 * SIMD (the kernel's register budget is cut to match). */
GRRT_EXPORT int grrt_ext_set_vol_tuning(GRRTContext* ctx, int waves_per_simd);
GRRT_EXPORT int grrt_ext_render_spectral_f32(GRRTContext* ctx, float* out, int width, int height);

/* Access to the volumetric disk set-up (src/volumetric_disk.cpp:33-73,
 * src/opacity.cpp:237-303): builds the opacity, radial and vertical tables
 * for the volumetric disk that grrt_create would build from `params`, on
 * host threads (no GPU needed). */
typedef struct GRRTVolTables GRRTVolTables;
GRRT_EXPORT GRRTVolTables* grrt_ext_vol_build(const GRRTParams* params);
/* The same with the opacity tables (10^4 independent Saha solves and their
 * absorption and Rosseland opacities) and the 500 vertical columns computed
 * on the GPU, as grrt_create does; NULL without a GPU. */
GRRT_EXPORT GRRTVolTables* grrt_ext_vol_build_device(const GRRTParams* params);
/* Either part on the GPU: flags bit 0, the opacity tables; bit 1, the
 * vertical columns (which needs bit 0).  0: grrt_ext_vol_build; 3:
 * grrt_ext_vol_build_device; 1: opacity tables on the GPU, columns on host
 * threads, which equal the GPU's columns bit for bit. */
GRRT_EXPORT GRRTVolTables* grrt_ext_vol_build_ex(const GRRTParams* params, int flags);
/* Scalars in this order: r_isco, r_horizon, r_min, r_outer, taper_width,
 * E_isco, L_isco, rho_scale, noise_scale, turbulence, octaves, n_r, n_z,
 * n_nu, log10 nu_min, log10 nu_max, M, a, t_peak, then the set-up's wall
 * time in ms: opacity tables, whole build, vertical columns, and the GPU
 * column kernel alone (0 for a host build).  Returns the count. */
GRRT_EXPORT int grrt_ext_vol_scalars(const GRRTVolTables* t, double* out, int cap);
/* Copies table `name` ("H", "z_max", "rho_mid", "T_eff", "rho_profile",
 * "T_profile", "kappa_abs", "kappa_es", "kappa_ross", "mu", "perm") into
 * `out` if it fits; returns the table length, or -1 for an unknown name. */
GRRT_EXPORT int64_t grrt_ext_vol_table(const GRRTVolTables* t, const char* name, double* out, int64_t cap);
GRRT_EXPORT void grrt_ext_vol_free(GRRTVolTables* t);
/* The scene grrt_create builds from `params` under the reference's
 * zero-means-default rules (src/api.cpp:43-126), without a GPU.  Values in
 * this order: width, height, M, a (dimensional), r_obs, theta_obs, phi_obs,
 * fov, max_steps, tol, thin disk on, volumetric disk on, r_in, r_out,
 * T_disk, stars, spp, r_horizon, r_isco, then the volumetric disk's own
 * wiring: its M and spin (the raw params fields), r_outer, T_peak, opacity
 * nu_min, nu_max.  Returns the count (25), or -1. */
GRRT_EXPORT int grrt_ext_resolve_params(const GRRTParams* params, double* out, int cap);
/* Saha ionisation of the H/He/C/O/Fe mixture: n_e, mu, n_ion_eff, n_HI,
 * n_Hminus (CGS). */
GRRT_EXPORT int grrt_ext_saha(double rho_cgs, double T, double out[5]);
/* Point opacities at (nu, rho, T): kappa_abs, kappa_es [cm^2/g],
 * alpha_ff, alpha_hminus, alpha_bf [1/cm], B_nu. */
GRRT_EXPORT int grrt_ext_opacity(double nu, double rho_cgs, double T, double out[6]);

/* ---- Learned surrogate of the ray transfer map ----------------------------
 * A 5 -> 256 -> 256 -> 256 -> 256 -> 9 MLP (SiLU, BF16 weights, FP32
 * accumulation on MFMA) from (pixel direction alpha, beta; spin a;
 * cos theta_obs; log r_obs) to (3 class logits captured/escaped/disk, ln r
 * and ln g of the first disk crossing, sin/cos of the escape theta and phi),
 * trained on rays of the context's own FP64 integrator.  The disk extent it
 * learns is the context's (r_in = ISCO of each spin, r_out = disk_outer). */
typedef struct {
    int steps;                /* Adam steps in this call */
    int batch;                /* rays per step (a multiple of 256) */
    double lr_max, lr_min;    /* linear warm-up to lr_max, cosine decay to lr_min */
    int warmup;               /* warm-up steps */
    int pad;
    double w_r, w_g, w_dir;   /* loss weights of ln r, ln g, escape direction */
    uint64_t seed;            /* batch shuffling */
} GRRTTrainConfig;

typedef struct {
    int64_t steps;            /* Adam steps taken so far (all calls) */
    int64_t rays;             /* rows evaluated */
    double loss, ce;          /* mean total loss and cross-entropy */
    double accuracy;          /* class accuracy (argmax) */
    double rmse_lnr, rmse_lng, rmse_dir;  /* on disk rays / escaped rays */
    double ms;                /* GPU time of the call */
} GRRTTrainStats;

typedef struct {
    double infer_ms;          /* GPU time from the frame's start until the network pass ends */
    double trace_ms;          /* FP64 fallback/judge tracing still running after that */
    double gpu_ms;            /* GPU time of the whole frame (infer_ms + trace_ms) */
    double total_ms;          /* wall time of the call */
    int64_t rays, fallback_rays, judged_rays, trace_attempts;
    double judge_class_agreement;     /* judged rays: surrogate class == FP64 class */
    double judge_median_rel_err;      /* judged lit rays: |L_s - L| / L of summed RGB */
} GRRTSurrogateStats;

/* Creates (or re-initialises) the network with seeded random weights. */
GRRT_EXPORT int grrt_ext_surrogate_init(GRRTContext* ctx, uint64_t seed);
/* Traces n_rays training rays (validation=0; n_rays a power of two) or
 * validation rays (validation=1) with the FP64 integrator: views of 32x32
 * pixels, 1 uniform jittered sample each, a ~ U[0, 0.998], theta_obs ~
 * U[5, 89] deg, r_obs log-uniform in [20, 100], 90 deg field. */
GRRT_EXPORT int grrt_ext_surrogate_generate(GRRTContext* ctx, int64_t n_rays, uint64_t seed, int validation);
/* Adam steps over the training set; stats are the running means of the
 * call's last 10% of steps. */
GRRT_EXPORT int grrt_ext_surrogate_train(GRRTContext* ctx, const GRRTTrainConfig* cfg, GRRTTrainStats* out);
/* Metrics on the validation set. */
GRRT_EXPORT int grrt_ext_surrogate_evaluate(GRRTContext* ctx, GRRTTrainStats* out);
/* FP32 master parameters (201225 floats): get (set=0) or set (set=1). */
GRRT_EXPORT int grrt_ext_surrogate_params(GRRTContext* ctx, float* params, int64_t n, int set);
/* Network outputs y[n][9] for inputs x[n][5] (host memory). */
GRRT_EXPORT int grrt_ext_surrogate_forward(GRRTContext* ctx, const float* x, int64_t n, float* y);
/* Loss gradient (201225 floats) and loss at the current parameters for rows
 * x[n][5] with targets t[n][8] = (class or -1, ln r, ln g, sin th, cos th,
 * sin ph, cos ph, escaped); n a multiple of 256; weights w[3] = (w_r, w_g,
 * w_dir). */
GRRT_EXPORT int grrt_ext_surrogate_gradients(GRRTContext* ctx, const float* x, const float* t, int64_t n,
                                             const float w[3], float* grad, double* loss);
/* Copies rows [0, n) of the training (validation=0) or validation set. */
GRRT_EXPORT int grrt_ext_surrogate_dataset(GRRTContext* ctx, int validation, float* x, float* t, int64_t n);
/* Surrogate frame of the context's view: every sample is shaded from the
 * network unless its class entropy exceeds entropy_max (nats), its ray lies
 * near the critical curve (grrt_ext_surrogate_set_critical_gate) or, for a
 * predicted disk hit, the network's ln g differs by more than lng_tol from
 * the redshift that the predicted radius and the ray's conserved p_t, p_phi
 * fix (lng_tol <= 0: no such check); those samples are traced by the FP64
 * integrator.  A disk hit is shaded with the exact redshift of its predicted
 * radius.  judge_fraction of the shaded samples is traced as well and
 * compared (stats).  out: RGBA float32 as grrt_render (host memory, or
 * device memory when device_out != 0; NULL with device_out != 0 keeps the
 * frame in the context's device memory). */
/* Critical-curve gate of grrt_ext_render_surrogate: a sample whose camera
 * ray lies near the critical curve (the edge of the black hole's shadow) is
 * traced by the FP64 integrator whatever the network predicts.  The test is
 * |s| < score_max with s the minimum of R(r) / (E^2 r^4) over the spherical
 * photon-orbit radii, R the Kerr radial potential of the ray's conserved
 * E, L and Carter constant (0 on the critical curve).  Default 0.03;
 * 0 turns the gate off. */
GRRT_EXPORT int grrt_ext_surrogate_set_critical_gate(GRRTContext* ctx, double score_max);
GRRT_EXPORT int grrt_ext_render_surrogate(GRRTContext* ctx, float* out, int device_out, double entropy_max,
                                          double lng_tol, double judge_fraction, GRRTSurrogateStats* stats);
/* The surrogate frame's network pass itself, writing the network's 9 raw
 * outputs per ray to y[n][9] instead of shading: with x the inputs are
 * x[n][5]; with x NULL they are computed from the frame's sample positions
 * as the surrogate frame does (n must then be the frame's sample count). */
GRRT_EXPORT int grrt_ext_surrogate_infer_raw(GRRTContext* ctx, const float* x, int64_t n, float* y);
/* One Adam step of the training loop's kernel on caller arrays of n floats,
 * in place on w, m, v (bias corrections 1 - b^t as the training loop forms
 * them). */
GRRT_EXPORT int grrt_ext_surrogate_adam_step(GRRTContext* ctx, float* w, const float* g, float* m, float* v,
                                             int64_t n, double lr, double b1, double b2, double eps, int64_t t);
/* Sample ids of the last surrogate frame: the first n_fallback rays sent to
 * the FP64 integrator and the first n_judged judged rays (counts at most
 * GRRTSurrogateStats fallback_rays / judged_rays; either pointer may be
 * NULL). */
GRRT_EXPORT int grrt_ext_surrogate_lists(GRRTContext* ctx, int64_t* fallback, int64_t n_fallback, int64_t* judged,
                                         int64_t n_judged);

/* ---- Output files --------------------------------------------------------
 * grrt_render_spectral_to_fits_cb (include/grrt/api.h) streams the double
 * cube (BITPIX -64) in bands of rows; this variant streams the float32 cube
 * of the context's precision setting (BITPIX -32).  Same header cards,
 * row order and return codes. */
GRRT_EXPORT int grrt_ext_render_spectral_to_fits_f32(GRRTContext* ctx, const char* path, int width, int height,
                                                     grrt_progress_fn progress, void* user_data);
/* Host-only: writes a cube in renderer layout [(j*width + i)*num_bins + k]
 * (double for bitpix -64, float for -32) as FITS, metadata cards from meta
 * (NULL: defaults).  Returns 0, or -1 with grrt_last_error(). */
GRRT_EXPORT int grrt_ext_write_fits_cube(const char* path, const void* cube, int bitpix, int width, int height,
                                         int num_bins, const double* frequencies_hz, const GRRTParams* meta);
/* Host-only: writes an RGBA float frame as a Radiance RGBE .hdr (top row
 * first), linear (normalised=0) or scaled by 0.18 over the log-average
 * luminance of the lit pixels (normalised=1), the CLI's two HDR outputs. */
GRRT_EXPORT int grrt_ext_write_hdr(const char* path, const float* rgba, int width, int height, int normalised);
/* Host-only: reads the GRRT_BG_TEXTURE image format (an equirectangular
 * Radiance RGBE .hdr or PFM, row 0 at theta = 0, column 0 at phi = -pi) into
 * linear RGB floats, the form the tracer samples.  Writes the size; with rgb
 * non-null also the width*height*3 values (capacity in floats).  The
 * reference declares GRRT_BG_TEXTURE (include/grrt/types.h:17) without a
 * reader.  Returns -1 with grrt_last_error() set on an unreadable file. */
GRRT_EXPORT int grrt_ext_read_texture(const char* path, int* width, int* height, float* rgb, int64_t capacity);
/* Renders the RGB frame with the rows streamed to the CLI's two HDR outputs
 * (replaces the render-then-write sequence of reference cli/main.cpp:386-434):
 * bands of rows are rendered on the GPU, and each finished band is encoded
 * and appended to path_linear while the next one renders.  path_scaled (may
 * be NULL) then receives the frame scaled by 0.18 over the log-average
 * luminance of the lit pixels.  framebuffer (may be NULL) receives the RGBA
 * frame, equal to grrt_render's.  progress (may be NULL) is called per band on
 * the caller's thread.  Returns 0; 1 when cancelled (path_linear is completed
 * with black rows, the unrendered rows of framebuffer are zero, path_scaled is
 * not written); -1 on error. */
GRRT_EXPORT int grrt_ext_render_hdr(GRRTContext* ctx, const char* path_linear, const char* path_scaled,
                                    float* framebuffer, grrt_progress_fn progress, void* user_data);

#ifdef __cplusplus
}
#endif
#endif
