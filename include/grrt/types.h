/*
 * GRRTParams and the enums of the C ABI.  Field order, types and enum
 * values are those of the reference's include/grrt/types.h:4-64, so a
 * caller compiled against the reference header links unchanged.  The
 * layout (240 bytes on LP64) is checked by static_asserts in the library.
 */
#ifndef GRRT_TYPES_H
#define GRRT_TYPES_H

typedef enum {
    GRRT_METRIC_SCHWARZSCHILD = 0,
    GRRT_METRIC_KERR = 1
} GRRTMetricType;

/* GRRT_BACKEND_CUDA keeps the reference's numbering and selects the GPU
 * (HIP, MI355X) backend; this library has only that backend. */
typedef enum {
    GRRT_BACKEND_CPU = 0,
    GRRT_BACKEND_CUDA = 1
} GRRTBackend;

/* GRRT_BG_TEXTURE with background_texture_path naming an equirectangular
 * Radiance .hdr or PFM image samples it at each escaped ray's direction;
 * with no path it is black, as the reference renders the mode. */
typedef enum {
    GRRT_BG_BLACK = 0,
    GRRT_BG_STARS = 1,
    GRRT_BG_TEXTURE = 2
} GRRTBackgroundType;

typedef struct {
    int width;
    int height;

    GRRTMetricType metric_type;
    double mass;
    double spin;                 /* dimensionless a/M; Kerr spin <= 0 means 0.998 */

    double observer_r;
    double observer_theta;
    double observer_phi;
    double fov;                  /* horizontal field of view, radians */

    double cam_yaw;              /* accepted, unused (as in the reference) */
    double cam_pitch;
    double cam_roll;

    int disk_enabled;
    double disk_inner;           /* 0 = ISCO */
    double disk_outer;           /* 0 = 20 M thin / 30 M volumetric */
    double disk_temperature;     /* 0 = 1e7 K */

    int disk_volumetric;
    double disk_alpha;
    double disk_turbulence;
    int disk_seed;
    double disk_noise_scale;
    int disk_noise_octaves;

    double mass_solar;
    double eddington_fraction;

    GRRTBackgroundType background_type;
    const char* background_texture_path;

    double integrator_tolerance; /* 0 = 1e-8 */
    int integrator_max_steps;    /* 0 = 10000 */

    int samples_per_pixel;
    int thread_count;            /* accepted, unused */
    GRRTBackend backend;

    int num_frequency_bins;
    const double* frequency_bins_hz;
} GRRTParams;

typedef struct GRRTContext GRRTContext;

#endif
