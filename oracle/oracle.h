/*
 * CPU oracle for the per-pixel null-geodesic + thin-disk path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the
 * reference CPU algorithm used as the checker for the HIP path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library (libgrrt.so) never links, loads or calls it.
 *
 * Each function cites the reference file:line it restates.  The reference
 * itself is not buildable in this image (its headers include a CMake-
 * generated export header), so every function here is pinned by closed
 * forms, invariants and known-answer tests instead (see tests/test_oracle_*).
 */
#ifndef GRRT_ORACLE_H
#define GRRT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene description shared by every oracle entry point (geometric units). */
typedef struct {
    int width, height;
    double mass;        /* M */
    double spin_a;      /* dimensional a = a_* M */
    double r_obs, theta_obs, phi_obs, fov;
    int disk_enabled;
    double r_in, r_out, t_peak;
    int stars;          /* 1 = 5000-star background (seed 42) */
    double tol;
    int max_steps;
    double r_escape;    /* reference: 1000 */
    int integrator;     /* 0: adaptive Dormand-Prince (src/rk4.cpp:87-136, the
                           reference tracer's); 1: step-doubling RK4
                           (src/rk4.cpp:33-80) */
} OracleScene;

/* Per-ray record.  term: 0 horizon, 1 escaped, 2 max steps.
 * cls: 0 captured, 1 escaped, 2 disk hit, 3 max steps (disk hit wins). */
typedef struct {
    double color[3];
    double x_final[4];
    double p_final[4];
    double r_hit, g_hit;      /* first in-range disk crossing (0 if none) */
    double h_max, h_final;    /* |H| max over accepted states, |H| at the end */
    int32_t term, cls;
    int32_t steps, attempts, n_cross, pad;
} OracleRay;

/* Volumetric disk tables (built by the library's host set-up, which the
 * tests pin separately) and scalars, consumed by the volumetric oracle. */
typedef struct {
    double M, a, r_horizon, r_min, r_outer, r_isco, taper_w, E_isco, L_isco;
    double rho_scale, noise_scale, turbulence;
    int octaves, n_r, n_z;
    const double *H, *z_max, *rho_mid, *rho_prof, *T_prof;
    const double *kabs, *kes;           /* kappa_abs [nu][rho][T], kappa_es [rho][T] */
    int n_nu, n_rho, n_T;
    double lnu0, lnu1, lrho0, lrho1, lT0, lT1;
    const int* perm;                    /* 512 */
} OracleVol;

/* Sample placement modes for render/trace_rays. */
enum { ORACLE_JITTER_HASH = 0, ORACLE_JITTER_CENTER = 1, ORACLE_JITTER_SEEDED = 2 };

void   oracle_fast_sincos(double x, double* s, double* c);
void   oracle_kerr_derivs(double M, double a, const double x[4], const double p[4],
                          double dx[4], double dp[4]);
void   oracle_g_lower(double M, double a, const double x[4], double g[16]);
double oracle_hamiltonian(double M, double a, const double x[4], const double p[4]);
double oracle_horizon(double M, double a);
double oracle_isco(double M, double a);
void   oracle_dp45_attempt(double M, double a, const double y[8], double dl,
                           double y5[8], double* err);
void   oracle_rk4_step(double M, double a, const double y[8], double dl, double out[8]);
double oracle_adaptive_step(double M, double a, const double y[8], double dl, double tol,
                            double out[8], int* attempts);
double oracle_adaptive_step_doubling(double M, double a, const double y[8], double dl, double tol,
                                     double out[8], int* attempts);
double oracle_pixel_hash(int i, int j, int s, int ch);
void   oracle_seeded_jitter(uint64_t seed, int64_t sample_index, double* jx, double* jy);
void   oracle_camera_ray(const OracleScene* sc, double px, double py, double state[8]);
void   oracle_flux_lut(const OracleScene* sc, double* lut500, double* flux_max);
void   oracle_spectrum_lut(double* rgb3000);
void   oracle_stars(double* theta_phi_b15000);
void   oracle_star_sample(double theta, double phi, double out[3]);
/* src/accretion_disk.cpp:132-147 for a crossing at rc with p_t, p_phi. */
void   oracle_disk_emission(const OracleScene* sc, double rc, double p_t, double p_phi,
                            double out[3], double* g, double* T);
void   oracle_trace(const OracleScene* sc, const double state[8], OracleRay* out);
/* Trace an explicit list of sample ids (pixel*spp + s); rays may be NULL. */
void   oracle_trace_samples(const OracleScene* sc, int mode, int spp, uint64_t seed,
                            const int64_t* ids, int64_t n, OracleRay* rays, int threads);
/* Full frame: RGBA float W*H*4, per-sample classes (may be NULL). Returns
 * the number of accepted steps summed over all rays. */
int64_t oracle_render(const OracleScene* sc, int mode, int spp, uint64_t seed,
                      float* rgba, uint8_t* cls, int threads);

/* Volumetric path (src/geodesic_tracer.cpp:23-337 with a volumetric disk,
 * :495-746 spectral).  nbins == 0: RGB ray-march, colours in rays[k].color;
 * nbins > 0: spectral, spectra[k*nbins + b] = I_nu [erg/s/cm^2/Hz/sr]. */
void   oracle_vol_trace_samples(const OracleScene* sc, const OracleVol* V, int mode, int spp,
                                uint64_t seed, const int64_t* ids, int64_t n, int nbins,
                                const double* nu, OracleRay* rays, double* spectra, int threads);
double oracle_vol_density(const OracleVol* V, double r, double z, double phi);
double oracle_vol_temperature(const OracleVol* V, double r, double z);
double oracle_vol_kappa_abs(const OracleVol* V, double nu, double rho, double T);
double oracle_vol_kappa_es(const OracleVol* V, double rho, double T);
double oracle_simplex(const int* perm, double x, double y, double z);
// Gas 4-velocity (u^t, u^r, u^phi) the ray-march uses at (r, theta).
void oracle_vol_velocity(const OracleVol* V, double r, double theta, double* u);
// n emitting samples of the backward transfer on one channel, in order.
void oracle_rt_steps(int n, const double* S, const double* dtau, double* J, double* T);

#ifdef __cplusplus
}
#endif
#endif
