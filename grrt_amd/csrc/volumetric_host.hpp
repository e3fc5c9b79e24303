// Host-side construction of the volumetric (Shakura–Sunyaev) disk tables:
// Saha/opacity tables, radial structure, vertical hydrostatic/radiative
// profiles, density normalisation and the noise permutation.  The tables
// are uploaded once per context and read by the GPU ray-marcher.
#pragma once
#include <cstdint>
#include <vector>

namespace grrt_amd {

struct VolParams {
    double alpha = 0.1;
    double turbulence = 0.4;
    uint32_t seed = 42;
    double tau_mid = 100.0;
    double noise_scale = 0.0;   // 0: twice the scale height at the peak-flux radius
    int noise_octaves = 2;
    double opacity_nu_min = 1e14;
    double opacity_nu_max = 1e16;
};

// Ionisation state of the H/He/C/O/Fe mixture (Saha equation).
struct Ionization {
    double n_e, mu, n_ion_eff, n_HI, n_Hminus;
    double n_pop[5][5];
};

Ionization saha(double rho_cgs, double T);
double kappa_abs_point(double nu, double rho_cgs, double T, const Ionization& ion);
double kappa_es_point(double rho_cgs, const Ionization& ion);
double alpha_ff_point(double nu, double T, const Ionization& ion);
double alpha_hminus_point(double nu, double T, const Ionization& ion);
double alpha_bf_point(double nu, double T, const Ionization& ion);
double planck_nu_host(double nu, double T);

struct OpacityTables {
    int n_nu = 0, n_rho = 100, n_T = 100;
    double log_nu_min = 0, log_nu_max = 0, log_rho_min = 0, log_rho_max = 0, log_T_min = 0,
           log_T_max = 0;
    std::vector<double> kappa_abs;   // [nu][rho][T]
    std::vector<double> kappa_es;    // [rho][T]
    std::vector<double> kappa_ross;  // [rho][T]
    std::vector<double> mu;          // [rho][T]
    double lookup_es(double rho, double T) const;
    double lookup_ross(double rho, double T) const;
    double lookup_mu(double rho, double T) const;
};

OpacityTables build_opacity_tables(double rho_min, double rho_max, double T_min, double T_max,
                                   double nu_min, double nu_max);

struct VolTables {
    double M = 1, a = 0, r_outer = 30, t_peak = 1e7;
    VolParams prm;
    double r_isco = 0, r_horizon = 0, r_min = 0, taper_width = 0, E_isco = 0, L_isco = 0;
    double rho_scale = 1.0, noise_scale = 1.0;
    int n_r = 500, n_z = 128;
    std::vector<double> H, rho_mid, T_eff, z_max;   // n_r
    std::vector<double> rho_profile, T_profile;     // n_r * n_z
    OpacityTables op;
    int perm[512];
};

// Builds everything (src/volumetric_disk.cpp:33-73 construction order).
void build_volumetric(double M, double a, double r_outer, double t_peak, const VolParams& p,
                      VolTables& out);

}  // namespace grrt_amd
