// Host-side construction of the volumetric (Shakura–Sunyaev) disk tables:
// Saha/opacity tables, radial structure, vertical hydrostatic/radiative
// profiles, density normalisation and the noise permutation.  The tables
// are uploaded once per context and read by the GPU ray-marcher.
#pragma once
#include <cstdint>
#include <vector>

#include "opacity_math.hpp"

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

// Table geometry (src/opacity.cpp:237-262): 100 x 100 log-spaced (rho, T)
// cells over the given ranges, max(20, 10 per decade) absorption
// frequencies; the tables are allocated and zero.
OpacityTables opacity_table_shape(double rho_min, double rho_max, double T_min, double T_max,
                                  double nu_min, double nu_max);
// Every cell of the tables (src/opacity.cpp:263-303), on host threads, or on
// the GPU (one lane per cell, results copied back; stream-ordered and
// synchronised before returning).
void fill_opacity_tables_host(OpacityTables& o);
hipError_t fill_opacity_tables_device(OpacityTables& o, hipStream_t st);

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
    double build_ms_opacity = 0.0, build_ms_total = 0.0;   // wall time of the set-up
    double build_ms_columns = 0.0;       // wall time of the vertical columns (host or GPU)
    double build_ms_columns_kernel = 0.0;   // GPU build: the column kernel alone (events)
};

// Builds everything (src/volumetric_disk.cpp:33-73 construction order).
// With a stream, the opacity tables are computed on the GPU on that stream,
// and the vertical columns too unless device_columns is false; without one,
// both on host threads.  The radial structure and the normalisation are
// host code.  Either way the tables end in host memory.
void build_volumetric(double M, double a, double r_outer, double t_peak, const VolParams& p,
                      VolTables& out, const hipStream_t* device_stream = nullptr, bool device_columns = true);

}  // namespace grrt_amd
