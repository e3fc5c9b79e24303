// Host-side scene construction for the GPU tracer: parameter defaulting,
// Kerr constants, camera tetrad, thin-disk flux table, blackbody chromaticity
// table and the star catalogue in cell-sorted (CSR) form.
#pragma once
#include <cstdint>
#include <vector>

#include "grrt/types.h"

namespace grrt_amd {

// Parameters after the reference's zero-means-default rules
// (src/api.cpp:43-173).
struct Resolved {
    int W, H;
    double M, a;              // a is dimensional (a_* M)
    double r_obs, th_obs, ph_obs, fov;
    int max_steps;
    double tol;
    int disk;                 // thin disk active (enabled and not volumetric)
    int volumetric;
    double r_in, r_out, t_disk;
    int stars;
    int texture;              // equirectangular background image (GRRT_BG_TEXTURE)
    int spp;
    double r_h, r_isco;
};

Resolved resolve_params(const GRRTParams& p);

double kerr_horizon(double M, double a);
double kerr_isco(double M, double a);

// Static-observer tetrad (src/camera.cpp:29-63) and covariant metric there.
struct CameraFrame {
    double g[4][4];
    double e0[4], e1[4], e2[4], e3[4];
};
CameraFrame build_camera(double M, double a, double r, double th, double ph);

// Novikov–Thorne flux table (src/accretion_disk.cpp:45-94); returns max.
double build_flux_table(double M, double a, double r_in, double r_out, std::vector<double>& lut);

// Normalised linear-sRGB chromaticity of a blackbody, 1000 temperatures in
// [1e3, 1e5] K (src/spectrum.cpp:56-96), interleaved RGB.
void build_chromaticity_table(std::vector<double>& lut);

// 5000-star catalogue from mt19937(42) (src/celestial_sphere.cpp:9-46),
// bucketed on the 180x360 grid; stars within a cell keep catalogue order.
struct StarTable {
    std::vector<int> start;        // 180*360 + 1 offsets
    std::vector<double> tpb;       // (theta, phi, brightness) per star, cell order
};
void build_star_table(StarTable& t);

}  // namespace grrt_amd
