// Pixel sample placement and the static-observer camera covector, shared by
// the thin-disk and volumetric tracers.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "kerr_rhs.hpp"
#include "trace_params.hpp"

namespace grrt_amd {

// Fractional pixel position of ray `id` (src/renderer.cpp:48-55 and the
// extension sampling modes).
__device__ __forceinline__ void sample_position(const TraceParams& P, long long id, double& px,
                                                double& py) {
    const long long pl = id / P.nS;
    const int s = static_cast<int>(id - pl * P.nS);
    const int i = P.x0 + static_cast<int>(pl % P.tw);
    const int j = P.y0 + static_cast<int>(pl / P.tw);
    if (P.mode == GRRT_SAMPLING_CENTER) {
        px = i + 0.5;
        py = j + 0.5;
    } else if (P.mode == GRRT_SAMPLING_HASH) {
        auto hash = [](int ii, int jj, int ss, int ch) -> double {
            uint32_t h = static_cast<uint32_t>(ii * 73856093u ^ jj * 19349663u ^ ss * 83492791u ^
                                               ch * 45678917u);
            h ^= h >> 16;
            h *= 0x45d9f3bu;
            h ^= h >> 16;
            return (h & 0xFFFFu) / 65536.0;
        };
        const int sy = s / P.grid, sx = s - sy * P.grid;
        px = i + (sx + hash(i, j, s, 0)) * P.cell;
        py = j + (sy + hash(i, j, s, 1)) * P.cell;
    } else {
        // Seeded: pixel-major global sample index over the full frame.
        const long long gid = (static_cast<long long>(j) * P.W + i) * P.nS + s;
        const uint64_t key = P.seed * 0xD1B54A32D192ED03ull;
        const uint64_t w = 0x9E3779B97F4A7C15ull;
        const uint64_t u = static_cast<uint64_t>(gid) * 2u;
        const double jx = static_cast<double>(mix64(key + (u + 1u) * w) >> 11) * 0x1.0p-53;
        const double jy = static_cast<double>(mix64(key + (u + 2u) * w) >> 11) * 0x1.0p-53;
        px = i + jx;
        py = j + jy;
    }
}

// Camera ray (src/camera.cpp:69-96): p^mu = -e0 + d, lowered with g at the
// observer; only the non-zero metric entries enter.  Writes the conserved
// p_t, p_phi and the evolving (r, theta, phi, p_r, p_theta).
__device__ __forceinline__ void camera_ray(const TraceParams& P, long long id, double& P0, double& P3, Vec5& y) {
    double px, py;
    sample_position(P, id, px, py);
    const double alpha = (px - P.W / 2.0) * P.fov / P.W;
    const double beta = (py - P.H / 2.0) * P.fov / P.W;
    double sa, ca_, sb, cb;
    sincos(alpha, &sa, &ca_);
    sincos(beta, &sb, &cb);
    const double ca = cb * sa, cc = cb * ca_;
    double pc[4];
    pc[0] = -P.e0t + (-ca * P.e1[0] - sb * P.e2[0] + cc * P.e3[0]);
#pragma unroll
    for (int m = 1; m < 4; ++m) pc[m] = -ca * P.e1[m] - sb * P.e2[m] + cc * P.e3[m];
    P0 = P.g00 * pc[0] + P.g03 * pc[3];
    P3 = P.g03 * pc[0] + P.g33 * pc[3];
    y.v[0] = P.r_obs;
    y.v[1] = P.th_obs;
    y.v[2] = P.ph_obs;
    y.v[3] = P.g11 * pc[1];
    y.v[4] = P.g22 * pc[2];
}

}  // namespace grrt_amd
