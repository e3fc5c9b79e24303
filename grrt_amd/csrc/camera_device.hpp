// Pixel sample placement and the static-observer camera covector, shared by
// the thin-disk and volumetric tracers.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "kerr_rhs.hpp"
#include "trace_params.hpp"

namespace grrt_amd {

struct SampleId {
    int i, j, s;           // pixel (frame coordinates) and sample
    long long local;       // pixel-major sample index within the view's window
};

__device__ __forceinline__ bool decode_ray(const ViewParams& V, long long q, SampleId& o) {
    const int nS = V.nS;
    if (V.tile <= 0) {
        const long long pl = q / nS;
        o.s = static_cast<int>(q - pl * nS);
        o.i = V.x0 + static_cast<int>(pl % V.tw);
        o.j = V.y0 + static_cast<int>(pl / V.tw);
        o.local = q;
        return true;
    }
    const int T = V.tile;
    const long long per = static_cast<long long>(T) * T * nS;
    const long long t = q / per;
    const int w = static_cast<int>(q - t * per);
    const int nx = (V.tw + T - 1) / T;
    const int tx = static_cast<int>(t % nx), ty = static_cast<int>(t / nx);
    const int wp = w / nS;
    o.s = w - wp * nS;
    const int lx = tx * T + wp % T, ly = ty * T + wp / T;
    if (lx >= V.tw || ly >= V.th) return false;
    o.i = V.x0 + lx;
    o.j = V.y0 + ly;
    o.local = (static_cast<long long>(ly) * V.tw + lx) * nS + o.s;
    return true;
}

// Fractional pixel position of a sample (src/renderer.cpp:48-55 and the
// extension sampling modes).
__device__ __forceinline__ void sample_position(const ViewParams& V, const SampleId& d, double& px,
                                                double& py) {
    const int i = d.i, j = d.j, s = d.s;
    if (V.mode == GRRT_SAMPLING_CENTER) {
        px = i + 0.5;
        py = j + 0.5;
    } else if (V.mode == GRRT_SAMPLING_HASH) {
        auto hash = [](int ii, int jj, int ss, int ch) -> double {
            uint32_t h = static_cast<uint32_t>(ii * 73856093u ^ jj * 19349663u ^ ss * 83492791u ^
                                               ch * 45678917u);
            h ^= h >> 16;
            h *= 0x45d9f3bu;
            h ^= h >> 16;
            return (h & 0xFFFFu) / 65536.0;
        };
        const int sy = s / V.grid, sx = s - sy * V.grid;
        px = i + (sx + hash(i, j, s, 0)) * V.cell;
        py = j + (sy + hash(i, j, s, 1)) * V.cell;
    } else {
        // Seeded: pixel-major global sample index over the full frame.
        const long long gid = (static_cast<long long>(j) * V.W + i) * V.nS + s;
        const uint64_t key = V.seed * 0xD1B54A32D192ED03ull;
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
__device__ __forceinline__ void camera_ray(const ViewParams& V, const SampleId& d, double& P0, double& P3,
                                           Vec5& y) {
    double px, py;
    sample_position(V, d, px, py);
    const double alpha = (px - V.W / 2.0) * V.fov / V.W;
    const double beta = (py - V.H / 2.0) * V.fov / V.W;
    double sa, ca_, sb, cb;
    sincos(alpha, &sa, &ca_);
    sincos(beta, &sb, &cb);
    const double ca = cb * sa, cc = cb * ca_;
    double pc[4];
    pc[0] = -V.e0t + (-ca * V.e1[0] - sb * V.e2[0] + cc * V.e3[0]);
#pragma unroll
    for (int m = 1; m < 4; ++m) pc[m] = -ca * V.e1[m] - sb * V.e2[m] + cc * V.e3[m];
    P0 = V.g00 * pc[0] + V.g03 * pc[3];
    P3 = V.g03 * pc[0] + V.g33 * pc[3];
    y.v[0] = V.r_obs;
    y.v[1] = V.th_obs;
    y.v[2] = V.ph_obs;
    y.v[3] = V.g11 * pc[1];
    y.v[4] = V.g22 * pc[2];
}

}  // namespace grrt_amd
