// Star background lookup shared by the tracers.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "trace_params.hpp"

namespace grrt_amd {

// Nearest star within 0.003 rad in the 3x3 neighbouring cells, first in
// cell-then-catalogue order (src/celestial_sphere.cpp:48-90).
__device__ __noinline__ double star_sample(const TraceParams& P, double theta, double phi) {
    const double twopi = 2.0 * kPi;
    theta = fmod(theta, twopi);
    if (theta < 0.0) theta += twopi;
    if (theta > kPi) {
        theta = twopi - theta;
        phi += kPi;
    }
    phi = fmod(phi, twopi);
    if (phi > kPi) phi -= twopi;
    if (phi < -kPi) phi += twopi;
    const int tb = static_cast<int>(theta * kStarGT / kPi);
    const int pb = static_cast<int>((phi + kPi) * kStarGP / twopi);
    const double st = sin(theta);
    for (int dt = -1; dt <= 1; ++dt) {
        int t = tb + dt;
        t = t < 0 ? 0 : (t > kStarGT - 1 ? kStarGT - 1 : t);
        for (int dp = -1; dp <= 1; ++dp) {
            const int p = (((pb + dp) % kStarGP) + kStarGP) % kStarGP;
            const int cell = t * kStarGP + p;
            const int e = P.star_start[cell + 1];
            for (int k = P.star_start[cell]; k < e; ++k) {
                const double sth = P.star_tpb[3 * k], sph = P.star_tpb[3 * k + 1];
                const double dth = theta - sth;
                double dph = phi - sph;
                if (dph > kPi) dph -= twopi;
                if (dph < -kPi) dph += twopi;
                const double d2 = dth * dth + dph * dph * st * st;
                if (d2 < 0.003 * 0.003) return P.star_tpb[3 * k + 2];
            }
        }
    }
    return 0.0;
}

struct TexRGB {
    double c[3];
};

// The texture background at an escape direction (theta, phi), folded as
// star_sample folds it (theta into [0, pi], phi into [-pi, pi]): bilinear
// between texel centres, wrapping in phi and clamped at the poles; linear
// RGB.  Returned by value so that callers keep their colour in registers.
__device__ __noinline__ TexRGB texture_sample(const TraceParams& P, double theta, double phi) {
    const double twopi = 2.0 * kPi;
    theta = fmod(theta, twopi);
    if (theta < 0.0) theta += twopi;
    if (theta > kPi) {
        theta = twopi - theta;
        phi += kPi;
    }
    phi = fmod(phi, twopi);
    if (phi > kPi) phi -= twopi;
    if (phi < -kPi) phi += twopi;
    const int W = P.bg_w, H = P.bg_h;
    const double x = (phi + kPi) / twopi * W - 0.5, y = theta / kPi * H - 0.5;
    const double fx0 = floor(x), fy0 = floor(y);
    const double tx = x - fx0;
    double ty = y - fy0;
    int x0 = static_cast<int>(fx0) % W;
    if (x0 < 0) x0 += W;
    const int x1 = x0 + 1 < W ? x0 + 1 : 0;
    int y0 = static_cast<int>(fy0), y1 = y0 + 1;
    if (y0 < 0) {
        y0 = y1 = 0;
        ty = 0.0;
    } else if (y1 > H - 1) {
        y0 = y1 = H - 1;
        ty = 0.0;
    }
    const float* r0 = P.bg_tex + static_cast<long long>(y0) * W * 3;
    const float* r1 = P.bg_tex + static_cast<long long>(y1) * W * 3;
    TexRGB out;
    for (int k = 0; k < 3; ++k) {
        const double a = (1.0 - tx) * r0[3 * x0 + k] + tx * r0[3 * x1 + k];
        const double b = (1.0 - tx) * r1[3 * x0 + k] + tx * r1[3 * x1 + k];
        out.c[k] = (1.0 - ty) * a + ty * b;
    }
    return out;
}

}  // namespace grrt_amd
