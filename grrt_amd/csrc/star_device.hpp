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

}  // namespace grrt_amd
