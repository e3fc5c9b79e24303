// Thin-disk emission pieces shared by the geodesic tracer and the surrogate
// shading pass (src/accretion_disk.cpp:96-147).
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "trace_params.hpp"

namespace grrt_amd {

// Redshift g = (p.u)_emit / (p.u)_obs for a circular emitter at rc and the
// static observer (src/accretion_disk.cpp:113-130).
__device__ __forceinline__ double disk_redshift(const ViewParams& V, double rc, double P0, double P3) {
    const double w = sqrt(V.M / (rc * rc * rc));
    const double aw = V.a * w;
    const double ut_e = 1.0 / sqrt(1.0 - 3.0 * V.M / rc + 2.0 * aw);
    const double up_e = (w / (1.0 + V.a * w)) * ut_e;
    const double pu_obs = P0 * V.ut_obs;
    const double pu_e = P0 * ut_e + P3 * up_e;
    return (__builtin_fabs(pu_obs) < 1e-30) ? 1.0 : pu_e / pu_obs;
}

// Colour added by a crossing at rc seen with redshift g: Novikov–Thorne
// flux (lerp of the view's table), T = T_peak (F/F_max)^(1/4), observed
// T g, chromaticity lerp (src/spectrum.cpp:115-124) times (T/1e5)^4 g^4.
__device__ __forceinline__ void thin_disk_color(const ViewParams& V, const double* flux, const double* s_spec,
                                                double rc, double g, double col[3]) {
    double F = 0.0;
    if (!(rc <= V.r_in || rc >= V.r_out || V.flux_max <= 0.0)) {
        const double frac = (rc - V.r_in) / (V.r_out - V.r_in) * (kFluxN - 1);
        const int idx = static_cast<int>(frac);
        const double t = frac - idx;
        F = (idx >= kFluxN - 1) ? flux[kFluxN - 1] : flux[idx] * (1.0 - t) + flux[idx + 1] * t;
    }
    if (F <= 0.0 || V.flux_max <= 0.0) return;
    const double T = V.t_peak * sqrt(sqrt(F / V.flux_max));
    if (T <= 0.0) return;
    const double Tobs = g * T;
    if (!(Tobs >= 100.0)) return;       // also drops a NaN temperature
    const double Tc = sclamp(Tobs, kSpecTmin, kSpecTmax);
    const double frac = (Tc - kSpecTmin) / (kSpecTmax - kSpecTmin) * (kSpecN - 1);
    const int idx = static_cast<int>(frac);
    const double t = frac - idx;
    const double ratio = T / kSpecTmax;
    const double lum = ratio * ratio * ratio * ratio;
    const double g4 = g * g * g * g;
    const double wgt = lum * g4;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const double cv = (idx >= kSpecN - 1)
                              ? s_spec[3 * (kSpecN - 1) + ch]
                              : s_spec[3 * idx + ch] * (1.0 - t) + s_spec[3 * (idx + 1) + ch] * t;
        col[ch] += cv * wgt;
    }
}

}  // namespace grrt_amd
