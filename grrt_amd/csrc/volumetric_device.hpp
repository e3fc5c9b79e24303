// Device-side volumetric disk queries, opacities and blackbody source, for
// the ray-marcher (src/volumetric_disk.cpp:79-273, src/noise.cpp:38-115,
// src/opacity.cpp:209-365).  Templated on the radiative-transfer float type
// so the same code serves the all-FP64 path and the FP32 path.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "trace_params.hpp"

namespace grrt_amd {

// Table index kept inside [0, hi] whatever the float-to-int conversion of a
// non-finite coordinate produced.
__device__ __forceinline__ int clamp_index(int i, int hi) { return i < 0 ? 0 : (i > hi ? hi : i); }

namespace vcgs {
constexpr double h = 6.626070e-27, k_B = 1.380649e-16, c = 2.997924e10;
}

// The radial tables (scale height, z_max, midplane density), read from
// wherever the caller staged them (LDS in the tracer).
struct VolRadial {
    const double *H, *z_max, *rho_mid;
};

__device__ __forceinline__ double vol_interp_radial(const VolDev& V, const double* lut, double r) {
    if (r <= V.r_min) return lut[0];
    if (r >= V.r_outer) return lut[V.n_r - 1];
    const double frac = (r - V.r_min) * V.r_scale;
    const int idx = clamp_index(static_cast<int>(frac), V.n_r - 2);
    const double t = frac - idx;
    return lut[idx] * (1.0 - t) + lut[idx + 1] * t;
}

// Per-column z-normalised 2-D lookup (src/volumetric_disk.cpp:175-214).
__device__ double vol_interp_2d(const VolDev& V, const VolRadial& R, const double* lut, double r, double za) {
    double rf = (r - V.r_min) * V.r_scale;
    rf = rf < 0.0 ? 0.0 : (rf > V.n_r - 1 ? static_cast<double>(V.n_r - 1) : rf);
    const int ri = clamp_index(static_cast<int>(rf), V.n_r - 2);
    const double tr = rf - ri;
    const int ri1 = ri + 1 < V.n_r - 1 ? ri + 1 : V.n_r - 1;
    const double zlo = R.z_max[ri], zhi = R.z_max[ri1];
    if ((zlo <= 0.0 || za >= zlo) && (zhi <= 0.0 || za >= zhi)) return 0.0;
    const int nz = V.n_z;
    auto column = [&](int rr, double zm) -> double {
        if (!(zm > 0.0 && za < zm)) return 0.0;
        double zf = za * rcp(zm) * (nz - 1);
        zf = zf < 0.0 ? 0.0 : (zf > nz - 1 ? static_cast<double>(nz - 1) : zf);
        const int zi = clamp_index(static_cast<int>(zf), nz - 2);
        const double tz = zf - zi;
        return (1.0 - tz) * lut[rr * nz + zi] + tz * lut[rr * nz + zi + 1];
    };
    return (1.0 - tr) * column(ri, zlo) + tr * column(ri1, zhi);
}

__device__ __forceinline__ double vol_H(const VolDev& V, const VolRadial& R, double r) {
    return vol_interp_radial(V, R.H, r);
}
__device__ __forceinline__ double vol_zmax(const VolDev& V, const VolRadial& R, double r) {
    return vol_interp_radial(V, R.z_max, r);
}

__device__ __forceinline__ bool vol_inside(const VolDev& V, const VolRadial& R, double r, double z) {
    if (r <= V.r_horizon || r > V.r_outer) return false;
    return __builtin_fabs(z) < vol_zmax(V, R, r) + 1.5 * vol_H(V, R, r);
}

__device__ __forceinline__ double vol_taper(const VolDev& V, double r) {
    if (r >= V.r_isco) return 1.0;
    const double dr = V.r_isco - r;
    return exp(-(dr * dr) / (V.taper_w * V.taper_w));
}

// 3-D simplex noise (Perlin 2001; the corner ranking, gradient set and
// (0.6 - r^2)^4 kernel of src/noise.cpp:38-96), permutation from LDS.
__device__ double simplex3(const int* perm, double x, double y, double z) {
    constexpr double F3 = 1.0 / 3.0, G3 = 1.0 / 6.0;
    const double s = (x + y + z) * F3;
    const int i = static_cast<int>(floor(x + s)), j = static_cast<int>(floor(y + s)),
              k = static_cast<int>(floor(z + s));
    const double t = (i + j + k) * G3;
    const double x0 = x - (i - t), y0 = y - (j - t), z0 = z - (k - t);
    int i1, j1, k1, i2, j2, k2;
    if (x0 >= y0) {
        if (y0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
        else if (x0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 0; k2 = 1; }
        else { i1 = 0; j1 = 0; k1 = 1; i2 = 1; j2 = 0; k2 = 1; }
    } else {
        if (y0 < z0) { i1 = 0; j1 = 0; k1 = 1; i2 = 0; j2 = 1; k2 = 1; }
        else if (x0 < z0) { i1 = 0; j1 = 1; k1 = 0; i2 = 0; j2 = 1; k2 = 1; }
        else { i1 = 0; j1 = 1; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
    }
    const int ii = i & 255, jj = j & 255, kk = k & 255;
    // 12 cube-edge gradients, gradient index g % 12.
    auto corner = [](double cx, double cy, double cz, int g) -> double {
        double t0 = 0.6 - cx * cx - cy * cy - cz * cz;
        if (t0 < 0.0) return 0.0;
        t0 *= t0;
        g %= 12;
        const double gx = (g < 8) ? ((g & 1) ? -1.0 : 1.0) : 0.0;
        const double gy = (g < 4) ? ((g & 2) ? -1.0 : 1.0) : ((g < 8) ? 0.0 : ((g & 1) ? -1.0 : 1.0));
        const double gz = (g < 4) ? 0.0 : ((g < 8) ? ((g & 2) ? -1.0 : 1.0) : ((g & 2) ? -1.0 : 1.0));
        return t0 * t0 * (gx * cx + gy * cy + gz * cz);
    };
    const double n0 = corner(x0, y0, z0, perm[ii + perm[jj + perm[kk]]]);
    const double n1 = corner(x0 - i1 + G3, y0 - j1 + G3, z0 - k1 + G3, perm[ii + i1 + perm[jj + j1 + perm[kk + k1]]]);
    const double n2 = corner(x0 - i2 + 2 * G3, y0 - j2 + 2 * G3, z0 - k2 + 2 * G3,
                             perm[ii + i2 + perm[jj + j2 + perm[kk + k2]]]);
    const double n3 = corner(x0 - 1 + 3 * G3, y0 - 1 + 3 * G3, z0 - 1 + 3 * G3,
                             perm[ii + 1 + perm[jj + 1 + perm[kk + 1]]]);
    return 32.0 * (n0 + n1 + n2 + n3);
}

__device__ double fbm3(const int* perm, double x, double y, double z, int octaves) {
    double res = 0.0, amp = 1.0, freq = 1.0;
    for (int o = 0; o < octaves; ++o) {
        res += amp * simplex3(perm, x * freq, y * freq, z * freq);
        amp *= 0.5;
        freq *= 2.0;
    }
    return res;
}

// Density [g/cm^3] with photosphere soft edge, ISCO taper and fBm turbulence.
__device__ double vol_density(const VolDev& V, const VolRadial& R, const int* perm, double r, double z, double phi) {
    if (r <= V.r_horizon || r > V.r_outer) return 0.0;
    const double za = __builtin_fabs(z);
    const double zm = vol_zmax(V, R, r), H = vol_H(V, R, r);
    double edge = 1.0, zl = za;
    if (za >= zm) {
        const double dz = za - zm;
        if (dz >= 1.5 * H) return 0.0;
        edge = exp(-2.0 * (dz / H) * (dz / H));
        zl = zm * 0.9999;
    }
    const double base = vol_interp_radial(V, R.rho_mid, r) * vol_interp_2d(V, R, V.rho_prof, r, zl) *
                        V.rho_scale * vol_taper(V, r) * edge;
    double sp, cp;
    sincos_cephes(phi, sp, cp);
    const double n = fbm3(perm, r * cp * V.inv_noise, r * sp * V.inv_noise, z * V.inv_noise, V.octaves);
    return smax(0.0, base * (1.0 + V.turbulence * n));
}

__device__ double vol_temperature(const VolDev& V, const VolRadial& R, double r, double za) {
    if (r <= V.r_horizon || r > V.r_outer) return 0.0;
    const double zm = vol_zmax(V, R, r), H = vol_H(V, R, r);
    if (za >= zm + 1.5 * H) return 0.0;
    return vol_interp_2d(V, R, V.T_prof, r, smin(za, zm * 0.9999));
}

// Emitter 4-velocity: circular prograde orbit outside the ISCO, plunging
// with the ISCO energy and angular momentum inside.
__device__ void vol_velocity(const VolDev& V, double r, double theta, double& ut, double& ur, double& up) {
    const double M = V.M, a = V.a;
    const double ir = rcp(r);                    // the divisions by r as products
    if (r >= V.r_isco) {
        const double sM = sqrt(M);
        const double Om = sM * rcp(r * sqrt(r) + a * sM);
        const double gtt = -(1.0 - 2.0 * M * ir), gtp = -2.0 * M * a * ir, gpp = r * r + a * a + 2.0 * M * a * a * ir;
        const double den = -(gtt + 2.0 * gtp * Om + gpp * Om * Om);
        ut = rcp(sqrt(smax(den, 1e-30)));
        ur = 0.0;
        up = Om * ut;
        return;
    }
    const double E = V.E_isco, L = V.L_isco;
    const double Del = r * r - 2.0 * M * r + a * a;
    const double iDel = rcp(Del);
    double st, ct;
    sincos_cephes(theta, st, ct);
    const double Sig = r * r + a * a * ct * ct;
    ut = (E * (r * r + a * a + 2.0 * M * a * a * ir) - 2.0 * M * a * L * ir) * iDel;
    up = (L * (1.0 - 2.0 * M * ir) + 2.0 * M * a * E * ir) * iDel;
    const double EaL = E * (r * r + a * a) - a * L, LaE = L - a * E;
    const double Rr = EaL * EaL - Del * (r * r + LaE * LaE);
    ur = -sqrt(smax(0.0, Rr)) * rcp(Sig);
}

// Cell index and fraction on a log-spaced axis, clamped to the table.
template <class T>
__device__ __forceinline__ void vol_log_cell(T lv, T lo, T scale, int n, int& idx, T& frac) {
    T t = (lv - lo) * scale;     // scale = (n - 1) / (hi - lo)
    t = t < T(0) ? T(0) : (t > static_cast<T>(n - 2) ? static_cast<T>(n - 2) : t);
    idx = clamp_index(static_cast<int>(t), n - 2);
    frac = t - static_cast<T>(idx);
}

// Precomputed per-step (rho, T) cell of the opacity tables.
template <class T>
struct OpCell {
    int irho, iT;
    T frho, fT;
};

template <class T>
__device__ __forceinline__ OpCell<T> vol_op_cell(const VolDev& V, double rho, double Temp) {
    OpCell<T> c;
    vol_log_cell<T>(static_cast<T>(log10(rho)), static_cast<T>(V.lrho0), static_cast<T>(V.rho_scale_ax), V.n_rho, c.irho,
                    c.frho);
    vol_log_cell<T>(static_cast<T>(log10(Temp)), static_cast<T>(V.lT0), static_cast<T>(V.T_scale), V.n_T, c.iT, c.fT);
    return c;
}

// kappa_es bilinear in (log rho, log T) cell coordinates (src/opacity.cpp:339-346).
// The opacity tables in the transfer's own precision.
__device__ __forceinline__ const double* vol_kes_table(const VolDev& V, double) { return V.kes; }
__device__ __forceinline__ const float* vol_kes_table(const VolDev& V, float) { return V.kes_f; }
__device__ __forceinline__ const double* vol_lnk_table(const VolDev& V, double) { return V.ln_kabs; }
__device__ __forceinline__ const float* vol_lnk_table(const VolDev& V, float) { return V.ln_kabs_f; }

template <class T>
__device__ __forceinline__ T vol_kes(const VolDev& V, const OpCell<T>& c) {
    const int n = V.n_T;
    const auto* q = vol_kes_table(V, T(0)) + c.irho * n + c.iT;
    return static_cast<T>(q[0]) * (T(1) - c.frho) * (T(1) - c.fT) + static_cast<T>(q[1]) * (T(1) - c.frho) * c.fT +
           static_cast<T>(q[n]) * c.frho * (T(1) - c.fT) + static_cast<T>(q[n + 1]) * c.frho * c.fT;
}

// kappa_abs trilinear in ln kappa (src/opacity.cpp:307-336); lognu = log10(nu).
template <class T>
__device__ __forceinline__ T vol_kabs(const VolDev& V, const OpCell<T>& c, double lognu) {
    int inu;
    T fnu;
    vol_log_cell<T>(static_cast<T>(lognu), static_cast<T>(V.lnu0), static_cast<T>(V.nu_scale), V.n_nu, inu, fnu);
    const int nrt = V.n_rho * V.n_T;
    const auto* q = vol_lnk_table(V, T(0)) + inu * nrt + c.irho * V.n_T + c.iT;
    const int n = V.n_T;
    const T l00 = static_cast<T>(q[0]) * (T(1) - c.fT) + static_cast<T>(q[1]) * c.fT;
    const T l01 = static_cast<T>(q[n]) * (T(1) - c.fT) + static_cast<T>(q[n + 1]) * c.fT;
    const T l10 = static_cast<T>(q[nrt]) * (T(1) - c.fT) + static_cast<T>(q[nrt + 1]) * c.fT;
    const T l11 = static_cast<T>(q[nrt + n]) * (T(1) - c.fT) + static_cast<T>(q[nrt + n + 1]) * c.fT;
    const T l0 = l00 * (T(1) - c.frho) + l01 * c.frho;
    const T l1 = l10 * (T(1) - c.frho) + l11 * c.frho;
    return exp(l0 * (T(1) - fnu) + l1 * fnu);
}

// Invariant source B_nu / nu^3 in units of 2h/c^2: 1/(exp(h nu/kT) - 1),
// zero for h nu/kT > 500 (src/opacity.cpp:209-214 divided by nu^3).  The
// scaling keeps the FP32 path clear of underflow (B_nu/nu^3 ~ 1e-47 cgs).
// hk_T = h / (k_B T) of the sample, formed once for all its bins.
__device__ __forceinline__ double vol_source(double nu, double hk_T) {
    if (nu <= 0.0 || hk_T <= 0.0) return 0.0;
    const double x = hk_T * nu;
    if (x > 500.0) return 0.0;
    return rcp(exp(x) - 1.0);
}
__device__ __forceinline__ float vol_source(float nu, float hk_T) {
    if (nu <= 0.0f || hk_T <= 0.0f) return 0.0f;
    const float x = hk_T * nu;
    if (x > 500.0f) return 0.0f;
    return __builtin_amdgcn_rcpf(expm1f(x));
}
__device__ __forceinline__ double vol_hk_over_T(double Temp) { return Temp > 0.0 ? (vcgs::h / vcgs::k_B) / Temp : 0.0; }

}  // namespace grrt_amd
