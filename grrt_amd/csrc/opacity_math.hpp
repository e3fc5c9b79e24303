// Point opacities of the volumetric disk's gas, shared by the host table
// builder and the device one (src/opacity.cpp:12-234): Saha ionisation of the
// H/He/C/O/Fe mixture, free-free, H- bound-free and free-free, and Kramers
// bound-free absorption, electron scattering, Planck functions.  The atomic
// data, H- cross-section tables and physical constants are the reference's
// values (include/grrt/color/atomic_data.h:24-128,
// include/grrt/math/constants.h:11-35); they are data and must match.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numbers>

namespace grrt_amd {

// Ionisation state of the H/He/C/O/Fe mixture (Saha equation).
struct Ionization {
    double n_e, mu, n_ion_eff, n_HI, n_Hminus;
    double n_pop[5][5];
};

namespace cgs {
constexpr double k_B = 1.380649e-16, sigma_SB = 5.670374e-5, m_p = 1.672622e-24,
                 m_e = 9.10938e-28, c = 2.997924e10, h = 6.626070e-27, sigma_T = 6.652e-25,
                 eV = 1.602e-12, gamma_E = 1.7811, C_ff = 3.69e8, X_H = 0.70;
constexpr double a_rad = 4.0 * sigma_SB / c;
}  // namespace cgs


struct Element {
    double x, A;
    int stages;
    double chi_eV[5], g0[5], Z_eff[5];
    int n_outer[5];
};

// H, He, C, O, Fe: mass fraction, atomic mass, stages, ionisation energies,
// ground-state weights, Kramers effective charges, outer shells.
constexpr Element kElements[5] = {
    {0.70, 1.0, 2, {13.60, 0, 0, 0, 0}, {2.0, 1.0, 0, 0, 0}, {1.0, 0, 0, 0, 0}, {1, 0, 0, 0, 0}},
    {0.28, 4.0, 3, {24.59, 54.42, 0, 0, 0}, {1.0, 2.0, 1.0, 0, 0}, {1.0, 2.0, 0, 0, 0}, {1, 1, 0, 0, 0}},
    {3.0e-3, 12.0, 5, {11.26, 24.38, 47.89, 64.49, 0}, {9.0, 6.0, 1.0, 6.0, 9.0},
     {1.0, 1.5, 2.0, 2.5, 0}, {2, 2, 2, 2, 0}},
    {6.6e-3, 16.0, 5, {13.62, 35.12, 54.93, 77.41, 0}, {9.0, 4.0, 9.0, 4.0, 1.0},
     {1.0, 1.5, 2.0, 2.5, 0}, {2, 2, 2, 2, 0}},
    {1.2e-3, 56.0, 5, {7.90, 16.19, 30.65, 54.80, 0}, {25.0, 30.0, 25.0, 28.0, 25.0},
     {1.0, 1.5, 2.0, 2.5, 0}, {4, 4, 3, 3, 0}},
};
constexpr double kChiHminus = 0.754 * 1.602e-12;   // H- electron affinity [erg]

// H- bound-free cross-section (Wishart 1979): wavelength [nm], sigma [cm^2].
constexpr double kBfLam[16] = {200, 300, 400, 450, 500, 550, 600, 650,
                               700, 750, 800, 850, 900, 1000, 1200, 1642};
constexpr double kBfSig[16] = {0.36e-17, 1.06e-17, 1.63e-17, 2.33e-17, 3.02e-17, 3.58e-17,
                               3.96e-17, 4.14e-17, 4.09e-17, 3.98e-17, 3.93e-17, 3.98e-17,
                               3.85e-17, 3.40e-17, 2.23e-17, 0.0};
// H- free-free cross-section (Bell & Berrington 1987) [cm^5], 8 T x 7 lambda.
constexpr double kFfT[8] = {2000, 3000, 4000, 5000, 6000, 7000, 8000, 10000};
constexpr double kFfLam[7] = {400, 500, 600, 700, 800, 900, 1000};
constexpr double kFfSig[56] = {
    2.21e-39, 3.44e-39, 4.91e-39, 6.61e-39, 8.55e-39, 1.07e-38, 1.31e-38,
    3.30e-39, 5.15e-39, 7.35e-39, 9.90e-39, 1.28e-38, 1.60e-38, 1.96e-38,
    4.39e-39, 6.86e-39, 9.79e-39, 1.32e-38, 1.70e-38, 2.13e-38, 2.61e-38,
    5.48e-39, 8.56e-39, 1.22e-38, 1.65e-38, 2.13e-38, 2.67e-38, 3.27e-38,
    6.57e-39, 1.03e-38, 1.47e-38, 1.98e-38, 2.56e-38, 3.20e-38, 3.92e-38,
    7.66e-39, 1.20e-38, 1.71e-38, 2.30e-38, 2.98e-38, 3.73e-38, 4.57e-38,
    8.76e-39, 1.37e-38, 1.95e-38, 2.63e-38, 3.41e-38, 4.27e-38, 5.22e-38,
    1.09e-38, 1.71e-38, 2.44e-38, 3.29e-38, 4.26e-38, 5.33e-38, 6.53e-38,
};

__host__ __device__ inline double sigma_bf_hminus(double lambda_cm) {
    const double nm = lambda_cm * 1e7;
    if (nm >= kBfLam[15] || nm <= 0.0) return 0.0;
    if (nm <= kBfLam[0]) return kBfSig[0];
    for (int i = 0; i < 15; ++i)
        if (nm >= kBfLam[i] && nm <= kBfLam[i + 1]) {
            const double f = (nm - kBfLam[i]) / (kBfLam[i + 1] - kBfLam[i]);
            return kBfSig[i] + f * (kBfSig[i + 1] - kBfSig[i]);
        }
    return 0.0;
}

__host__ __device__ inline double sigma_ff_hminus(double lambda_cm, double T) {
    const double lam = std::clamp(lambda_cm * 1e7, kFfLam[0], kFfLam[6]);
    const double tc = std::clamp(T, kFfT[0], kFfT[7]);
    int il = 0, it = 0;
    for (int i = 0; i < 6; ++i)
        if (lam >= kFfLam[i]) il = i;
    for (int i = 0; i < 7; ++i)
        if (tc >= kFfT[i]) it = i;
    const int il1 = std::min(il + 1, 6), it1 = std::min(it + 1, 7);
    const double fl = (il == il1) ? 0.0 : (lam - kFfLam[il]) / (kFfLam[il1] - kFfLam[il]);
    const double ft = (it == it1) ? 0.0 : (tc - kFfT[it]) / (kFfT[it1] - kFfT[it]);
    auto v = [](int ti, int li) { return kFfSig[ti * 7 + li]; };
    return v(it, il) * (1 - fl) * (1 - ft) + v(it, il1) * fl * (1 - ft) + v(it1, il) * (1 - fl) * ft +
           v(it1, il1) * fl * ft;
}

__host__ __device__ inline double stim_factor(double nu, double T) { return 1.0 - std::exp(-cgs::h * nu / (cgs::k_B * T)); }

__host__ __device__ inline double dplanck_dT(double nu, double T) {
    const double x = cgs::h * nu / (cgs::k_B * T);
    if (x > 500.0) return 0.0;
    const double ex = std::exp(x);
    const double den = (ex - 1.0) * (ex - 1.0);
    return (2.0 * cgs::h * cgs::h * nu * nu * nu * nu / (cgs::c * cgs::c * cgs::k_B * T * T)) * ex / den;
}


__host__ __device__ inline Ionization saha(double rho, double T) {
    Ionization s{};
    if (T < 100.0 || rho <= 0.0) {
        s.mu = 1.3;
        for (int e = 0; e < 5; ++e) {
            s.n_pop[e][0] = rho * kElements[e].x / (kElements[e].A * cgs::m_p);
            if (e == 0) s.n_HI = s.n_pop[e][0];
        }
        return s;
    }
    const double ldb = cgs::h / std::sqrt(2.0 * std::numbers::pi * cgs::m_e * cgs::k_B * T);
    const double ldb3 = ldb * ldb * ldb;
    const double pref = 2.0 / ldb3;
    double n_tot[5];
    for (int e = 0; e < 5; ++e) n_tot[e] = rho * kElements[e].x / (kElements[e].A * cgs::m_p);
    double n_e = rho * (1.0 + cgs::X_H) / (2.0 * cgs::m_p);
    for (int iter = 0; iter < 50; ++iter) {
        double n_e_new = 0.0;
        for (int e = 0; e < 5; ++e) {
            const Element& el = kElements[e];
            double ratio[5] = {};
            for (int i = 0; i < el.stages - 1; ++i)
                ratio[i] = (pref / n_e) * (el.g0[i + 1] / el.g0[i]) *
                           std::exp(-(el.chi_eV[i] * cgs::eV) / (cgs::k_B * T));
            double denom = 1.0, prod = 1.0;
            for (int i = 0; i < el.stages - 1; ++i) {
                prod *= ratio[i];
                denom += prod;
            }
            const double f0 = 1.0 / denom;
            s.n_pop[e][0] = f0 * n_tot[e];
            prod = f0;
            for (int i = 0; i < el.stages - 1; ++i) {
                prod *= ratio[i];
                s.n_pop[e][i + 1] = prod * n_tot[e];
            }
            for (int i = 1; i < el.stages; ++i) n_e_new += i * s.n_pop[e][i];
        }
        const double rel = std::abs(n_e_new - n_e) / std::max(n_e, 1e-30);
        n_e = n_e_new;
        if (rel < 1e-6) break;
    }
    s.n_e = n_e;
    s.n_HI = s.n_pop[0][0];
    s.n_ion_eff = 0.0;
    double n_part = n_e;
    for (int e = 0; e < 5; ++e) {
        for (int i = 1; i < kElements[e].stages; ++i) s.n_ion_eff += static_cast<double>(i * i) * s.n_pop[e][i];
        for (int i = 0; i < kElements[e].stages; ++i) n_part += s.n_pop[e][i];
    }
    s.mu = rho / (n_part * cgs::m_p);
    s.n_Hminus = s.n_HI * (n_e * ldb3 * 0.25 * std::exp(kChiHminus / (cgs::k_B * T)));
    return s;
}

__host__ __device__ inline double alpha_ff_point(double nu, double T, const Ionization& ion) {
    if (ion.n_e <= 0.0 || ion.n_ion_eff <= 0.0 || nu <= 0.0 || T <= 0.0) return 0.0;
    // thermally averaged Gaunt factor (Rybicki & Lightman)
    const double arg = 4.0 * cgs::k_B * T / (cgs::gamma_E * cgs::h * nu);
    const double gff = std::max((std::sqrt(3.0) / std::numbers::pi) * std::log(std::max(arg, 1.0)), 1.0);
    return cgs::C_ff * (ion.n_e * ion.n_ion_eff / (nu * nu * nu)) * gff * std::pow(T, -0.5) *
           stim_factor(nu, T);
}

__host__ __device__ inline double alpha_hminus_point(double nu, double T, const Ionization& ion) {
    if (T < 1500.0 || ion.n_HI <= 0.0) return 0.0;
    const double lam = cgs::c / nu;
    const double bf = ion.n_Hminus * sigma_bf_hminus(lam) * stim_factor(nu, T);
    const double ff = ion.n_HI * ion.n_e * sigma_ff_hminus(lam, T);
    return bf + ff;
}

__host__ __device__ inline double alpha_bf_point(double nu, double T, const Ionization& ion) {
    const double stim = stim_factor(nu, T);
    double total = 0.0;
    for (int e = 0; e < 5; ++e) {
        const Element& el = kElements[e];
        for (int i = 0; i < el.stages - 1; ++i) {
            if (el.Z_eff[i] <= 0.0 || el.n_outer[i] <= 0) continue;
            const double nu_th = el.chi_eV[i] * cgs::eV / cgs::h;
            if (nu < nu_th) continue;
            const double n = static_cast<double>(el.n_outer[i]);
            const double Z = el.Z_eff[i];
            const double sigma = 7.91e-18 * n / (Z * Z) * std::pow(nu_th / nu, 3.0);   // Kramers
            total += ion.n_pop[e][i] * sigma * stim;
        }
    }
    return total;
}

__host__ __device__ inline double kappa_es_point(double rho, const Ionization& ion) {
    return rho <= 0.0 ? 0.0 : cgs::sigma_T * ion.n_e / rho;
}

__host__ __device__ inline double kappa_abs_point(double nu, double rho, double T, const Ionization& ion) {
    const double alpha = alpha_ff_point(nu, T, ion) + alpha_hminus_point(nu, T, ion) + alpha_bf_point(nu, T, ion);
    return rho <= 0.0 ? 0.0 : alpha / rho;
}

__host__ __device__ inline double planck_nu_host(double nu, double T) {
    if (nu <= 0.0 || T <= 0.0) return 0.0;
    const double x = cgs::h * nu / (cgs::k_B * T);
    if (x > 500.0) return 0.0;
    return (2.0 * cgs::h * nu * nu * nu / (cgs::c * cgs::c)) / (std::exp(x) - 1.0);
}

// One cell of the opacity tables at (rho, T): electron scattering, mean
// molecular weight and Rosseland mean (50 log-spaced frequencies over
// 1e13..1e16 Hz, src/opacity.cpp:272-297) returned, and the absorption
// opacity at the table's n_nu log-spaced frequencies over
// [10^lnu0, 10^lnu1] written to kabs[i * stride].
__host__ __device__ inline void opacity_cell(double rho, double T, int n_nu, double lnu0, double lnu1,
                                             double* kabs, long long stride, double& es, double& mu,
                                             double& ross) {
    const Ionization ion = saha(rho, T);
    es = kappa_es_point(rho, ion);
    mu = ion.mu;
    for (int i = 0; i < n_nu; ++i) {
        const double nu = std::pow(10.0, lnu0 + i * (lnu1 - lnu0) / (n_nu - 1));
        kabs[i * stride] = kappa_abs_point(nu, rho, T, ion);
    }
    constexpr int n_ross = 50;
    const double lr0 = 13.0, lr1 = 16.0;
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n_ross; ++i) {
        const double lnu = lr0 + i * (lr1 - lr0) / (n_ross - 1);
        const double nu = std::pow(10.0, lnu);
        const double dnu = nu * (lr1 - lr0) / (n_ross - 1) * std::numbers::ln10;
        const double dB = dplanck_dT(nu, T);
        const double ka = kappa_abs_point(nu, rho, T, ion);
        if (ka > 1e-30 && dB > 0.0) num += (1.0 / ka) * dB * dnu;
        den += dB * dnu;
    }
    ross = (num > 0.0 && den > 0.0) ? den / num : 0.0;
}

}  // namespace grrt_amd
