// Host-side volumetric disk construction (C++23 language, libstdc++ library).
//
// Restates, in this library's own code, the setup algorithms of
//   src/opacity.cpp:12-303         Saha ionisation, opacities, opacity tables
//   src/volumetric_disk.cpp:33-849 radial structure, vertical profiles,
//                                  density normalisation
//   src/noise.cpp:23-36            noise permutation
// The atomic data, H- cross-section tables and physical constants are the
// reference's values (include/grrt/color/atomic_data.h:24-128,
// include/grrt/math/constants.h:11-35); they are data and must match.
#include "volumetric_host.hpp"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <numeric>
#include <random>
#include <thread>

namespace grrt_amd {

namespace cgs {
constexpr double k_B = 1.380649e-16, sigma_SB = 5.670374e-5, m_p = 1.672622e-24,
                 m_e = 9.10938e-28, c = 2.997924e10, h = 6.626070e-27, sigma_T = 6.652e-25,
                 eV = 1.602e-12, gamma_E = 1.7811, C_ff = 3.69e8, X_H = 0.70;
constexpr double a_rad = 4.0 * sigma_SB / c;
}  // namespace cgs

namespace {

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

double sigma_bf_hminus(double lambda_cm) {
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

double sigma_ff_hminus(double lambda_cm, double T) {
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

double stim_factor(double nu, double T) { return 1.0 - std::exp(-cgs::h * nu / (cgs::k_B * T)); }

double dplanck_dT(double nu, double T) {
    const double x = cgs::h * nu / (cgs::k_B * T);
    if (x > 500.0) return 0.0;
    const double ex = std::exp(x);
    const double den = (ex - 1.0) * (ex - 1.0);
    return (2.0 * cgs::h * cgs::h * nu * nu * nu * nu / (cgs::c * cgs::c * cgs::k_B * T * T)) * ex / den;
}

void log_cell(double v, double lo, double hi, int n, int& idx, double& frac) {
    double t = (v - lo) / (hi - lo) * (n - 1);
    t = std::clamp(t, 0.0, static_cast<double>(n - 2));
    idx = static_cast<int>(t);
    frac = t - idx;
}

double bilinear(const std::vector<double>& lut, const OpacityTables& o, double rho, double T) {
    int ir, iT;
    double fr, fT;
    log_cell(std::log10(rho), o.log_rho_min, o.log_rho_max, o.n_rho, ir, fr);
    log_cell(std::log10(T), o.log_T_min, o.log_T_max, o.n_T, iT, fT);
    const int n = o.n_T;
    return lut[ir * n + iT] * (1 - fr) * (1 - fT) + lut[ir * n + iT + 1] * (1 - fr) * fT +
           lut[(ir + 1) * n + iT] * fr * (1 - fT) + lut[(ir + 1) * n + iT + 1] * fr * fT;
}

double clamp_rho(double r) { return std::clamp(r, 1e-18, 1e-6); }
double clamp_T(double t) { return std::clamp(t, 3000.0, 1e8); }

}  // namespace

Ionization saha(double rho, double T) {
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

double alpha_ff_point(double nu, double T, const Ionization& ion) {
    if (ion.n_e <= 0.0 || ion.n_ion_eff <= 0.0 || nu <= 0.0 || T <= 0.0) return 0.0;
    // thermally averaged Gaunt factor (Rybicki & Lightman)
    const double arg = 4.0 * cgs::k_B * T / (cgs::gamma_E * cgs::h * nu);
    const double gff = std::max((std::sqrt(3.0) / std::numbers::pi) * std::log(std::max(arg, 1.0)), 1.0);
    return cgs::C_ff * (ion.n_e * ion.n_ion_eff / (nu * nu * nu)) * gff * std::pow(T, -0.5) *
           stim_factor(nu, T);
}

double alpha_hminus_point(double nu, double T, const Ionization& ion) {
    if (T < 1500.0 || ion.n_HI <= 0.0) return 0.0;
    const double lam = cgs::c / nu;
    const double bf = ion.n_Hminus * sigma_bf_hminus(lam) * stim_factor(nu, T);
    const double ff = ion.n_HI * ion.n_e * sigma_ff_hminus(lam, T);
    return bf + ff;
}

double alpha_bf_point(double nu, double T, const Ionization& ion) {
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

double kappa_es_point(double rho, const Ionization& ion) {
    return rho <= 0.0 ? 0.0 : cgs::sigma_T * ion.n_e / rho;
}

double kappa_abs_point(double nu, double rho, double T, const Ionization& ion) {
    const double alpha = alpha_ff_point(nu, T, ion) + alpha_hminus_point(nu, T, ion) + alpha_bf_point(nu, T, ion);
    return rho <= 0.0 ? 0.0 : alpha / rho;
}

double planck_nu_host(double nu, double T) {
    if (nu <= 0.0 || T <= 0.0) return 0.0;
    const double x = cgs::h * nu / (cgs::k_B * T);
    if (x > 500.0) return 0.0;
    return (2.0 * cgs::h * nu * nu * nu / (cgs::c * cgs::c)) / (std::exp(x) - 1.0);
}

double OpacityTables::lookup_es(double rho, double T) const { return bilinear(kappa_es, *this, rho, T); }
double OpacityTables::lookup_ross(double rho, double T) const { return bilinear(kappa_ross, *this, rho, T); }
double OpacityTables::lookup_mu(double rho, double T) const { return bilinear(mu, *this, rho, T); }

OpacityTables build_opacity_tables(double rho_min, double rho_max, double T_min, double T_max,
                                   double nu_min, double nu_max) {
    OpacityTables o;
    const double decades = std::log10(nu_max) - std::log10(nu_min);
    o.n_nu = std::max(20, static_cast<int>(decades * 10));
    o.log_nu_min = std::log10(nu_min);
    o.log_nu_max = std::log10(nu_max);
    o.log_rho_min = std::log10(rho_min);
    o.log_rho_max = std::log10(rho_max);
    o.log_T_min = std::log10(T_min);
    o.log_T_max = std::log10(T_max);
    const size_t n2 = static_cast<size_t>(o.n_rho) * o.n_T;
    o.kappa_abs.assign(n2 * o.n_nu, 0.0);
    o.kappa_es.assign(n2, 0.0);
    o.kappa_ross.assign(n2, 0.0);
    o.mu.assign(n2, 0.0);
    // Rosseland mean over 50 log-spaced frequencies in [1e13, 1e16] Hz.
    constexpr int n_ross = 50;
    const double lr0 = std::log10(1e13), lr1 = std::log10(1e16);
    auto row = [&](int j) {
        const double rho = std::pow(10.0, o.log_rho_min + j * (o.log_rho_max - o.log_rho_min) / (o.n_rho - 1));
        for (int k = 0; k < o.n_T; ++k) {
            const double T = std::pow(10.0, o.log_T_min + k * (o.log_T_max - o.log_T_min) / (o.n_T - 1));
            const Ionization ion = saha(rho, T);
            const size_t i2 = static_cast<size_t>(j) * o.n_T + k;
            o.kappa_es[i2] = kappa_es_point(rho, ion);
            o.mu[i2] = ion.mu;
            for (int i = 0; i < o.n_nu; ++i) {
                const double nu = std::pow(10.0, o.log_nu_min + i * (o.log_nu_max - o.log_nu_min) / (o.n_nu - 1));
                o.kappa_abs[static_cast<size_t>(i) * n2 + i2] = kappa_abs_point(nu, rho, T, ion);
            }
            double num = 0.0, den = 0.0;
            for (int i = 0; i < n_ross; ++i) {
                const double lnu = lr0 + i * (lr1 - lr0) / (n_ross - 1);
                const double nu = std::pow(10.0, lnu);
                const double dnu = nu * (lr1 - lr0) / (n_ross - 1) * std::log(10.0);
                const double dB = dplanck_dT(nu, T);
                const double ka = kappa_abs_point(nu, rho, T, ion);
                if (ka > 1e-30 && dB > 0.0) num += (1.0 / ka) * dB * dnu;
                den += dB * dnu;
            }
            o.kappa_ross[i2] = (num > 0.0 && den > 0.0) ? den / num : 0.0;
        }
    };
    // Rows of the (rho, T) grid are independent: spread them over host threads.
    const int nt = std::clamp(static_cast<int>(std::thread::hardware_concurrency()), 1, 16);
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int j = t; j < o.n_rho; j += nt) row(j);
        });
    for (auto& th : pool) th.join();
    return o;
}

namespace {

struct Builder {
    VolTables& V;
    double omega_orb(double r) const {
        return std::sqrt(V.M) / (r * std::sqrt(r) + V.a * std::sqrt(V.M));
    }
    double omega_z_sq(double r) const {
        const double O = omega_orb(r);
        const double sqM = std::sqrt(V.M);
        const double r3 = r * r * r;
        return O * O * (1.0 - 4.0 * V.a * sqM / std::sqrt(r3) + 3.0 * V.a * V.a / (r * r));
    }

    // Novikov–Thorne flux over [r_isco, r_outer] on the radial grid size
    // (src/volumetric_disk.cpp:279-346).
    double flux_table(std::vector<double>& f) const {
        const int N = V.n_r;
        const double M = V.M, a = V.a, ri = V.r_isco, ro = V.r_outer;
        f.assign(N, 0.0);
        auto wk = [&](double r) { return std::sqrt(M / (r * r * r)); };
        auto Om = [&](double r) { const double w = wk(r); return w / (1.0 + a * w); };
        auto E = [&](double r) {
            const double aw = a * wk(r);
            return (1.0 - 2.0 * M / r + aw) / std::sqrt(1.0 - 3.0 * M / r + 2.0 * aw);
        };
        auto L = [&](double r) {
            const double aw = a * wk(r);
            return std::sqrt(M * r) * (1.0 - 2.0 * aw + a * a / (r * r)) / std::sqrt(1.0 - 3.0 * M / r + 2.0 * aw);
        };
        const double Ei = E(ri), Li = L(ri), h = 1e-6, dr = (ro - ri) / (N - 1);
        double I = 0.0, prev = 0.0, fmax = 0.0;
        for (int i = 1; i < N; ++i) {
            const double r = ri + (ro - ri) * i / (N - 1);
            const double Ep = (E(r + h) - E(r - h)) / (2.0 * h);
            const double Lp = (L(r + h) - L(r - h)) / (2.0 * h);
            const double g = (E(r) - Ei) * Lp - (L(r) - Li) * Ep;
            I += 0.5 * (prev + g) * dr;
            prev = g;
            const double dOm = (Om(r + h) - Om(r - h)) / (2.0 * h);
            const double den = E(r) - Om(r) * L(r);
            if (std::abs(den) < 1e-20) continue;
            const double F = (3.0 * M / (8.0 * std::numbers::pi * r * r * r)) * (1.0 / den) * (-dOm) * I;
            f[i] = std::max(F, 0.0);
            fmax = std::max(fmax, f[i]);
        }
        return fmax;
    }

    // src/volumetric_disk.cpp:348-483
    void radial() {
        std::vector<double> fl;
        const double fmax = flux_table(fl);
        const int nf = static_cast<int>(fl.size());
        auto flux_at = [&](double r) {
            if (r <= V.r_isco || r >= V.r_outer || fmax <= 0.0) return 0.0;
            const double frac = (r - V.r_isco) / (V.r_outer - V.r_isco) * (nf - 1);
            const int idx = static_cast<int>(frac);
            const double t = frac - idx;
            if (idx >= nf - 1) return fl.back();
            return fl[idx] * (1.0 - t) + fl[idx + 1] * t;
        };
        const int N = V.n_r;
        V.H.assign(N, 0.0);
        V.rho_mid.assign(N, 0.0);
        V.T_eff.assign(N, 0.0);
        const double T_plunge = fmax > 0.0 ? V.t_peak * std::pow(flux_at(V.r_isco + 0.01) / fmax, 0.25) : 0.0;
        double H_isco = 0.0;
        int isco_idx = -1;
        for (int i = 0; i < N; ++i) {
            const double r = V.r_min + (V.r_outer - V.r_min) * i / (N - 1);
            double Te = 0.0;
            if (r >= V.r_isco) {
                const double F = flux_at(r);
                if (F > 0.0 && fmax > 0.0) Te = V.t_peak * std::pow(F / fmax, 0.25);
            } else {
                Te = T_plunge;
            }
            V.T_eff[i] = Te;
            if (r >= V.r_isco && isco_idx < 0) isco_idx = i;
            if (Te <= 0.0) {
                V.H[i] = H_isco > 0.0 ? H_isco : 0.01 * V.M;
                V.rho_mid[i] = 0.0;
                continue;
            }
            const double Omz2 = omega_z_sq(r);
            if (r < V.r_isco || Omz2 <= 0.0) {
                V.H[i] = H_isco > 0.0 ? H_isco : 0.01 * V.M;
                const int src = isco_idx >= 0 ? isco_idx : i - 1;
                V.rho_mid[i] = (i > 0 && V.rho_mid[src] > 0.0) ? V.rho_mid[src] : 1.0;
                continue;
            }
            const double Omz = std::sqrt(Omz2);
            double mu = V.op.lookup_mu(clamp_rho(1e-10), clamp_T(Te));
            if (mu <= 0.0 || !std::isfinite(mu)) mu = 0.6;
            const double cg2 = cgs::k_B * Te / (mu * cgs::m_p * cgs::c * cgs::c);
            double H = std::sqrt(cg2) / Omz;
            const double sqM = std::sqrt(V.M);
            const double ds = r * std::sqrt(r) + V.a * sqM;
            const double dOm = -1.5 * sqM * std::sqrt(r) / (ds * ds);
            const double shear2 = r * r * dOm * dOm;
            const double nu_visc = V.prm.alpha * (std::sqrt(cg2) * cgs::c) * H;
            const double F = r >= V.r_isco ? flux_at(r) : flux_at(V.r_isco + 0.01);
            double rm = 1.0;
            if (shear2 > 0.0 && nu_visc > 0.0 && H > 0.0 && F > 0.0)
                rm = (F / (nu_visc * shear2)) / (std::sqrt(2.0 * std::numbers::pi) * H);
            H = std::clamp(H, 0.001 * V.M, 5.0 * V.M);
            V.H[i] = H;
            V.rho_mid[i] = rm;
            if (r >= V.r_isco && (H_isco <= 0.0 || (isco_idx >= 0 && i == isco_idx))) H_isco = H;
        }
        if (H_isco > 0.0)
            for (int i = 0; i < N; ++i)
                if (V.r_min + (V.r_outer - V.r_min) * i / (N - 1) < V.r_isco) V.H[i] = H_isco;
    }

    // Levermore & Pomraning (1981) flux limiter and Eddington factor.
    static double eddington_factor(double R) {
        const double lam = (2.0 + R) / (6.0 + 3.0 * R + R * R);
        return lam + lam * lam * R * R;
    }

    // src/volumetric_disk.cpp:505-777: per column, Eddington T(tau) and a
    // hydrostatic density profile with gas + flux-limited radiation
    // pressure, extending the column top until the density falls below
    // 1e-10 of the midplane (up to 20 H), at most 8 passes.
    void vertical() {
        const int NR = V.n_r, NZ = V.n_z;
        V.z_max.assign(NR, 0.0);
        V.rho_profile.assign(static_cast<size_t>(NR) * NZ, 0.0);
        V.T_profile.assign(static_cast<size_t>(NR) * NZ, 0.0);
        const double c2 = cgs::c * cgs::c;
        auto column = [&](int ri) {
            std::vector<double> rho(NZ), T(NZ), tau(NZ), E(NZ), f(NZ), mu(NZ), dcs2(NZ), dfE(NZ);
            const double r = V.r_min + (V.r_outer - V.r_min) * ri / (NR - 1);
            const double H = V.H[ri], Te = V.T_eff[ri], rm = V.rho_mid[ri];
            double* rp = &V.rho_profile[static_cast<size_t>(ri) * NZ];
            double* tp = &V.T_profile[static_cast<size_t>(ri) * NZ];
            if (H <= 0.0 || Te <= 0.0 || rm <= 0.0) {
                V.z_max[ri] = 3.0 * H;
                rp[0] = 1.0;
                for (int z = 0; z < NZ; ++z) tp[z] = Te;
                return;
            }
            double Omz2 = omega_z_sq(r);
            if (r < V.r_isco || Omz2 <= 0.0) {
                Omz2 = omega_z_sq(V.r_isco);
                if (Omz2 <= 0.0) Omz2 = omega_orb(V.r_isco) * omega_orb(V.r_isco);
            }
            const double kref = std::max(V.op.lookup_ross(1e-10, clamp_T(Te)) + V.op.lookup_es(1e-10, clamp_T(Te)), 1.0);
            const double rho_ref = std::clamp(V.prm.tau_mid / (kref * 3.0 * H), 1e-18, 1e-6);
            const double Te4 = Te * Te * Te * Te;
            const double T_mid = std::pow(0.75 * Te4 * (V.prm.tau_mid + 2.0 / 3.0), 0.25);
            double zmax = 3.0 * H;
            const double floor_rho = 1e-15;
            std::vector<double> prev(NZ, 1.0);
            for (int iter = 0; iter < 8; ++iter) {
                const double dz = zmax / (NZ - 1);
                std::fill(rho.begin(), rho.end(), 1.0);
                std::fill(T.begin(), T.end(), T_mid);
                std::fill(tau.begin(), tau.end(), 0.0);
                for (int z = NZ - 2; z >= 0; --z) {
                    const double rh = rho[z] * rho_ref, rn = rho[z + 1] * rho_ref;
                    const double kh = V.op.lookup_ross(clamp_rho(rh), clamp_T(T[z])) + V.op.lookup_es(clamp_rho(rh), clamp_T(T[z]));
                    const double kn = V.op.lookup_ross(clamp_rho(rn), clamp_T(T[z + 1])) + V.op.lookup_es(clamp_rho(rn), clamp_T(T[z + 1]));
                    tau[z] = tau[z + 1] + 0.5 * (kh * rh + kn * rn) * dz;
                }
                for (int z = 0; z < NZ; ++z)
                    T[z] = std::pow(std::max(0.75 * Te4 * (tau[z] + 2.0 / 3.0), 0.0), 0.25);
                for (int z = 0; z < NZ; ++z) {
                    E[z] = cgs::a_rad * T[z] * T[z] * T[z] * T[z];
                    mu[z] = V.op.lookup_mu(clamp_rho(rho[z] * rho_ref), clamp_T(T[z]));
                    if (mu[z] <= 0.0 || !std::isfinite(mu[z])) mu[z] = 0.6;
                }
                for (int z = 0; z < NZ; ++z) {
                    double dE = 0.0;
                    if (z == NZ - 1) dE = (E[z] - E[z - 1]) / dz;
                    else if (z > 0) dE = (E[z + 1] - E[z - 1]) / (2.0 * dz);
                    const double rc = rho[z] * rho_ref;
                    const double den = V.op.lookup_ross(clamp_rho(rc), clamp_T(T[z])) * rc * E[z];
                    f[z] = eddington_factor(den < 1e-30 ? 1e30 : std::abs(dE) / den);
                }
                auto cs2 = [&](int z) { return cgs::k_B * T[z] / (mu[z] * cgs::m_p); };
                for (int z = 0; z < NZ; ++z) {
                    if (z == 0) { dcs2[z] = 0.0; dfE[z] = 0.0; continue; }
                    if (z == NZ - 1) {
                        dcs2[z] = (cs2(z) - cs2(z - 1)) / dz;
                        dfE[z] = (f[z] * E[z] - f[z - 1] * E[z - 1]) / dz;
                    } else {
                        dcs2[z] = (cs2(z + 1) - cs2(z - 1)) / (2.0 * dz);
                        dfE[z] = (f[z + 1] * E[z + 1] - f[z - 1] * E[z - 1]) / (2.0 * dz);
                    }
                }
                // dRho/dz = [-rho Omz^2 z - rho d(cs^2)/dz - d(f E)/dz / rho_ref] / cs^2,
                // geometric z with CGS thermodynamics converted by c^2.
                auto rhs = [&](double ze, double re) {
                    const double zf = ze / dz;
                    const int idx = std::clamp(static_cast<int>(zf), 0, NZ - 2);
                    const double t = zf - idx;
                    const double cs = cgs::k_B * ((1.0 - t) * T[idx] + t * T[idx + 1]) /
                                      (((1.0 - t) * mu[idx] + t * mu[idx + 1]) * cgs::m_p);
                    const double dc = (1.0 - t) * dcs2[idx] + t * dcs2[idx + 1];
                    const double df = (1.0 - t) * dfE[idx] + t * dfE[idx + 1];
                    if (cs < 1e-30) return 0.0;
                    return (-re * (Omz2 * ze) - re * (dc / c2) - df / (rho_ref * c2)) / std::max(cs / c2, 1e-30);
                };
                rho[0] = 1.0;
                for (int z = 0; z < NZ - 1; ++z) {
                    const double zh = z * dz, rh = rho[z];
                    const double k1 = dz * rhs(zh, rh);
                    const double k2 = dz * rhs(zh + 0.5 * dz, std::max(rh + 0.5 * k1, floor_rho));
                    const double k3 = dz * rhs(zh + 0.5 * dz, std::max(rh + 0.5 * k2, floor_rho));
                    const double k4 = dz * rhs(zh + dz, std::max(rh + k3, floor_rho));
                    rho[z + 1] = std::max(rh + (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0, floor_rho);
                }
                if (rho[NZ - 1] > 1e-10 && zmax < 20.0 * H) {
                    zmax = std::min(zmax + H, 20.0 * H);
                    prev = rho;
                    continue;
                }
                double dmax = 0.0;
                for (int z = 0; z < NZ; ++z)
                    if (prev[z] > floor_rho * 10.0) dmax = std::max(dmax, std::abs(rho[z] - prev[z]) / prev[z]);
                prev = rho;
                if (iter > 0 && dmax < 0.001) break;
            }
            V.z_max[ri] = zmax;
            for (int z = 0; z < NZ; ++z) {
                rp[z] = rho[z];
                tp[z] = T[z];
            }
        };
        const int nt = std::clamp(static_cast<int>(std::thread::hardware_concurrency()), 1, 16);
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                for (int ri = t; ri < NR; ri += nt) column(ri);
            });
        for (auto& th : pool) th.join();
    }

    // src/volumetric_disk.cpp:783-849
    void normalize() {
        int peak = 0;
        double peak_rho = 0.0;
        for (int i = 0; i < V.n_r; ++i) {
            const double r = V.r_min + (V.r_outer - V.r_min) * i / (V.n_r - 1);
            if (r >= V.r_isco && V.rho_mid[i] > peak_rho) {
                peak_rho = V.rho_mid[i];
                peak = i;
            }
        }
        if (peak_rho <= 0.0) {
            V.rho_scale = 1.0;
            return;
        }
        V.noise_scale = V.prm.noise_scale > 0.0 ? V.prm.noise_scale : 2.0 * V.H[peak];
        if (V.noise_scale < 0.01) V.noise_scale = 0.01;
        const double dz = V.z_max[peak] / (V.n_z - 1);
        const double* col = &V.rho_profile[static_cast<size_t>(peak) * V.n_z];
        double integral = 0.0;
        for (int z = 0; z < V.n_z - 1; ++z) integral += 0.5 * (col[z] + col[z + 1]) * dz;
        integral *= 2.0;
        const double Tp = V.T_eff[peak];
        double guess = 1e-10;
        for (int it = 0; it < 3; ++it) {
            const double kap = V.op.lookup_ross(clamp_rho(guess), clamp_T(Tp)) + V.op.lookup_es(clamp_rho(guess), clamp_T(Tp));
            if (kap <= 0.0 || integral <= 0.0) {
                V.rho_scale = 1.0;
                return;
            }
            V.rho_scale = V.prm.tau_mid / (kap * peak_rho * integral);
            guess = clamp_rho(V.rho_scale * peak_rho);
        }
    }
};

}  // namespace

void build_volumetric(double M, double a, double r_outer, double t_peak, const VolParams& p,
                      VolTables& V) {
    V.M = M;
    V.a = a;
    V.r_outer = r_outer;
    V.t_peak = t_peak;
    V.prm = p;
    const double as = a / M;
    const double z1 = 1.0 + std::cbrt(1.0 - as * as) * (std::cbrt(1.0 + as) + std::cbrt(1.0 - as));
    const double z2 = std::sqrt(3.0 * as * as + z1 * z1);
    V.r_isco = M * (3.0 + z2 - std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2)));
    V.r_horizon = M + std::sqrt(M * M - a * a);
    V.r_min = V.r_horizon + 0.01 * M;
    V.taper_width = (V.r_isco - V.r_horizon) / 3.0;
    {
        const double v = std::sqrt(M / V.r_isco), v3 = v * v * v;
        const double den = std::sqrt(1.0 - 3.0 * v * v + 2.0 * as * v3);
        V.E_isco = (1.0 - 2.0 * v * v + as * v3) / den;
        V.L_isco = std::sqrt(M * V.r_isco) * (1.0 - 2.0 * as * v3 + as * as * v * v * v * v) / den;
    }
    // Noise permutation: Fisher–Yates over 0..255 with mt19937(seed) and
    // uniform_int_distribution, duplicated to 512 entries.
    {
        int base[256];
        std::iota(base, base + 256, 0);
        std::mt19937 rng(p.seed);
        for (int i = 255; i > 0; --i) {
            std::uniform_int_distribution<int> d(0, i);
            std::swap(base[i], base[d(rng)]);
        }
        for (int i = 0; i < 256; ++i) V.perm[i] = V.perm[i + 256] = base[i];
    }
    V.op = build_opacity_tables(1e-18, 1e-6, 3000.0, 1e8, p.opacity_nu_min, p.opacity_nu_max);
    Builder b{V};
    b.radial();
    b.vertical();
    b.normalize();
}

}  // namespace grrt_amd
