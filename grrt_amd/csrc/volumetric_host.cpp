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

#include "vol_column.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numbers>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>

namespace grrt_amd {

namespace {

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

double OpacityTables::lookup_es(double rho, double T) const { return bilinear(kappa_es, *this, rho, T); }
double OpacityTables::lookup_ross(double rho, double T) const { return bilinear(kappa_ross, *this, rho, T); }
double OpacityTables::lookup_mu(double rho, double T) const { return bilinear(mu, *this, rho, T); }

OpacityTables opacity_table_shape(double rho_min, double rho_max, double T_min, double T_max,
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
    return o;
}

void fill_opacity_tables_host(OpacityTables& o) {
    const size_t n2 = static_cast<size_t>(o.n_rho) * o.n_T;
    auto row = [&](int j) {
        const double rho = std::pow(10.0, o.log_rho_min + j * (o.log_rho_max - o.log_rho_min) / (o.n_rho - 1));
        for (int k = 0; k < o.n_T; ++k) {
            const double T = std::pow(10.0, o.log_T_min + k * (o.log_T_max - o.log_T_min) / (o.n_T - 1));
            const size_t i2 = static_cast<size_t>(j) * o.n_T + k;
            opacity_cell(rho, T, o.n_nu, o.log_nu_min, o.log_nu_max, &o.kappa_abs[i2], static_cast<long long>(n2),
                         o.kappa_es[i2], o.mu[i2], o.kappa_ross[i2]);
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

    // src/volumetric_disk.cpp:505-777: per column, Eddington T(tau) and a
    // hydrostatic density profile with gas + flux-limited radiation
    // pressure, extending the column top until the density falls below
    // 1e-10 of the midplane (up to 20 H), at most 8 passes.  The node steps
    // are vol_column.hpp's, shared with the GPU build (device != nullptr),
    // which gives the same tables bit for bit.
    void vertical(const hipStream_t* device) {
        const int NR = V.n_r, NZ = V.n_z;
        if (NZ < 2 || NZ > kColMaxZ) throw std::runtime_error("vertical columns: n_z out of range");
        V.z_max.assign(NR, 0.0);
        V.rho_profile.assign(static_cast<size_t>(NR) * NZ, 0.0);
        V.T_profile.assign(static_cast<size_t>(NR) * NZ, 0.0);
        std::vector<ColumnIn> in(NR);
        for (int ri = 0; ri < NR; ++ri) {
            const double r = V.r_min + (V.r_outer - V.r_min) * ri / (NR - 1);
            double Omz2 = omega_z_sq(r);
            if (r < V.r_isco || Omz2 <= 0.0) {
                Omz2 = omega_z_sq(V.r_isco);
                if (Omz2 <= 0.0) Omz2 = omega_orb(V.r_isco) * omega_orb(V.r_isco);
            }
            in[ri] = ColumnIn{V.H[ri], V.T_eff[ri], V.rho_mid[ri], Omz2};
        }
        const OpacityTables& o = V.op;
        ColumnTables t{o.kappa_ross.data(), o.kappa_es.data(), o.mu.data(), o.n_rho, o.n_T,
                       o.log_rho_min, o.log_rho_max, o.log_T_min, o.log_T_max};
        if (device) {
            columns_on_device(t, in, *device);
            return;
        }
        auto column = [&](int ri) {
#pragma clang fp contract(off)
            double rho[kColMaxZ], T[kColMaxZ], tau[kColMaxZ], E[kColMaxZ], f[kColMaxZ], mu[kColMaxZ],
                dcs2[kColMaxZ], dfE[kColMaxZ];
            RhsTerms q0[kColMaxZ], qh[kColMaxZ], q1[kColMaxZ];
            const ColumnIn& c = in[ri];
            double* rp = &V.rho_profile[static_cast<size_t>(ri) * NZ];
            double* tp = &V.T_profile[static_cast<size_t>(ri) * NZ];
            if (c.H <= 0.0 || c.Te <= 0.0 || c.rm <= 0.0) {
                V.z_max[ri] = 3.0 * c.H;
                rp[0] = 1.0;
                for (int z = 0; z < NZ; ++z) tp[z] = c.Te;
                return;
            }
            const double rho_ref = col_rho_ref(t, c, V.prm.tau_mid);
            const double Te4 = c.Te * c.Te * c.Te * c.Te;
            const double T_mid = col_T(Te4, V.prm.tau_mid);
            double zmax = 3.0 * c.H;
            for (int iter = 0; iter < 8; ++iter) {
                const double dz = zmax / (NZ - 1);
                const double step = col_tau_step(t, rho_ref, T_mid, dz);
                tau[NZ - 1] = 0.0;
                for (int z = NZ - 2; z >= 0; --z) tau[z] = tau[z + 1] + step;
                for (int z = 0; z < NZ; ++z) {
                    T[z] = col_T(Te4, tau[z]);
                    E[z] = col_E(T[z]);
                    mu[z] = col_mu(t, rho_ref, T[z]);
                }
                for (int z = 0; z < NZ; ++z) f[z] = col_f(t, E, T, z, NZ, dz, rho_ref);
                for (int z = 0; z < NZ; ++z) col_grads(T, mu, f, E, z, NZ, dz, dcs2[z], dfE[z]);
                for (int z = 0; z < NZ - 1; ++z)
                    col_step_terms(T, mu, dcs2, dfE, NZ, dz, c.Omz2, rho_ref, z, q0[z], qh[z], q1[z]);
                col_rk4(q0, qh, q1, NZ, dz, rho);
                if (!col_raise_top(rho, NZ, c.H, zmax)) break;
            }
            V.z_max[ri] = zmax;
            for (int z = 0; z < NZ; ++z) {
                rp[z] = rho[z];
                tp[z] = T[z];
            }
        };
        const int nt = std::clamp(static_cast<int>(std::thread::hardware_concurrency()), 1, 16);
        std::vector<std::thread> pool;
        for (int th = 0; th < nt; ++th)
            pool.emplace_back([&, th] {
                for (int ri = th; ri < NR; ri += nt) column(ri);
            });
        for (auto& w : pool) w.join();
    }

    // The columns on the GPU: inputs and the three (rho, T) tables up,
    // z_max and the two profiles back; stream-ordered, synchronised here.
    void columns_on_device(const ColumnTables& th, const std::vector<ColumnIn>& in, hipStream_t st) {
        const int NR = V.n_r, NZ = V.n_z;
        const size_t n2 = static_cast<size_t>(V.op.n_rho) * V.op.n_T;
        const size_t prof = static_cast<size_t>(NR) * NZ;
        const size_t bytes = sizeof(ColumnIn) * NR + sizeof(double) * (3 * n2 + NR + 2 * prof);
        char* d = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), bytes);
        if (e != hipSuccess) throw std::runtime_error(std::string("vertical columns: ") + hipGetErrorString(e));
        auto* din = reinterpret_cast<ColumnIn*>(d);
        double* dt = reinterpret_cast<double*>(d + sizeof(ColumnIn) * NR);
        double* dz = dt + 3 * n2;
        double* drho = dz + NR;
        double* dT = drho + prof;
        ColumnTables t = th;
        t.ross = dt;
        t.es = dt + n2;
        t.mu = dt + 2 * n2;
        e = hipMemcpyAsync(din, in.data(), sizeof(ColumnIn) * NR, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dt, th.ross, sizeof(double) * n2, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dt + n2, th.es, sizeof(double) * n2, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dt + 2 * n2, th.mu, sizeof(double) * n2, hipMemcpyHostToDevice, st);
        hipEvent_t ev[2] = {nullptr, nullptr};
        if (e == hipSuccess) e = hipEventCreate(&ev[0]);
        if (e == hipSuccess) e = hipEventCreate(&ev[1]);
        if (e == hipSuccess) e = hipEventRecord(ev[0], st);
        if (e == hipSuccess) e = launch_vol_columns(t, din, NR, NZ, V.prm.tau_mid, dz, drho, dT, st);
        if (e == hipSuccess) e = hipEventRecord(ev[1], st);
        if (e == hipSuccess) e = hipMemcpyAsync(V.z_max.data(), dz, sizeof(double) * NR, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(V.rho_profile.data(), drho, sizeof(double) * prof, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(V.T_profile.data(), dT, sizeof(double) * prof, hipMemcpyDeviceToHost, st);
        const hipError_t es = hipStreamSynchronize(st);
        if (e == hipSuccess) e = es;
        float ms = 0.0f;
        if (e == hipSuccess && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) V.build_ms_columns_kernel = ms;
        for (auto& x : ev)
            if (x) (void)hipEventDestroy(x);
        const hipError_t ef = hipFree(d);
        if (e == hipSuccess) e = ef;
        if (e != hipSuccess) throw std::runtime_error(std::string("vertical columns on the GPU: ") + hipGetErrorString(e));
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
                      VolTables& V, const hipStream_t* device_stream, bool device_columns) {
    const auto t0 = std::chrono::steady_clock::now();
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
    V.op = opacity_table_shape(1e-18, 1e-6, 3000.0, 1e8, p.opacity_nu_min, p.opacity_nu_max);
    const auto t1 = std::chrono::steady_clock::now();
    if (device_stream) {
        const hipError_t e = fill_opacity_tables_device(V.op, *device_stream);
        if (e != hipSuccess) throw std::runtime_error(std::string("opacity tables on the GPU: ") + hipGetErrorString(e));
    } else {
        fill_opacity_tables_host(V.op);
    }
    const auto t2 = std::chrono::steady_clock::now();
    Builder b{V};
    b.radial();
    const auto tc0 = std::chrono::steady_clock::now();
    b.vertical(device_columns ? device_stream : nullptr);
    V.build_ms_columns = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc0).count();
    b.normalize();
    const auto t3 = std::chrono::steady_clock::now();
    V.build_ms_opacity = std::chrono::duration<double, std::milli>(t2 - t1).count();
    V.build_ms_total = std::chrono::duration<double, std::milli>(t3 - t0).count();
}

}  // namespace grrt_amd
