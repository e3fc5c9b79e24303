// Host-side scene construction (C++23 language, libstdc++ library).
#include "scene_host.hpp"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <random>

namespace grrt_amd {

Resolved resolve_params(const GRRTParams& p) {
    Resolved r{};
    r.W = p.width;
    r.H = p.height;
    r.M = p.mass > 0.0 ? p.mass : 1.0;
    r.r_obs = p.observer_r > 0.0 ? p.observer_r : 50.0;
    r.th_obs = p.observer_theta > 0.0 ? p.observer_theta : 1.396;
    r.ph_obs = p.observer_phi;
    r.fov = p.fov > 0.0 ? p.fov : 1.047;
    r.max_steps = p.integrator_max_steps > 0 ? p.integrator_max_steps : 10000;
    r.tol = p.integrator_tolerance > 0.0 ? p.integrator_tolerance : 1e-8;
    // Kerr with spin <= 0 falls back to 0.998; Schwarzschild is Kerr a = 0.
    r.a = 0.0;
    if (p.metric_type == GRRT_METRIC_KERR) r.a = (p.spin > 0.0 ? p.spin : 0.998) * r.M;
    r.r_h = kerr_horizon(r.M, r.a);
    r.r_isco = kerr_isco(r.M, r.a);
    r.volumetric = (p.disk_enabled && p.disk_volumetric) ? 1 : 0;
    r.disk = (p.disk_enabled && !p.disk_volumetric) ? 1 : 0;
    r.t_disk = p.disk_temperature > 0.0 ? p.disk_temperature : 1e7;
    r.r_out = p.disk_outer > 0.0 ? p.disk_outer : 20.0;
    r.r_in = p.disk_inner > 0.0 ? p.disk_inner : r.r_isco;
    r.stars = p.background_type == GRRT_BG_STARS ? 1 : 0;
    r.texture = p.background_type == GRRT_BG_TEXTURE ? 1 : 0;
    r.spp = p.samples_per_pixel > 0 ? p.samples_per_pixel : 1;
    return r;
}

double kerr_horizon(double M, double a) { return M + std::sqrt(M * M - a * a); }

double kerr_isco(double M, double a) {
    // Prograde ISCO, Bardeen, Press & Teukolsky (1972).
    const double s = a / M;
    const double z1 = 1.0 + std::cbrt(1.0 - s * s) * (std::cbrt(1.0 + s) + std::cbrt(1.0 - s));
    const double z2 = std::sqrt(3.0 * s * s + z1 * z1);
    return M * (3.0 + z2 - std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2)));
}

namespace {

void metric_lower(double M, double a, double r, double th, double g[4][4]) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) g[i][j] = 0.0;
    const double st = std::sin(th);
    const double s2 = std::max(st * st, 1e-20);
    const double ct = std::cos(th);
    const double sig = std::max(r * r + a * a * ct * ct, 1e-20);
    const double del = r * r - 2.0 * M * r + a * a;
    g[0][0] = -(1.0 - 2.0 * M * r / sig);
    g[0][3] = g[3][0] = -2.0 * M * a * r * s2 / sig;
    g[1][1] = sig / del;
    g[2][2] = sig;
    g[3][3] = (r * r + a * a + 2.0 * M * a * a * r * s2 / sig) * s2;
}

double inner(const double g[4][4], const double* u, const double* v) {
    double acc = 0.0;
    for (int m = 0; m < 4; ++m)
        for (int n = 0; n < 4; ++n) acc += g[m][n] * u[m] * v[n];
    return acc;
}

void remove_along(const double g[4][4], double* v, const double* e) {
    const double f = inner(g, v, e) / inner(g, e, e);
    for (int i = 0; i < 4; ++i) v[i] = v[i] - e[i] * f;
}

void unit(const double g[4][4], const double* v, double* out) {
    const double s = 1.0 / std::sqrt(std::abs(inner(g, v, v)));
    for (int i = 0; i < 4; ++i) out[i] = v[i] * s;
}

}  // namespace

CameraFrame build_camera(double M, double a, double r, double th, double ph) {
    (void)ph;
    CameraFrame f{};
    metric_lower(M, a, r, th, f.g);
    f.e0[0] = 1.0 / std::sqrt(-f.g[0][0]);
    // forward = -r, up = -theta, right = +phi, each made orthogonal to the
    // previous legs by the metric and normalised.
    double v[4] = {0.0, -1.0, 0.0, 0.0};
    remove_along(f.g, v, f.e0);
    unit(f.g, v, f.e3);
    double u[4] = {0.0, 0.0, -1.0, 0.0};
    remove_along(f.g, u, f.e0);
    remove_along(f.g, u, f.e3);
    unit(f.g, u, f.e2);
    double w[4] = {0.0, 0.0, 0.0, 1.0};
    remove_along(f.g, w, f.e0);
    remove_along(f.g, w, f.e3);
    remove_along(f.g, w, f.e2);
    unit(f.g, w, f.e1);
    return f;
}

double build_flux_table(double M, double a, double r_in, double r_out, std::vector<double>& lut) {
    constexpr int N = 500;
    lut.assign(N, 0.0);
    auto omega_k = [&](double r) { return std::sqrt(M / (r * r * r)); };
    auto omega = [&](double r) { const double w = omega_k(r); return w / (1.0 + a * w); };
    auto energy = [&](double r) {
        const double aw = a * omega_k(r);
        return (1.0 - 2.0 * M / r + aw) / std::sqrt(1.0 - 3.0 * M / r + 2.0 * aw);
    };
    auto ang_mom = [&](double r) {
        const double aw = a * omega_k(r);
        return std::sqrt(M * r) * (1.0 - 2.0 * aw + a * a / (r * r)) /
               std::sqrt(1.0 - 3.0 * M / r + 2.0 * aw);
    };
    const double e_in = energy(r_in), l_in = ang_mom(r_in);
    constexpr double h = 1e-6;   // central-difference step
    const double dr = (r_out - r_in) / (N - 1);
    double fmax = 0.0, integral = 0.0, prev = 0.0;
    for (int i = 1; i < N; ++i) {
        const double r = r_in + (r_out - r_in) * i / (N - 1);
        const double de = (energy(r + h) - energy(r - h)) / (2.0 * h);
        const double dlm = (ang_mom(r + h) - ang_mom(r - h)) / (2.0 * h);
        const double f = (energy(r) - e_in) * dlm - (ang_mom(r) - l_in) * de;
        integral += 0.5 * (prev + f) * dr;
        prev = f;
        const double om = omega(r);
        const double dom = (omega(r + h) - omega(r - h)) / (2.0 * h);
        const double den = energy(r) - om * ang_mom(r);
        if (std::abs(den) < 1e-20) continue;
        const double F = (3.0 * M / (8.0 * std::numbers::pi * r * r * r)) * (1.0 / den) * (-dom) * integral;
        lut[i] = std::max(F, 0.0);
        fmax = std::max(fmax, lut[i]);
    }
    return fmax;
}

namespace {
// CIE 1931 2-degree colour-matching functions, 380..780 nm in 5 nm steps,
// rounded to four decimals (the table of src/spectrum.cpp:16-44).
constexpr double kCmf[81][3] = {
    {0.0014,0.0000,0.0065}, {0.0022,0.0001,0.0105}, {0.0042,0.0001,0.0201},
    {0.0076,0.0002,0.0362}, {0.0143,0.0004,0.0679}, {0.0232,0.0006,0.1102},
    {0.0435,0.0012,0.2074}, {0.0776,0.0022,0.3713}, {0.1344,0.0040,0.6456},
    {0.2148,0.0073,1.0391}, {0.2839,0.0116,1.3856}, {0.3285,0.0168,1.6230},
    {0.3483,0.0230,1.7471}, {0.3481,0.0298,1.7826}, {0.3362,0.0380,1.7721},
    {0.3187,0.0480,1.7441}, {0.2908,0.0600,1.6692}, {0.2511,0.0739,1.5281},
    {0.1954,0.0910,1.2876}, {0.1421,0.1126,1.0419}, {0.0956,0.1390,0.8130},
    {0.0580,0.1693,0.6162}, {0.0320,0.2080,0.4652}, {0.0147,0.2586,0.3533},
    {0.0049,0.3230,0.2720}, {0.0024,0.4073,0.2123}, {0.0093,0.5030,0.1582},
    {0.0291,0.6082,0.1117}, {0.0633,0.7100,0.0782}, {0.1096,0.7932,0.0573},
    {0.1655,0.8620,0.0422}, {0.2257,0.9149,0.0298}, {0.2904,0.9540,0.0203},
    {0.3597,0.9803,0.0134}, {0.4334,0.9950,0.0087}, {0.5121,1.0002,0.0057},
    {0.5945,0.9950,0.0039}, {0.6784,0.9786,0.0027}, {0.7621,0.9520,0.0021},
    {0.8425,0.9154,0.0018}, {0.9163,0.8700,0.0017}, {0.9786,0.8163,0.0014},
    {1.0263,0.7570,0.0011}, {1.0567,0.6949,0.0010}, {1.0622,0.6310,0.0008},
    {1.0456,0.5668,0.0006}, {1.0026,0.5030,0.0003}, {0.9384,0.4412,0.0002},
    {0.8544,0.3810,0.0002}, {0.7514,0.3210,0.0001}, {0.6424,0.2650,0.0000},
    {0.5419,0.2170,0.0000}, {0.4479,0.1750,0.0000}, {0.3608,0.1382,0.0000},
    {0.2835,0.1070,0.0000}, {0.2187,0.0816,0.0000}, {0.1649,0.0610,0.0000},
    {0.1212,0.0446,0.0000}, {0.0874,0.0320,0.0000}, {0.0636,0.0232,0.0000},
    {0.0468,0.0170,0.0000}, {0.0329,0.0119,0.0000}, {0.0227,0.0082,0.0000},
    {0.0158,0.0057,0.0000}, {0.0114,0.0041,0.0000}, {0.0081,0.0029,0.0000},
    {0.0058,0.0021,0.0000}, {0.0041,0.0015,0.0000}, {0.0029,0.0010,0.0000},
    {0.0020,0.0007,0.0000}, {0.0014,0.0005,0.0000}, {0.0010,0.0004,0.0000},
    {0.0007,0.0002,0.0000}, {0.0005,0.0002,0.0000}, {0.0003,0.0001,0.0000},
    {0.0002,0.0001,0.0000}, {0.0002,0.0001,0.0000}, {0.0001,0.0000,0.0000},
    {0.0001,0.0000,0.0000}, {0.0001,0.0000,0.0000}, {0.0000,0.0000,0.0000},
};
}  // namespace

void build_chromaticity_table(std::vector<double>& lut) {
    constexpr int N = 1000;
    constexpr double t_lo = 1000.0, t_hi = 100000.0;
    constexpr double hP = 6.62607015e-34, cL = 2.99792458e8, kB = 1.380649e-23;  // SI
    lut.assign(3 * N, 0.0);
    for (int i = 0; i < N; ++i) {
        const double T = t_lo + (t_hi - t_lo) * i / (N - 1);
        double X = 0.0, Y = 0.0, Z = 0.0;
        for (int w = 0; w < 81; ++w) {
            const double lam = 380e-9 + w * 5e-9;
            const double x = hP * cL / (lam * kB * T);
            const double B = (x > 500.0) ? 0.0
                : (2.0 * hP * cL * cL) / (lam * lam * lam * lam * lam * (std::exp(x) - 1.0));
            X += B * kCmf[w][0] * 5e-9;
            Y += B * kCmf[w][1] * 5e-9;
            Z += B * kCmf[w][2] * 5e-9;
        }
        // XYZ -> linear sRGB (IEC 61966-2-1), clipped at zero, max-normalised.
        double rgb[3] = {std::max(3.2406 * X - 1.5372 * Y - 0.4986 * Z, 0.0),
                         std::max(-0.9689 * X + 1.8758 * Y + 0.0415 * Z, 0.0),
                         std::max(0.0557 * X - 0.2040 * Y + 1.0570 * Z, 0.0)};
        const double mx = std::fmax(rgb[0], std::fmax(rgb[1], rgb[2]));
        const double s = mx > 0.0 ? 1.0 / mx : 1.0;
        for (int c = 0; c < 3; ++c) lut[3 * i + c] = mx > 0.0 ? rgb[c] * s : rgb[c];
    }
}

void build_star_table(StarTable& t) {
    constexpr int NS = 5000, GT = 180, GP = 360;
    constexpr double pi = std::numbers::pi;
    std::mt19937 gen(42u);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    std::vector<double> th(NS), ph(NS), br(NS);
    std::vector<int> cell(NS);
    for (int i = 0; i < NS; ++i) {
        th[i] = std::acos(1.0 - 2.0 * u01(gen));
        ph[i] = 2.0 * pi * u01(gen) - pi;
        const double u = u01(gen);
        br[i] = std::min(0.1 * std::pow(std::max(u, 0.001), -2.5), 50.0);
        const int tb = std::clamp(static_cast<int>(th[i] * GT / pi), 0, GT - 1);
        const int pb = std::clamp(static_cast<int>((ph[i] + pi) * GP / (2.0 * pi)), 0, GP - 1);
        cell[i] = tb * GP + pb;
    }
    t.start.assign(GT * GP + 1, 0);
    for (int i = 0; i < NS; ++i) ++t.start[cell[i] + 1];
    for (int c = 0; c < GT * GP; ++c) t.start[c + 1] += t.start[c];
    std::vector<int> fill(t.start.begin(), t.start.end() - 1);
    t.tpb.assign(3 * NS, 0.0);
    for (int i = 0; i < NS; ++i) {   // catalogue order within each cell
        const int k = fill[cell[i]]++;
        t.tpb[3 * k] = th[i];
        t.tpb[3 * k + 1] = ph[i];
        t.tpb[3 * k + 2] = br[i];
    }
}

}  // namespace grrt_amd
