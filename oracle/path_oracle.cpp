// CPU oracle: restatement of the reference per-pixel geodesic path (thin
// disk, volumetric disk and spectral ray-march).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg as the checker; never by the product library.
//
// Written to follow the reference's arithmetic literally (operation order as
// the C++ expressions parse), compiled strict IEEE (-ffp-contract=off) for the
// canonical oracle and with the reference's release flags (-O3 -ffast-math
// -march=native) for the self-noise variant.  Citations are to the reference
// tree (SethSamuelCode/GRRT).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------- sincos ---
// include/grrt/math/fast_trig.h:21-82 — Cephes sin/cos polynomials (public
// domain, S. Moshier) with three-term Cody–Waite reduction by pi/4.
inline void fast_sincos(double x, double& s, double& c) {
    const double DP1 = 7.85398125648498535156e-1;
    const double DP2 = 3.77489470793079817668e-8;
    const double DP3 = 2.69515142907905952645e-15;
    const double FOPI = 1.27323954473516268615;
    const double S0 = -1.66666666666666324348e-01, S1 = 8.33333333332248946124e-03,
                 S2 = -1.98412698298579493134e-04, S3 = 2.75573137070700676789e-06,
                 S4 = -2.50507602534068634195e-08, S5 = 1.58969099521155010221e-10;
    const double C0 = 4.16666666666666019037e-02, C1 = -1.38888888888741095749e-03,
                 C2 = 2.48015872894767294178e-05, C3 = -2.75573143513906633035e-07,
                 C4 = 2.08757232129817482790e-09, C5 = -1.13596475577881948265e-11;
    double ax = std::fabs(x);
    int ss = x < 0.0 ? -1 : 1;
    int sc = 1;
    int j = static_cast<int>(ax * FOPI);
    j = (j + 1) & ~1;
    const double y = static_cast<double>(j);
    ax = ((ax - y * DP1) - y * DP2) - y * DP3;
    int oct = j & 7;
    if (oct > 3) { ss = -ss; sc = -sc; oct -= 4; }
    if (oct > 1) sc = -sc;
    const double z = ax * ax;
    const double ps = ((((S5 * z + S4) * z + S3) * z + S2) * z + S1) * z + S0;
    const double pc = (((((C5 * z + C4) * z + C3) * z + C2) * z + C1) * z + C0);
    const double sr = ax + ax * z * ps;
    const double cr = 1.0 - 0.5 * z + z * z * pc;
    if (oct == 1 || oct == 2) { s = ss * cr; c = sc * sr; }
    else                      { s = ss * sr; c = sc * cr; }
}

// std::min/max/clamp semantics (NaN in the second operand is ignored).
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }
inline double sclamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// ------------------------------------------------------------- Kerr (BL) ---
struct State { double x[4]; double p[4]; };

// include/grrt/spacetime/kerr.h:31-108 — fused dx^mu = g^{mu nu} p_nu and
// analytic dp_r, dp_theta; dp_t = dp_phi = 0.
inline void kerr_derivs(double M, double a, const State& s, State& d) {
    const double r = s.x[1];
    const double theta = s.x[2];
    const double r2 = r * r;
    const double a2 = a * a;
    double sin_t, cos_t;
    fast_sincos(theta, sin_t, cos_t);
    if (std::fabs(sin_t) < 1e-10) sin_t = (sin_t >= 0.0) ? 1e-10 : -1e-10;
    const double sin2 = sin_t * sin_t;
    const double cos2 = cos_t * cos_t;
    const double sin4 = sin2 * sin2;
    const double sin2theta = 2.0 * sin_t * cos_t;
    const double Sigma = r2 + a2 * cos2;
    const double Delta = r2 - 2.0 * M * r + a2;
    const double Sigma2 = Sigma * Sigma;
    const double Delta2 = Delta * Delta;
    const double Mr = M * r;
    const double g_tt = -(1.0 - 2.0 * Mr / Sigma);
    const double g_tphi = -2.0 * M * a * r * sin2 / Sigma;
    const double g_rr = Sigma / Delta;
    const double g_thth = Sigma;
    const double g_phph = (r2 + a2 + 2.0 * M * a2 * r * sin2 / Sigma) * sin2;
    const double det_tf = g_tt * g_phph - g_tphi * g_tphi;
    const double gi_tt = g_phph / det_tf;
    const double gi_tphi = -g_tphi / det_tf;
    const double gi_rr = 1.0 / g_rr;
    const double gi_thth = 1.0 / g_thth;
    const double gi_phph = g_tt / det_tf;
    const double* p = s.p;
    d.x[0] = gi_tt * p[0] + gi_tphi * p[3];
    d.x[1] = gi_rr * p[1];
    d.x[2] = gi_thth * p[2];
    d.x[3] = gi_tphi * p[0] + gi_phph * p[3];
    const double Sm2r2 = Sigma - 2.0 * r2;
    const double dgtt_r = 2.0 * M * Sm2r2 / Sigma2;
    const double dgtp_r = -2.0 * M * a * sin2 * Sm2r2 / Sigma2;
    const double dgrr_r = (2.0 * r * Delta - 2.0 * Sigma * (r - M)) / Delta2;
    const double dgpp_r = 2.0 * r * sin2 + 2.0 * M * a2 * sin4 * Sm2r2 / Sigma2;
    const double* dx = d.x;
    const double dp_r = 0.5 * (dgtt_r * dx[0] * dx[0] + 2.0 * dgtp_r * dx[0] * dx[3]
                               + dgrr_r * dx[1] * dx[1] + 2.0 * r * dx[2] * dx[2]
                               + dgpp_r * dx[3] * dx[3]);
    const double dgtt_t = 2.0 * Mr * a2 * sin2theta / Sigma2;
    const double dgtp_t = -2.0 * Mr * a * sin2theta * (Sigma + a2 * sin2) / Sigma2;
    const double dgrr_t = -a2 * sin2theta / Delta;
    const double dgpp_t = (r2 + a2) * sin2theta
        + 2.0 * Mr * a2 * sin2theta * sin2 * (2.0 * Sigma + a2 * sin2) / Sigma2;
    const double dp_t = 0.5 * (dgtt_t * dx[0] * dx[0] + 2.0 * dgtp_t * dx[0] * dx[3]
                               + dgrr_t * dx[1] * dx[1] + (-a2 * sin2theta) * dx[2] * dx[2]
                               + dgpp_t * dx[3] * dx[3]);
    d.p[0] = 0.0;
    d.p[1] = dp_r;
    d.p[2] = dp_t;
    d.p[3] = 0.0;
#ifdef ORACLE_RHS_ULP_NOISE
    // Rounding-noise variant (lib/liboracle_ulp.so): every evolving
    // component of the right-hand side moved by one ulp up or down (a hash of
    // its bits picks the direction).  Compared with the strict build it
    // measures how far rounding-level differences in the geodesic carry into
    // each output, the floor for any other arithmetic of the same path.
    auto ulp_step = [](double v) {
        uint64_t b;
        std::memcpy(&b, &v, sizeof b);
        uint64_t h = b * 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
        return (h & 1) ? std::nextafter(v, 1e300) : std::nextafter(v, -1e300);
    };
    d.x[1] = ulp_step(d.x[1]);
    d.x[2] = ulp_step(d.x[2]);
    d.x[3] = ulp_step(d.x[3]);
    d.p[1] = ulp_step(d.p[1]);
    d.p[2] = ulp_step(d.p[2]);
#endif
}

// src/kerr.cpp:19-40 — covariant metric with libm sin/cos and pole guards.
void g_lower(double M, double a, const double* x, double g[4][4]) {
    std::memset(g, 0, sizeof(double) * 16);
    const double r = x[1], theta = x[2];
    const double sin_t = std::sin(theta);
    double sin2 = smax(sin_t * sin_t, 1e-20);
    const double cos_t = std::cos(theta);
    const double S = smax(r * r + a * a * cos_t * cos_t, 1e-20);
    const double D = r * r - 2.0 * M * r + a * a;
    g[0][0] = -(1.0 - 2.0 * M * r / S);
    g[0][3] = -2.0 * M * a * r * sin2 / S;
    g[3][0] = g[0][3];
    g[1][1] = S / D;
    g[2][2] = S;
    g[3][3] = (r * r + a * a + 2.0 * M * a * a * r * sin2 / S) * sin2;
}

// include/grrt/math/matrix4.h:49-64 — block inverse of the BL metric.
void g_upper(double M, double a, const double* x, double gi[4][4]) {
    double g[4][4];
    g_lower(M, a, x, g);
    std::memset(gi, 0, sizeof(double) * 16);
    const double det = g[0][0] * g[3][3] - g[0][3] * g[3][0];
    gi[0][0] = g[3][3] / det;
    gi[3][3] = g[0][0] / det;
    gi[0][3] = -g[0][3] / det;
    gi[3][0] = -g[3][0] / det;
    gi[1][1] = 1.0 / g[1][1];
    gi[2][2] = 1.0 / g[2][2];
}

double hamiltonian(double M, double a, const double* x, const double* p) {
    double gi[4][4];
    g_upper(M, a, x, gi);
    double h = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) h += gi[i][j] * p[i] * p[j];
    return 0.5 * h;
}

double horizon(double M, double a) { return M + std::sqrt(M * M - a * a); }  // src/kerr.cpp:46-50

double isco(double M, double a) {  // src/kerr.cpp:52-63 (Bardeen, Press & Teukolsky 1972)
    const double as = a / M;
    const double Z1 = 1.0 + std::cbrt(1.0 - as * as) * (std::cbrt(1.0 + as) + std::cbrt(1.0 - as));
    const double Z2 = std::sqrt(3.0 * as * as + Z1 * Z1);
    return M * (3.0 + Z2 - std::sqrt((3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2)));
}

// ------------------------------------------------------------ integrators ---
// include/grrt/geodesic/rk4.h:73-177 — Dormand–Prince 5(4) with local
// extrapolation; k7 evaluated at y5; relative max-norm error estimate.
void dp45_attempt(double M, double a, const State& y, double dl, State& y5, double& err) {
    State k1, k2, k3, k4, k5, k6, k7, s;
    kerr_derivs(M, a, y, k1);
    const double h21 = dl / 5.0;
    for (int i = 0; i < 4; ++i) {
        s.x[i] = y.x[i] + k1.x[i] * h21;
        s.p[i] = y.p[i] + k1.p[i] * h21;
    }
    kerr_derivs(M, a, s, k2);
    const double h31 = dl * 3.0 / 40.0, h32 = dl * 9.0 / 40.0;
    for (int i = 0; i < 4; ++i) {
        s.x[i] = y.x[i] + k1.x[i] * h31 + k2.x[i] * h32;
        s.p[i] = y.p[i] + k1.p[i] * h31 + k2.p[i] * h32;
    }
    kerr_derivs(M, a, s, k3);
    const double h41 = dl * 44.0 / 45.0, h42 = dl * -56.0 / 15.0, h43 = dl * 32.0 / 9.0;
    for (int i = 0; i < 4; ++i) {
        s.x[i] = y.x[i] + k1.x[i] * h41 + k2.x[i] * h42 + k3.x[i] * h43;
        s.p[i] = y.p[i] + k1.p[i] * h41 + k2.p[i] * h42 + k3.p[i] * h43;
    }
    kerr_derivs(M, a, s, k4);
    const double h51 = dl * 19372.0 / 6561.0, h52 = dl * -25360.0 / 2187.0,
                 h53 = dl * 64448.0 / 6561.0, h54 = dl * -212.0 / 729.0;
    for (int i = 0; i < 4; ++i) {
        s.x[i] = y.x[i] + k1.x[i] * h51 + k2.x[i] * h52 + k3.x[i] * h53 + k4.x[i] * h54;
        s.p[i] = y.p[i] + k1.p[i] * h51 + k2.p[i] * h52 + k3.p[i] * h53 + k4.p[i] * h54;
    }
    kerr_derivs(M, a, s, k5);
    const double h61 = dl * 9017.0 / 3168.0, h62 = dl * -355.0 / 33.0,
                 h63 = dl * 46732.0 / 5247.0, h64 = dl * 49.0 / 176.0,
                 h65 = dl * -5103.0 / 18656.0;
    for (int i = 0; i < 4; ++i) {
        s.x[i] = y.x[i] + k1.x[i] * h61 + k2.x[i] * h62 + k3.x[i] * h63 + k4.x[i] * h64
                 + k5.x[i] * h65;
        s.p[i] = y.p[i] + k1.p[i] * h61 + k2.p[i] * h62 + k3.p[i] * h63 + k4.p[i] * h64
                 + k5.p[i] * h65;
    }
    kerr_derivs(M, a, s, k6);
    const double b1 = dl * 35.0 / 384.0, b3 = dl * 500.0 / 1113.0, b4 = dl * 125.0 / 192.0,
                 b5 = dl * -2187.0 / 6784.0, b6 = dl * 11.0 / 84.0;
    for (int i = 0; i < 4; ++i) {
        y5.x[i] = y.x[i] + k1.x[i] * b1 + k3.x[i] * b3 + k4.x[i] * b4 + k5.x[i] * b5 + k6.x[i] * b6;
        y5.p[i] = y.p[i] + k1.p[i] * b1 + k3.p[i] * b3 + k4.p[i] * b4 + k5.p[i] * b5 + k6.p[i] * b6;
    }
    kerr_derivs(M, a, y5, k7);
    const double e1 = dl * 71.0 / 57600.0, e3 = dl * -71.0 / 16695.0, e4 = dl * 71.0 / 1920.0,
                 e5 = dl * -17253.0 / 339200.0, e6 = dl * 22.0 / 525.0, e7 = dl * -1.0 / 40.0;
    double ex[4], ep[4];
    for (int i = 0; i < 4; ++i) {
        ex[i] = k1.x[i] * e1 + k3.x[i] * e3 + k4.x[i] * e4 + k5.x[i] * e5 + k6.x[i] * e6 + k7.x[i] * e7;
        ep[i] = k1.p[i] * e1 + k3.p[i] * e3 + k4.p[i] * e4 + k5.p[i] * e5 + k6.p[i] * e6 + k7.p[i] * e7;
    }
    const double eps = 1e-10;
    double e = 0.0;
    for (int i = 1; i < 4; ++i) e = smax(e, std::fabs(ex[i]) / (std::fabs(y5.x[i]) + eps));
    for (int i = 0; i < 4; ++i) e = smax(e, std::fabs(ep[i]) / (std::fabs(y5.p[i]) + eps));
    err = e;
}

// include/grrt/geodesic/rk4.h:34-52 — classic RK4.
void rk4_step(double M, double a, const State& y, double dl, State& out) {
    State k1, k2, k3, k4, s;
    kerr_derivs(M, a, y, k1);
    for (int i = 0; i < 4; ++i) { s.x[i] = y.x[i] + k1.x[i] * (dl * 0.5); s.p[i] = y.p[i] + k1.p[i] * (dl * 0.5); }
    kerr_derivs(M, a, s, k2);
    for (int i = 0; i < 4; ++i) { s.x[i] = y.x[i] + k2.x[i] * (dl * 0.5); s.p[i] = y.p[i] + k2.p[i] * (dl * 0.5); }
    kerr_derivs(M, a, s, k3);
    for (int i = 0; i < 4; ++i) { s.x[i] = y.x[i] + k3.x[i] * dl; s.p[i] = y.p[i] + k3.p[i] * dl; }
    kerr_derivs(M, a, s, k4);
    for (int i = 0; i < 4; ++i) {
        out.x[i] = y.x[i] + (k1.x[i] + k2.x[i] * 2.0 + k3.x[i] * 2.0 + k4.x[i]) * (dl / 6.0);
        out.p[i] = y.p[i] + (k1.p[i] + k2.p[i] * 2.0 + k3.p[i] * 2.0 + k4.p[i]) * (dl / 6.0);
    }
}

// src/rk4.cpp:87-136 — PI-controlled adaptive DP45; returns next dl.
double adaptive_step(double M, double a, const State& y, double dlambda, double tol,
                     State& out, int& attempts) {
    const double dl_min = 1e-6, safety = 0.9, p_exp = 0.2, grow_max = 5.0, shrink_max = 0.2;
    double dl = dlambda;
    State y5;
    double err;
    for (int retry = 0; retry < 20; ++retry) {
        dp45_attempt(M, a, y, dl, y5, err);
        ++attempts;
        if (err <= tol) {
            double f = (err > 1e-30) ? safety * std::pow(tol / err, p_exp) : grow_max;
            f = sclamp(f, 1.0, grow_max);
            double next = dl * f;
            next = smin(next, 5.0 * smax(std::fabs(y5.x[1]), 1.0));
            out = y5;
            return next;
        }
        double f = safety * std::pow(tol / err, p_exp);
        f = sclamp(f, shrink_max, 1.0);
        dl *= f;
        if (dl < dl_min) {
            out = y5;
            return smin(dl_min, 5.0 * smax(std::fabs(y5.x[1]), 1.0));
        }
    }
    dp45_attempt(M, a, y, dl, y5, err);
    ++attempts;
    out = y5;
    return dl;
}

// src/rk4.cpp:33-80 — step-doubling RK4: one step of dl against two of
// dl/2, relative max-norm error over r, theta, phi and the four momenta
// (floor 1e-10); accept and double dl when the error is below 1% of tol;
// otherwise halve dl, keeping the half-step result once dl < 1e-6; after 20
// halvings take the two half steps unchecked.  Returns the next dl.
double adaptive_step_doubling(double M, double a, const State& y, double dlambda, double tol,
                              State& out, int& attempts) {
    const double dl_min = 1e-6, eps = 1e-10;
    double dl = dlambda;
    for (int retry = 0; retry < 20; ++retry) {
        State full, mid, half;
        rk4_step(M, a, y, dl, full);
        rk4_step(M, a, y, dl * 0.5, mid);
        rk4_step(M, a, mid, dl * 0.5, half);
        ++attempts;
        double err = 0.0;
        for (int i = 1; i < 4; ++i) {
            const double diff = std::fabs(full.x[i] - half.x[i]);
            const double scale = std::fabs(half.x[i]) + eps;
            err = smax(err, diff / scale);
        }
        for (int i = 0; i < 4; ++i) {
            const double diff = std::fabs(full.p[i] - half.p[i]);
            const double scale = std::fabs(half.p[i]) + eps;
            err = smax(err, diff / scale);
        }
        if (err <= tol) {
            double next = dl;
            if (err < tol * 0.01) next = dl * 2.0;
            next = smin(next, 5.0 * smax(std::fabs(half.x[1]), 1.0));
            out = half;
            return next;
        }
        dl *= 0.5;
        if (dl < dl_min) {
            out = half;
            return smin(dl_min, 5.0 * smax(std::fabs(half.x[1]), 1.0));
        }
    }
    State mid;
    rk4_step(M, a, y, dl * 0.5, mid);
    rk4_step(M, a, mid, dl * 0.5, out);
    ++attempts;
    return dl;
}

// The outer loops' step: the tracer's Dormand–Prince controller, or the
// step-doubling one when the scene asks for it.
double outer_step(const OracleScene& p, const State& y, double dl, State& out, int& attempts) {
    return p.integrator == 1 ? adaptive_step_doubling(p.mass, p.spin_a, y, dl, p.tol, out, attempts)
                             : adaptive_step(p.mass, p.spin_a, y, dl, p.tol, out, attempts);
}

// ------------------------------------------------------------ sampling ---
// src/renderer.cpp:14-20.
double pixel_hash(int i, int j, int s, int ch) {
    uint32_t h = static_cast<uint32_t>(i * 73856093u ^ j * 19349663u ^ s * 83492791u ^ ch * 45678917u);
    h ^= h >> 16;
    h *= 0x45d9f3bu;
    h ^= h >> 16;
    return (h & 0xFFFFu) / 65536.0;
}

// Seeded jitter for the "uniform [0,1)^2 from seed" sampling mode (no
// counterpart in the reference): SplitMix64 finaliser (Steele, Lea & Flood,
// 2014) over a Weyl sequence keyed by the seed, two draws per sample.
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
void seeded_jitter(uint64_t seed, int64_t idx, double& jx, double& jy) {
    const uint64_t key = seed * 0xD1B54A32D192ED03ull;
    const uint64_t w = 0x9E3779B97F4A7C15ull;
    const uint64_t u = static_cast<uint64_t>(idx) * 2u;
    jx = static_cast<double>(mix64(key + (u + 1u) * w) >> 11) * 0x1.0p-53;
    jy = static_cast<double>(mix64(key + (u + 2u) * w) >> 11) * 0x1.0p-53;
}

// -------------------------------------------------------------- camera ---
// src/camera.cpp:7-63 — static-observer tetrad by metric Gram–Schmidt.
struct Camera {
    double pos[4];
    double g[4][4];
    double e0[4], e1[4], e2[4], e3[4];
    double fov;
    int W, H;
};

double mdot(const double g[4][4], const double* u, const double* v) {
    double s = 0.0;
    for (int m = 0; m < 4; ++m)
        for (int n = 0; n < 4; ++n) s += g[m][n] * u[m] * v[n];
    return s;
}
void mnormalize(const double g[4][4], const double* v, double* out) {
    const double n = std::sqrt(std::fabs(mdot(g, v, v)));
    const double inv = 1.0 / n;
    for (int i = 0; i < 4; ++i) out[i] = v[i] * inv;
}
void project_out(const double g[4][4], const double* v, const double* e, double* out) {
    const double ve = mdot(g, v, e), ee = mdot(g, e, e);
    const double f = ve / ee;
    for (int i = 0; i < 4; ++i) out[i] = v[i] - e[i] * f;
}

void build_camera(const OracleScene& sc, Camera& cam) {
    cam.pos[0] = 0.0; cam.pos[1] = sc.r_obs; cam.pos[2] = sc.theta_obs; cam.pos[3] = sc.phi_obs;
    cam.fov = sc.fov; cam.W = sc.width; cam.H = sc.height;
    g_lower(sc.mass, sc.spin_a, cam.pos, cam.g);
    const double gtt = cam.g[0][0];
    cam.e0[0] = 1.0 / std::sqrt(-gtt); cam.e0[1] = cam.e0[2] = cam.e0[3] = 0.0;
    double v[4], t[4];
    const double rdir[4] = {0.0, -1.0, 0.0, 0.0};
    project_out(cam.g, rdir, cam.e0, v);
    mnormalize(cam.g, v, cam.e3);
    const double thdir[4] = {0.0, 0.0, -1.0, 0.0};
    project_out(cam.g, thdir, cam.e0, v);
    project_out(cam.g, v, cam.e3, t);
    mnormalize(cam.g, t, cam.e2);
    const double phdir[4] = {0.0, 0.0, 0.0, 1.0};
    project_out(cam.g, phdir, cam.e0, v);
    project_out(cam.g, v, cam.e3, t);
    project_out(cam.g, t, cam.e2, v);
    mnormalize(cam.g, v, cam.e1);
}

// src/camera.cpp:69-96 — past-directed null covector for a fractional pixel.
void camera_ray(const Camera& cam, double x, double y, State& st) {
    const double alpha = (x - cam.W / 2.0) * cam.fov / cam.W;
    const double beta = (y - cam.H / 2.0) * cam.fov / cam.W;
    const double ca = std::cos(beta) * std::sin(alpha);
    const double cb = std::sin(beta);
    const double cc = std::cos(beta) * std::cos(alpha);
    double pc[4];
    for (int m = 0; m < 4; ++m) {
        const double d = -ca * cam.e1[m] - cb * cam.e2[m] + cc * cam.e3[m];
        pc[m] = -cam.e0[m] + d;
    }
    for (int m = 0; m < 4; ++m) {
        double s = 0.0;
        for (int n = 0; n < 4; ++n) s += cam.g[m][n] * pc[n];
        st.p[m] = s;
        st.x[m] = cam.pos[m];
    }
}

// --------------------------------------------------------------- scene ---
// src/spectrum.cpp:8-131 — CIE 1931 2-degree observer, 380–780 nm by 5 nm.
const double kCIE[81][3] = {
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

constexpr int kSpecN = 1000;
constexpr double kTmin = 1000.0, kTmax = 100000.0;

double planck_lambda(double lambda, double T) {  // src/spectrum.cpp:49-54 (SI)
    const double h = 6.62607015e-34, c = 2.99792458e8, k = 1.380649e-23;
    const double x = h * c / (lambda * k * T);
    if (x > 500.0) return 0.0;
    return (2.0 * h * c * c) / (lambda * lambda * lambda * lambda * lambda * (std::exp(x) - 1.0));
}

void build_spectrum(std::vector<double>& lut) {  // src/spectrum.cpp:56-96
    lut.assign(3 * kSpecN, 0.0);
    for (int i = 0; i < kSpecN; ++i) {
        const double t = kTmin + (kTmax - kTmin) * i / (kSpecN - 1);
        double X = 0.0, Y = 0.0, Z = 0.0;
        for (int w = 0; w < 81; ++w) {
            const double lambda = 380e-9 + w * 5e-9;
            const double B = planck_lambda(lambda, t);
            X += B * kCIE[w][0] * 5e-9;
            Y += B * kCIE[w][1] * 5e-9;
            Z += B * kCIE[w][2] * 5e-9;
        }
        double R = 3.2406 * X - 1.5372 * Y - 0.4986 * Z;
        double G = -0.9689 * X + 1.8758 * Y + 0.0415 * Z;
        double Bv = 0.0557 * X - 0.2040 * Y + 1.0570 * Z;
        R = smax(R, 0.0); G = smax(G, 0.0); Bv = smax(Bv, 0.0);
        const double mx = std::fmax(R, std::fmax(G, Bv));
        if (mx > 0.0) { const double s = 1.0 / mx; R *= s; G *= s; Bv *= s; }
        lut[3 * i + 0] = R; lut[3 * i + 1] = G; lut[3 * i + 2] = Bv;
    }
}

struct Disk {  // src/accretion_disk.cpp:9-147
    double M, a, r_in, r_out, t_peak;
    std::vector<double> flux;
    double flux_max;
    static constexpr int N = 500;

    double wk(double r) const { return std::sqrt(M / (r * r * r)); }
    double Om(double r) const { const double w = wk(r); return w / (1.0 + a * w); }
    double E(double r) const {
        const double w = wk(r), aw = a * w;
        return (1.0 - 2.0 * M / r + aw) / std::sqrt(1.0 - 3.0 * M / r + 2.0 * aw);
    }
    double L(double r) const {
        const double w = wk(r), aw = a * w;
        return std::sqrt(M * r) * (1.0 - 2.0 * aw + a * a / (r * r)) / std::sqrt(1.0 - 3.0 * M / r + 2.0 * aw);
    }
    void build() {
        flux.assign(N, 0.0);
        const double Ei = E(r_in), Li = L(r_in);
        flux_max = 0.0;
        double I = 0.0, prev = 0.0;
        const double fe = 1e-6;
        for (int i = 0; i < N; ++i) {
            const double r = r_in + (r_out - r_in) * i / (N - 1);
            if (i == 0) { flux[i] = 0.0; continue; }
            const double Ep = (E(r + fe) - E(r - fe)) / (2.0 * fe);
            const double Lp = (L(r + fe) - L(r - fe)) / (2.0 * fe);
            const double integrand = (E(r) - Ei) * Lp - (L(r) - Li) * Ep;
            const double dr = (r_out - r_in) / (N - 1);
            I += 0.5 * (prev + integrand) * dr;
            prev = integrand;
            const double om = Om(r), Er = E(r), Lr = L(r);
            const double dOm = (Om(r + fe) - Om(r - fe)) / (2.0 * fe);
            const double den = Er - om * Lr;
            if (std::fabs(den) < 1e-20) { flux[i] = 0.0; continue; }
            const double F = (3.0 * M / (8.0 * kPi * r * r * r)) * (1.0 / den) * (-dOm) * I;
            flux[i] = smax(F, 0.0);
            if (flux[i] > flux_max) flux_max = flux[i];
        }
    }
    double F(double r) const {
        if (r <= r_in || r >= r_out || flux_max <= 0.0) return 0.0;
        const double frac = (r - r_in) / (r_out - r_in) * (N - 1);
        const int idx = static_cast<int>(frac);
        const double t = frac - idx;
        if (idx >= N - 1) return flux[N - 1];
        return flux[idx] * (1.0 - t) + flux[idx + 1] * t;
    }
    double T(double r) const {
        const double f = F(r);
        if (f <= 0.0 || flux_max <= 0.0) return 0.0;
        return t_peak * std::pow(f / flux_max, 0.25);
    }
    double redshift(double rc, const double* p, double r_obs) const {
        const double ut_obs = 1.0 / std::sqrt(1.0 - 2.0 * M / r_obs);
        const double pu_obs = p[0] * ut_obs;
        const double w = wk(rc), aw = a * w;
        const double ut_e = 1.0 / std::sqrt(1.0 - 3.0 * M / rc + 2.0 * aw);
        const double up_e = Om(rc) * ut_e;
        const double pu_e = p[0] * ut_e + p[3] * up_e;
        if (std::fabs(pu_obs) < 1e-30) return 1.0;
        return pu_e / pu_obs;
    }
};

struct Stars {  // src/celestial_sphere.cpp:9-90
    static constexpr int GT = 180, GP = 360;
    std::vector<double> th, ph, b;
    std::vector<std::vector<int>> grid;
    void build() {
        std::mt19937 rng(42u);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        th.resize(5000); ph.resize(5000); b.resize(5000);
        for (int i = 0; i < 5000; ++i) {
            th[i] = std::acos(1.0 - 2.0 * U(rng));
            ph[i] = 2.0 * kPi * U(rng) - kPi;
            const double u = U(rng);
            b[i] = smin(0.1 * std::pow(smax(u, 0.001), -2.5), 50.0);
        }
        grid.assign(GT * GP, {});
        for (int i = 0; i < 5000; ++i) {
            int tb = static_cast<int>(th[i] * GT / kPi);
            int pb = static_cast<int>((ph[i] + kPi) * GP / (2.0 * kPi));
            tb = std::clamp(tb, 0, GT - 1);
            pb = std::clamp(pb, 0, GP - 1);
            grid[tb * GP + pb].push_back(i);
        }
    }
    static int cell(int t, int p) {
        t = std::clamp(t, 0, GT - 1);
        p = ((p % GP) + GP) % GP;
        return t * GP + p;
    }
    double sample(double theta, double phi) const {
        theta = std::fmod(theta, 2.0 * kPi);
        if (theta < 0.0) theta += 2.0 * kPi;
        if (theta > kPi) { theta = 2.0 * kPi - theta; phi += kPi; }
        phi = std::fmod(phi, 2.0 * kPi);
        if (phi > kPi) phi -= 2.0 * kPi;
        if (phi < -kPi) phi += 2.0 * kPi;
        const int tb = static_cast<int>(theta * GT / kPi);
        const int pb = static_cast<int>((phi + kPi) * GP / (2.0 * kPi));
        for (int dt = -1; dt <= 1; ++dt)
            for (int dp = -1; dp <= 1; ++dp)
                for (int si : grid[cell(tb + dt, pb + dp)]) {
                    double dth = theta - th[si];
                    double dph = phi - ph[si];
                    if (dph > kPi) dph -= 2.0 * kPi;
                    if (dph < -kPi) dph += 2.0 * kPi;
                    const double st = std::sin(theta);
                    const double d2 = dth * dth + dph * dph * st * st;
                    if (d2 < 0.003 * 0.003) return b[si];
                }
        return 0.0;
    }
};

const std::vector<double>& spectrum_lut() {
    static const std::vector<double> lut = [] { std::vector<double> v; build_spectrum(v); return v; }();
    return lut;
}
const Stars& star_catalog() {
    static const Stars s = [] { Stars t; t.build(); return t; }();
    return s;
}

void chroma(double T, double out[3]) {  // src/spectrum.cpp:115-124
    const auto& lut = spectrum_lut();
    T = sclamp(T, kTmin, kTmax);
    const double frac = (T - kTmin) / (kTmax - kTmin) * (kSpecN - 1);
    const int idx = static_cast<int>(frac);
    const double t = frac - idx;
    if (idx >= kSpecN - 1) { for (int c = 0; c < 3; ++c) out[c] = lut[3 * (kSpecN - 1) + c]; return; }
    for (int c = 0; c < 3; ++c) out[c] = lut[3 * idx + c] * (1.0 - t) + lut[3 * (idx + 1) + c] * t;
}

struct Scene {
    OracleScene p;
    Camera cam;
    Disk disk;
    bool has_disk;
    double r_h;
};

void build_scene(const OracleScene& sc, Scene& s) {
    s.p = sc;
    build_camera(sc, s.cam);
    s.has_disk = sc.disk_enabled != 0;
    if (s.has_disk) {
        s.disk.M = sc.mass; s.disk.a = sc.spin_a; s.disk.r_in = sc.r_in; s.disk.r_out = sc.r_out;
        s.disk.t_peak = sc.t_peak;
        s.disk.build();
    }
    s.r_h = horizon(sc.mass, sc.spin_a);
    (void)spectrum_lut();
    if (sc.stars) (void)star_catalog();
}

// src/geodesic_tracer.cpp:23-181 (thin-disk branch; the volumetric branches
// are inactive when no volumetric disk exists) plus src/renderer.cpp:56-61.
void trace(const Scene& S, const State& init, OracleRay& o, bool diag) {
    const double M = S.p.mass, a = S.p.spin_a;
    const double half_pi = kPi / 2.0;
    State st = init, prev;
    double col[3] = {0.0, 0.0, 0.0};
    double dl = 0.01 * S.p.r_obs;
    o.steps = o.attempts = o.n_cross = 0;
    o.r_hit = o.g_hit = 0.0;
    o.h_max = 0.0;
    if (diag) o.h_max = std::fabs(hamiltonian(M, a, st.x, st.p));
    int term = 2;
    for (int step = 0; step < S.p.max_steps; ++step) {
        const double r = st.x[1];
        if (r < S.r_h + 0.01) { term = 0; break; }
        if (r > S.p.r_escape) { term = 1; break; }
        prev = st;
        State nxt;
        dl = outer_step(S.p, st, dl, nxt, o.attempts);
        st = nxt;
        ++o.steps;
        if (diag) o.h_max = smax(o.h_max, std::fabs(hamiltonian(M, a, st.x, st.p)));
        if (S.has_disk) {
            const double dp = prev.x[2] - half_pi, dn = st.x[2] - half_pi;
            if (dp * dn < 0.0 && std::fabs(dp - dn) > 1e-12) {
                const double frac = -dp / (dn - dp);
                const double rc = prev.x[1] + frac * (st.x[1] - prev.x[1]);
                double pcr[4];
                for (int m = 0; m < 4; ++m) pcr[m] = prev.p[m] + frac * (st.p[m] - prev.p[m]);
                if (rc >= S.disk.r_in && rc <= S.disk.r_out) {
                    // src/accretion_disk.cpp:132-147
                    const double T = S.disk.T(rc);
                    const double g = S.disk.redshift(rc, pcr, S.p.r_obs);
                    if (o.n_cross == 0) { o.r_hit = rc; o.g_hit = g; }
                    ++o.n_cross;
                    if (T > 0.0) {
                        const double Tobs = g * T;
                        if (!(Tobs < 100.0)) {
                            double ch[3];
                            chroma(Tobs, ch);
                            const double ratio = T / kTmax;
                            const double lum = ratio * ratio * ratio * ratio;
                            const double g4 = g * g * g * g;
                            const double w = lum * g4;
                            for (int c = 0; c < 3; ++c) col[c] += ch[c] * w;
                        }
                    }
                }
            }
        }
    }
    // If the loop ran out, the last state was never checked; the reference
    // returns MaxSteps in that case (src/geodesic_tracer.cpp:179-180).
    if (term == 1 && S.p.stars) {
        const double b = star_catalog().sample(st.x[2], st.x[3]);
        for (int c = 0; c < 3; ++c) col[c] += b;
    }
    for (int c = 0; c < 3; ++c) o.color[c] = col[c];
    for (int m = 0; m < 4; ++m) { o.x_final[m] = st.x[m]; o.p_final[m] = st.p[m]; }
    o.term = term;
    o.cls = (o.n_cross > 0) ? 2 : (term == 0 ? 0 : (term == 1 ? 1 : 3));
    o.h_final = diag ? std::fabs(hamiltonian(M, a, st.x, st.p)) : 0.0;
    o.pad = 0;
}

int samples_per_pixel(int mode, int spp, int& grid) {
    if (mode == ORACLE_JITTER_CENTER) { grid = 1; return 1; }
    if (mode == ORACLE_JITTER_HASH) {
        grid = std::max(1, static_cast<int>(std::sqrt(static_cast<double>(spp < 1 ? 1 : spp))));
        return grid * grid;
    }
    grid = 0;
    return spp < 1 ? 1 : spp;
}

// Sample position for sample id (pixel*S + s); src/renderer.cpp:48-55.
void sample_position_of(const Scene& S, int mode, int grid, int nS, uint64_t seed, int64_t id,
                double& px, double& py) {
    const int64_t pixel = id / nS;
    const int s = static_cast<int>(id % nS);
    const int j = static_cast<int>(pixel / S.p.width);
    const int i = static_cast<int>(pixel % S.p.width);
    if (mode == ORACLE_JITTER_CENTER) { px = i + 0.5; py = j + 0.5; return; }
    if (mode == ORACLE_JITTER_HASH) {
        const double cell = 1.0 / grid;
        const int sy = s / grid, sx = s % grid;
        const double jx = pixel_hash(i, j, s, 0), jy = pixel_hash(i, j, s, 1);
        px = i + (sx + jx) * cell;
        py = j + (sy + jy) * cell;
        return;
    }
    double jx, jy;
    seeded_jitter(seed, id, jx, jy);
    px = i + jx;
    py = j + jy;
}


// ------------------------------------------------------- volumetric disk ---
constexpr double kC = 2.997924e10, kH = 6.626070e-27, kKB = 1.380649e-16;   // constants.h:11-35

struct Vol {
    const OracleVol& v;
    // src/volumetric_disk.cpp:164-171
    double interp_radial(const double* lut, double r) const {
        if (r <= v.r_min) return lut[0];
        if (r >= v.r_outer) return lut[v.n_r - 1];
        const double frac = (r - v.r_min) / (v.r_outer - v.r_min) * (v.n_r - 1);
        const int idx = std::clamp(static_cast<int>(frac), 0, v.n_r - 2);
        const double t = frac - idx;
        return lut[idx] * (1.0 - t) + lut[idx + 1] * t;
    }
    // src/volumetric_disk.cpp:173-214
    double interp_2d(const double* lut, double r, double z_abs) const {
        const double r_frac = sclamp((r - v.r_min) / (v.r_outer - v.r_min) * (v.n_r - 1), 0.0,
                                     static_cast<double>(v.n_r - 1));
        const int ri = std::min(static_cast<int>(r_frac), v.n_r - 2);
        const double tr = r_frac - ri;
        const double zm_lo = v.z_max[ri];
        const double zm_hi = v.z_max[std::min(ri + 1, v.n_r - 1)];
        if ((zm_lo <= 0.0 || z_abs >= zm_lo) && (zm_hi <= 0.0 || z_abs >= zm_hi)) return 0.0;
        double lo = 0.0;
        if (zm_lo > 0.0 && z_abs < zm_lo) {
            const double zf = sclamp(z_abs / zm_lo * (v.n_z - 1), 0.0, static_cast<double>(v.n_z - 1));
            const int zi = std::min(static_cast<int>(zf), v.n_z - 2);
            const double tz = zf - zi;
            lo = (1.0 - tz) * lut[ri * v.n_z + zi] + tz * lut[ri * v.n_z + zi + 1];
        }
        double hi = 0.0;
        const int ri1 = std::min(ri + 1, v.n_r - 1);
        if (zm_hi > 0.0 && z_abs < zm_hi) {
            const double zf = sclamp(z_abs / zm_hi * (v.n_z - 1), 0.0, static_cast<double>(v.n_z - 1));
            const int zi = std::min(static_cast<int>(zf), v.n_z - 2);
            const double tz = zf - zi;
            hi = (1.0 - tz) * lut[ri1 * v.n_z + zi] + tz * lut[ri1 * v.n_z + zi + 1];
        }
        return (1.0 - tr) * lo + tr * hi;
    }
    double scale_height(double r) const { return interp_radial(v.H, r); }
    double z_max_at(double r) const { return interp_radial(v.z_max, r); }
    bool inside(double r, double z) const {   // :153-160
        if (r <= v.r_horizon || r > v.r_outer) return false;
        return std::fabs(z) < z_max_at(r) + 1.5 * scale_height(r);
    }
    double taper(double r) const {   // :143-147
        if (r >= v.r_isco) return 1.0;
        const double dr = v.r_isco - r;
        return std::exp(-(dr * dr) / (v.taper_w * v.taper_w));
    }
    // src/noise.cpp:38-115
    double simplex(double x, double y, double z) const { return simplex_noise(v.perm, x, y, z); }
    static double simplex_noise(const int* perm, double x, double y, double z) {
        static const double grad3[12][3] = {{1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0},
                                            {1, 0, 1}, {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
                                            {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
        const double F3 = 1.0 / 3.0, G3 = 1.0 / 6.0;
        const double s = (x + y + z) * F3;
        const int i = static_cast<int>(std::floor(x + s)), j = static_cast<int>(std::floor(y + s)),
                  k = static_cast<int>(std::floor(z + s));
        const double t = (i + j + k) * G3;
        const double X0 = i - t, Y0 = j - t, Z0 = k - t;
        const double x0 = x - X0, y0 = y - Y0, z0 = z - Z0;
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
        const double x1 = x0 - i1 + G3, y1 = y0 - j1 + G3, z1 = z0 - k1 + G3;
        const double x2 = x0 - i2 + 2 * G3, y2 = y0 - j2 + 2 * G3, z2 = z0 - k2 + 2 * G3;
        const double x3 = x0 - 1 + 3 * G3, y3 = y0 - 1 + 3 * G3, z3 = z0 - 1 + 3 * G3;
        const int ii = i & 255, jj = j & 255, kk = k & 255;
        auto contrib = [&](double cx, double cy, double cz, int gi) {
            double t0 = 0.6 - cx * cx - cy * cy - cz * cz;
            if (t0 < 0.0) return 0.0;
            t0 *= t0;
            const double* g = grad3[gi % 12];
            return t0 * t0 * (g[0] * cx + g[1] * cy + g[2] * cz);
        };
        const double n0 = contrib(x0, y0, z0, perm[ii + perm[jj + perm[kk]]]);
        const double n1 = contrib(x1, y1, z1, perm[ii + i1 + perm[jj + j1 + perm[kk + k1]]]);
        const double n2 = contrib(x2, y2, z2, perm[ii + i2 + perm[jj + j2 + perm[kk + k2]]]);
        const double n3 = contrib(x3, y3, z3, perm[ii + 1 + perm[jj + 1 + perm[kk + 1]]]);
        return 32.0 * (n0 + n1 + n2 + n3);
    }
    double fbm(double x, double y, double z, int oct) const {
        double res = 0.0, amp = 1.0, freq = 1.0;
        for (int i = 0; i < oct; ++i) {
            res += amp * simplex(x * freq, y * freq, z * freq);
            amp *= 0.5;
            freq *= 2.0;
        }
        return res;
    }
    // :226-257
    double density(double r, double z, double phi) const {
        if (r <= v.r_horizon || r > v.r_outer) return 0.0;
        const double z_abs = std::fabs(z);
        const double zm = z_max_at(r), H = scale_height(r);
        double edge = 1.0, zl = z_abs;
        if (z_abs >= zm) {
            const double dz = z_abs - zm;
            if (dz >= 1.5 * H) return 0.0;
            edge = std::exp(-2.0 * (dz / H) * (dz / H));
            zl = zm * 0.9999;
        }
        const double base = interp_radial(v.rho_mid, r) * interp_2d(v.rho_prof, r, zl) * v.rho_scale *
                            taper(r) * edge;
        const double nx = r * std::cos(phi) / v.noise_scale;
        const double ny = r * std::sin(phi) / v.noise_scale;
        const double nz = z / v.noise_scale;
        const double n = fbm(nx, ny, nz, v.octaves);
        return smax(0.0, base * (1.0 + v.turbulence * n));
    }
    // :264-273
    double temperature(double r, double z) const {
        if (r <= v.r_horizon || r > v.r_outer) return 0.0;
        const double z_abs = std::fabs(z);
        const double zm = z_max_at(r), H = scale_height(r);
        if (z_abs >= zm + 1.5 * H) return 0.0;
        return interp_2d(v.T_prof, r, smin(z_abs, zm * 0.9999));
    }
    double omega_orb(double r) const { return std::sqrt(v.M) / (r * std::sqrt(r) + v.a * std::sqrt(v.M)); }
    // :95-133
    void circular(double r, double& ut, double& uphi) const {
        const double M = v.M, a = v.a, Om = omega_orb(r);
        const double gtt = -(1.0 - 2.0 * M / r), gtp = -2.0 * M * a / r, gpp = r * r + a * a + 2.0 * M * a * a / r;
        const double den = -(gtt + 2.0 * gtp * Om + gpp * Om * Om);
        ut = 1.0 / std::sqrt(smax(den, 1e-30));
        uphi = Om * ut;
    }
    void plunging(double r, double theta, double& ut, double& ur, double& uphi) const {
        const double M = v.M, a = v.a, E = v.E_isco, L = v.L_isco;
        const double Delta = r * r - 2.0 * M * r + a * a;
        const double Sigma = r * r + a * a * std::cos(theta) * std::cos(theta);
        ut = (E * (r * r + a * a + 2.0 * M * a * a / r) - 2.0 * M * a * L / r) / Delta;
        uphi = (L * (1.0 - 2.0 * M / r) + 2.0 * M * a * E / r) / Delta;
        const double EaL = E * (r * r + a * a) - a * L, LaE = L - a * E;
        const double R = EaL * EaL - Delta * (r * r + LaE * LaE);
        ur = -std::sqrt(smax(0.0, R)) / Sigma;
    }
    // src/opacity.cpp:218-346
    static void log_cell(double lv, double lo, double hi, int n, int& idx, double& frac) {
        double t = (lv - lo) / (hi - lo) * (n - 1);
        t = sclamp(t, 0.0, static_cast<double>(n - 2));
        idx = static_cast<int>(t);
        frac = t - idx;
    }
    double kappa_abs(double nu, double rho, double T) const {
        int inu, irho, iT;
        double fnu, frho, fT;
        log_cell(std::log10(nu), v.lnu0, v.lnu1, v.n_nu, inu, fnu);
        log_cell(std::log10(rho), v.lrho0, v.lrho1, v.n_rho, irho, frho);
        log_cell(std::log10(T), v.lT0, v.lT1, v.n_T, iT, fT);
        auto idx = [&](int i, int j, int k) { return static_cast<size_t>(i) * v.n_rho * v.n_T + j * v.n_T + k; };
        auto sl = [](double x) { return std::log(smax(x, 1e-100)); };
        const double l00 = sl(v.kabs[idx(inu, irho, iT)]) * (1 - fT) + sl(v.kabs[idx(inu, irho, iT + 1)]) * fT;
        const double l01 = sl(v.kabs[idx(inu, irho + 1, iT)]) * (1 - fT) + sl(v.kabs[idx(inu, irho + 1, iT + 1)]) * fT;
        const double l10 = sl(v.kabs[idx(inu + 1, irho, iT)]) * (1 - fT) + sl(v.kabs[idx(inu + 1, irho, iT + 1)]) * fT;
        const double l11 = sl(v.kabs[idx(inu + 1, irho + 1, iT)]) * (1 - fT) + sl(v.kabs[idx(inu + 1, irho + 1, iT + 1)]) * fT;
        const double l0 = l00 * (1 - frho) + l01 * frho;
        const double l1 = l10 * (1 - frho) + l11 * frho;
        return std::exp(l0 * (1 - fnu) + l1 * fnu);
    }
    double kappa_es(double rho, double T) const {
        int irho, iT;
        double frho, fT;
        log_cell(std::log10(rho), v.lrho0, v.lrho1, v.n_rho, irho, frho);
        log_cell(std::log10(T), v.lT0, v.lT1, v.n_T, iT, fT);
        const double* q = v.kes;
        const int n = v.n_T;
        return q[irho * n + iT] * (1 - frho) * (1 - fT) + q[irho * n + iT + 1] * (1 - frho) * fT +
               q[(irho + 1) * n + iT] * frho * (1 - fT) + q[(irho + 1) * n + iT + 1] * frho * fT;
    }
};

// src/opacity.cpp:209-214
double planck_nu(double nu, double T) {
    if (nu <= 0.0 || T <= 0.0) return 0.0;
    const double x = kH * nu / (kKB * T);
    if (x > 500.0) return 0.0;
    return (2.0 * kH * nu * nu * nu / (kC * kC)) / (std::exp(x) - 1.0);
}

// One emitting sample of the backward transfer (src/geodesic_tracer.cpp:
// 311-315 RGB, :724-728 spectral): J += T S (1 - e^-dtau), T *= e^-dtau.
inline void rt_step(double& J, double& T, double S, double dtau) {
    const double ex = std::exp(-dtau);
    J += T * S * (1.0 - ex);
    T *= ex;
}

// Gas 4-velocity at (r, theta) as the ray-march takes it: circular orbit
// outside the ISCO, plunging geodesic inside (src/volumetric_disk.cpp:95-133).
inline void gas_velocity(const Vol& VD, double r, double theta, double& ut, double& ur, double& uphi) {
    ut = ur = uphi = 0.0;
    if (r >= VD.v.r_isco) VD.circular(r, ut, uphi);
    else VD.plunging(r, theta, ut, ur, uphi);
}

// Backward ray-march (src/geodesic_tracer.cpp:183-337 RGB, :605-746
// spectral).  nb channels at frequencies nu; RGB mode exits on opacity,
// spectral mode on tau > 10 in every bin; step control on bin `med`.
// Returns the RK4 steps taken.
int raymarch(const Scene& S, const Vol& VD, State& state, int nb, const double* nu, bool spectral,
             double* J, double* T, double* tau, bool& emitted) {
    const double M = S.p.mass, a = S.p.spin_a;
    const double ut_obs = 1.0 / std::sqrt(1.0 - 2.0 / S.p.r_obs);
    const int med = spectral ? nb / 2 : 1;
    double r = state.x[1];
    const double z_start = r * std::cos(state.x[2]);
    const double H_start = VD.scale_height(r);
    double ds = VD.inside(r, z_start) ? H_start / 16.0 : smin(std::fabs(z_start) / 8.0, H_start * 2.0);
    bool been_inside = VD.inside(r, z_start);
    int step_count = 0;
    while (step_count < 4096) {
        State ns;
        rk4_step(M, a, state, ds, ns);
        ++step_count;
        r = ns.x[1];
        const double theta = ns.x[2], phi = ns.x[3];
        const double z = r * std::cos(theta);
        if (r < VD.v.r_horizon) { state = ns; break; }
        if (r > VD.v.r_outer) { state = ns; break; }
        if (!spectral) {
            if (T[0] < 1e-6 && T[1] < 1e-6 && T[2] < 1e-6) { state = ns; break; }
        } else {
            bool thick = true;
            for (int c = 0; c < nb; ++c)
                if (tau[c] <= 10.0) { thick = false; break; }
            if (thick) { state = ns; break; }
        }
        const double H = VD.scale_height(r);
        if (!VD.inside(r, z)) {
            const double zm = VD.z_max_at(r);
            if (been_inside && std::fabs(z) > zm + 3.0 * H) { state = ns; break; }
            if (!been_inside) {
                ds = smin(std::fabs(z) / 8.0, H * 2.0);
                ds = smax(ds, H / 64.0);
            } else {
                ds = sclamp(H / 4.0, H / 64.0, H);
            }
            state = ns;
            continue;
        }
        been_inside = true;
        const double rho = VD.density(r, z, phi);
        const double Tl = VD.temperature(r, std::fabs(z));
        if (rho <= 0.0 || Tl <= 0.0) { state = ns; continue; }
        double ut, ur, uphi;
        gas_velocity(VD, r, theta, ut, ur, uphi);
        const double pue = ns.p[0] * ut + ns.p[1] * ur + ns.p[3] * uphi;
        const double puo = ns.p[0] * ut_obs;
        const double g = pue / puo;
        const double dsp = std::fabs(pue) * std::fabs(ds);
        for (int c = 0; c < nb; ++c) {
            const double nue = std::fabs(g) * nu[c];
            const double kabs = VD.kappa_abs(nue, rho, Tl);
            const double kes = VD.kappa_es(rho, Tl);
            const double ktot = kabs + kes;
            const double eps = (ktot > 0.0) ? kabs / ktot : 1.0;
            const double dtau = ktot * rho * dsp;
            if (spectral) tau[c] += dtau;
            const double Bnu = planck_nu(nue, Tl);
            const double Sv = eps * Bnu / (nue * nue * nue);
            rt_step(J[c], T[c], Sv, dtau);
        }
        const double alpha_tot = (VD.kappa_abs(std::fabs(g) * nu[med], rho, Tl) + VD.kappa_es(rho, Tl)) * rho;
        const double ds_tau = (alpha_tot > 0.0) ? 0.05 / alpha_tot : ds * 2.0;
        const double ds_geo = 0.1 * smax(r - VD.v.r_horizon, 0.5);
        ds = smin(ds_tau, ds_geo);
        ds = sclamp(ds, H / 64.0, H);
        state = ns;
        emitted = true;
    }
    return step_count;
}

// Outer loop with a volumetric disk (src/geodesic_tracer.cpp:23-181 RGB,
// :495-603 spectral).  Writes the ray record (RGB colour incl. stars) or the
// spectrum.
void vol_trace(const Scene& S, const Vol& VD, const State& init, int nbins, const double* nu_spec,
               OracleRay& o, double* spectrum) {
    const double M = S.p.mass, a = S.p.spin_a;
    const double half_pi = kPi / 2.0;
    const bool spectral = nbins > 0;
    const double nu_rgb[3] = {kC / 450e-7, kC / 550e-7, kC / 650e-7};
    const int nb = spectral ? nbins : 3;
    const double* nu = spectral ? nu_spec : nu_rgb;
    std::vector<double> J(nb, 0.0), T(nb, 1.0), tau(nb, 0.0);
    State st = init, prev;
    double dl = 0.01 * S.p.r_obs;
    int term = 2;
    bool emitted = false;
    int march = 0;
    o.steps = o.attempts = o.n_cross = 0;
    for (int step = 0; step < S.p.max_steps; ++step) {
        const double r = st.x[1];
        if (r < S.r_h + 0.01) { term = 0; break; }
        if (r > S.p.r_escape) { term = 1; break; }
        prev = st;
        {
            const double theta = st.x[2];
            const double z = r * std::cos(theta);
            if (r >= VD.v.r_horizon * 0.9 && r <= VD.v.r_outer * 1.5) {
                const double H = VD.scale_height(r), zm = VD.z_max_at(r);
                if (std::fabs(z) < zm + 6.0 * H) {
                    State d;
                    kerr_derivs(M, a, st, d);
                    const double dz = std::fabs(std::cos(theta) * d.x[1] - r * std::sin(theta) * d.x[2]);
                    if (dz > 1e-20) dl = smin(dl, smax(0.25 * H / dz, H / 64.0));
                }
            }
        }
        State nxt;
        dl = outer_step(S.p, st, dl, nxt, o.attempts);
        st = nxt;
        ++o.steps;
        const bool opaque = !spectral && (T[0] < 1e-6 && T[1] < 1e-6 && T[2] < 1e-6);
        if (!opaque) {
            const double tp = prev.x[2], tn = st.x[2];
            const double dp = tp - half_pi, dn = tn - half_pi;
            const double rn = st.x[1], rp = prev.x[1];
            const double zn = rn * std::cos(tn), zp = rp * std::cos(tp);
            const bool crossed = (dp * dn < 0.0) && std::fabs(dp - dn) > 1e-12;
            const bool inside_now = VD.inside(rn, zn);
            const double zmn = VD.z_max_at(rn), Hn = VD.scale_height(rn), Hp = VD.scale_height(rp);
            const bool near = (std::fabs(zn) < zmn + 2.0 * Hn || std::fabs(zp) < VD.z_max_at(rp) + 2.0 * Hp) &&
                              rn >= VD.v.r_horizon && rn <= VD.v.r_outer;
            if (crossed || inside_now || near) {
                const double rlo = smin(rp, rn), rhi = smax(rp, rn);
                if (rhi >= VD.v.r_horizon && rlo <= VD.v.r_outer) {
                    State entry = prev;
                    const double re = entry.x[1];
                    if (re >= VD.v.r_horizon * 0.9 && re <= VD.v.r_outer * 1.5) {
                        march += raymarch(S, VD, entry, nb, nu, spectral, J.data(), T.data(), tau.data(), emitted);
                        ++o.n_cross;
                        st = entry;
                        continue;
                    }
                }
            }
        }
    }
    double col[3] = {0.0, 0.0, 0.0};
    if (spectral) {
        for (int c = 0; c < nb; ++c) spectrum[c] = J[c] * nu[c] * nu[c] * nu[c];
    } else {
        for (int c = 0; c < 3; ++c) col[c] += J[c] * nu[c] * nu[c] * nu[c];
        if (term == 1 && S.p.stars) {
            const double b = star_catalog().sample(st.x[2], st.x[3]);
            for (int c = 0; c < 3; ++c) col[c] += b;
        }
    }
    for (int c = 0; c < 3; ++c) o.color[c] = col[c];
    for (int m = 0; m < 4; ++m) { o.x_final[m] = st.x[m]; o.p_final[m] = st.p[m]; }
    o.term = term;
    o.cls = emitted ? 2 : (term == 0 ? 0 : (term == 1 ? 1 : 3));
    o.r_hit = o.g_hit = o.h_max = o.h_final = 0.0;
    o.pad = march;                   // ray-march RK4 steps over all entries
}

}  // namespace

extern "C" {

void oracle_fast_sincos(double x, double* s, double* c) { fast_sincos(x, *s, *c); }

void oracle_kerr_derivs(double M, double a, const double x[4], const double p[4], double dx[4],
                        double dp[4]) {
    State s, d;
    for (int i = 0; i < 4; ++i) { s.x[i] = x[i]; s.p[i] = p[i]; }
    kerr_derivs(M, a, s, d);
    for (int i = 0; i < 4; ++i) { dx[i] = d.x[i]; dp[i] = d.p[i]; }
}

void oracle_g_lower(double M, double a, const double x[4], double g[16]) {
    double gg[4][4];
    g_lower(M, a, x, gg);
    std::memcpy(g, gg, sizeof(gg));
}

double oracle_hamiltonian(double M, double a, const double x[4], const double p[4]) {
    return hamiltonian(M, a, x, p);
}
double oracle_horizon(double M, double a) { return horizon(M, a); }
double oracle_isco(double M, double a) { return isco(M, a); }

void oracle_dp45_attempt(double M, double a, const double y[8], double dl, double y5[8], double* err) {
    State s, o;
    for (int i = 0; i < 4; ++i) { s.x[i] = y[i]; s.p[i] = y[4 + i]; }
    dp45_attempt(M, a, s, dl, o, *err);
    for (int i = 0; i < 4; ++i) { y5[i] = o.x[i]; y5[4 + i] = o.p[i]; }
}

void oracle_rk4_step(double M, double a, const double y[8], double dl, double out[8]) {
    State s, o;
    for (int i = 0; i < 4; ++i) { s.x[i] = y[i]; s.p[i] = y[4 + i]; }
    rk4_step(M, a, s, dl, o);
    for (int i = 0; i < 4; ++i) { out[i] = o.x[i]; out[4 + i] = o.p[i]; }
}

double oracle_adaptive_step(double M, double a, const double y[8], double dl, double tol,
                            double out[8], int* attempts) {
    State s, o;
    for (int i = 0; i < 4; ++i) { s.x[i] = y[i]; s.p[i] = y[4 + i]; }
    int n = 0;
    const double next = adaptive_step(M, a, s, dl, tol, o, n);
    for (int i = 0; i < 4; ++i) { out[i] = o.x[i]; out[4 + i] = o.p[i]; }
    if (attempts) *attempts = n;
    return next;
}

double oracle_adaptive_step_doubling(double M, double a, const double y[8], double dl, double tol,
                                     double out[8], int* attempts) {
    State st, o;
    for (int i = 0; i < 4; ++i) { st.x[i] = y[i]; st.p[i] = y[4 + i]; }
    int n = 0;
    const double next = adaptive_step_doubling(M, a, st, dl, tol, o, n);
    for (int i = 0; i < 4; ++i) { out[i] = o.x[i]; out[4 + i] = o.p[i]; }
    if (attempts) *attempts = n;
    return next;
}

double oracle_pixel_hash(int i, int j, int s, int ch) { return pixel_hash(i, j, s, ch); }

void oracle_seeded_jitter(uint64_t seed, int64_t idx, double* jx, double* jy) {
    seeded_jitter(seed, idx, *jx, *jy);
}

void oracle_camera_ray(const OracleScene* sc, double px, double py, double state[8]) {
    Camera cam;
    build_camera(*sc, cam);
    State st;
    camera_ray(cam, px, py, st);
    for (int i = 0; i < 4; ++i) { state[i] = st.x[i]; state[4 + i] = st.p[i]; }
}

void oracle_flux_lut(const OracleScene* sc, double* lut, double* fmax) {
    Disk d;
    d.M = sc->mass; d.a = sc->spin_a; d.r_in = sc->r_in; d.r_out = sc->r_out; d.t_peak = sc->t_peak;
    d.build();
    std::memcpy(lut, d.flux.data(), sizeof(double) * Disk::N);
    *fmax = d.flux_max;
}

void oracle_spectrum_lut(double* rgb) {
    const auto& l = spectrum_lut();
    std::memcpy(rgb, l.data(), sizeof(double) * l.size());
}

void oracle_stars(double* out) {
    const Stars& s = star_catalog();
    for (int i = 0; i < 5000; ++i) { out[3 * i] = s.th[i]; out[3 * i + 1] = s.ph[i]; out[3 * i + 2] = s.b[i]; }
}

void oracle_star_sample(double theta, double phi, double out[3]) {
    const double b = star_catalog().sample(theta, phi);
    out[0] = out[1] = out[2] = b;
}

void oracle_disk_emission(const OracleScene* sc, double rc, double p_t, double p_phi, double out[3],
                          double* g_out, double* t_out) {
    Scene S;
    build_scene(*sc, S);
    const double p[4] = {p_t, 0.0, 0.0, p_phi};
    const double T = S.disk.T(rc);
    const double g = S.disk.redshift(rc, p, sc->r_obs);
    out[0] = out[1] = out[2] = 0.0;
    *g_out = g;
    *t_out = T;
    if (T <= 0.0) return;
    const double Tobs = g * T;
    if (Tobs < 100.0) return;
    double ch[3];
    chroma(Tobs, ch);
    const double ratio = T / kTmax;
    const double w = ratio * ratio * ratio * ratio * (g * g * g * g);
    for (int c = 0; c < 3; ++c) out[c] = ch[c] * w;
}

void oracle_trace(const OracleScene* sc, const double state[8], OracleRay* out) {
    Scene S;
    build_scene(*sc, S);
    State st;
    for (int i = 0; i < 4; ++i) { st.x[i] = state[i]; st.p[i] = state[4 + i]; }
    trace(S, st, *out, true);
}

void oracle_trace_samples(const OracleScene* sc, int mode, int spp, uint64_t seed,
                          const int64_t* ids, int64_t n, OracleRay* rays, int threads) {
    Scene S;
    build_scene(*sc, S);
    int grid;
    const int nS = samples_per_pixel(mode, spp, grid);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(guided)
#endif
    for (int64_t k = 0; k < n; ++k) {
        double px, py;
        sample_position_of(S, mode, grid, nS, seed, ids[k], px, py);
        State st;
        camera_ray(S.cam, px, py, st);
        trace(S, st, rays[k], true);
    }
}

int64_t oracle_render(const OracleScene* sc, int mode, int spp, uint64_t seed, float* rgba,
                      uint8_t* cls, int threads) {
    Scene S;
    build_scene(*sc, S);
    int grid;
    const int nS = samples_per_pixel(mode, spp, grid);
    const double inv = 1.0 / nS;
    const int64_t npix = static_cast<int64_t>(sc->width) * sc->height;
    int64_t total_steps = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(guided) reduction(+ : total_steps)
#endif
    for (int64_t pix = 0; pix < npix; ++pix) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < nS; ++s) {
            const int64_t id = pix * nS + s;
            double px, py;
            sample_position_of(S, mode, grid, nS, seed, id, px, py);
            State st;
            camera_ray(S.cam, px, py, st);
            OracleRay o;
            trace(S, st, o, false);
            total_steps += o.steps;
            for (int c = 0; c < 3; ++c) acc[c] = acc[c] + o.color[c];
            if (cls) cls[id] = static_cast<uint8_t>(o.cls);
        }
        rgba[4 * pix + 0] = static_cast<float>(acc[0] * inv);
        rgba[4 * pix + 1] = static_cast<float>(acc[1] * inv);
        rgba[4 * pix + 2] = static_cast<float>(acc[2] * inv);
        rgba[4 * pix + 3] = 1.0f;
    }
    return total_steps;
}

void oracle_vol_trace_samples(const OracleScene* sc, const OracleVol* V, int mode, int spp, uint64_t seed,
                              const int64_t* ids, int64_t n, int nbins, const double* nu, OracleRay* rays,
                              double* spectra, int threads) {
    Scene S;
    build_scene(*sc, S);
    const Vol VD{*V};
    int grid;
    const int nS = samples_per_pixel(mode, spp, grid);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic)
#endif
    for (int64_t k = 0; k < n; ++k) {
        double px, py;
        sample_position_of(S, mode, grid, nS, seed, ids[k], px, py);
        State st;
        camera_ray(S.cam, px, py, st);
        OracleRay tmp;
        vol_trace(S, VD, st, nbins, nu, rays ? rays[k] : tmp, spectra ? spectra + k * nbins : nullptr);
    }
}

double oracle_vol_density(const OracleVol* V, double r, double z, double phi) { return Vol{*V}.density(r, z, phi); }
double oracle_vol_temperature(const OracleVol* V, double r, double z) { return Vol{*V}.temperature(r, z); }
double oracle_vol_kappa_abs(const OracleVol* V, double nu, double rho, double T) {
    return Vol{*V}.kappa_abs(nu, rho, T);
}
double oracle_vol_kappa_es(const OracleVol* V, double rho, double T) { return Vol{*V}.kappa_es(rho, T); }
double oracle_simplex(const int* perm, double x, double y, double z) { return Vol::simplex_noise(perm, x, y, z); }
void oracle_vol_velocity(const OracleVol* V, double r, double theta, double* u) {
    gas_velocity(Vol{*V}, r, theta, u[0], u[1], u[2]);
}
void oracle_rt_steps(int n, const double* S, const double* dtau, double* J, double* T) {
    for (int i = 0; i < n; ++i) rt_step(*J, *T, S[i], dtau[i]);
}

}  // extern "C"
