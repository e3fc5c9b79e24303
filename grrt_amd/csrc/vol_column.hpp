// One vertical column of the volumetric disk (src/volumetric_disk.cpp:505-777):
// Eddington T(tau) over a uniform column, then the hydrostatic density
// profile with gas and flux-limited radiation pressure by classic RK4 with a
// 1e-15 floor, the column top raised by H while the density at the top
// stays above 1e-10 (up to 20 H), at most eight passes.
//
// The node-level steps are written once here and run on host threads
// (volumetric_host.cpp) and on the GPU (vol_columns.hip, one wavefront per
// column, the nodes over its lanes and the two running sums on lane 0).
// The columns are radiation dominated: T is constant along a column to
// ~1e-7, so the pressure gradients are differences of nearly equal numbers
// and the profile is a rounding-sensitive realisation of the floored map.
// Both builds therefore use only operations that round the same way on both
// sides: IEEE double +, -, *, / and sqrt, no fused multiply-add (contraction
// is off in every function here), the fourth root as sqrt(sqrt(x)) and a
// log10 of this file's own.  Given the same opacity tables, the two builds
// are equal bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include "opacity_math.hpp"

namespace grrt_amd {

constexpr int kColMaxZ = 128;   // nodes per column (the tables' n_z)

// The (rho, T) tables a column reads: Rosseland mean and electron
// scattering opacity and mean molecular weight, rho-major, with their
// log10 ranges.
struct ColumnTables {
    const double* ross;
    const double* es;
    const double* mu;
    int n_rho, n_T;
    double lr0, lr1, lt0, lt1;
};

// One column's radial inputs: scale height, effective temperature,
// midplane density (0: no column) and the vertical epicyclic frequency
// squared (already the ISCO value inside the ISCO).
struct ColumnIn {
    double H, Te, rm, Omz2;
};

// log10 of a positive finite x: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, by its series
// through s^21 (truncation below 3e-17 relative), then
// log10 x = (e ln 2 + ln m) / ln 10 with ln 2 in two parts.
__host__ __device__ inline double col_log10(double x) {
#pragma clang fp contract(off)
    int e = 0;
    double m = frexp(x, &e);                       // m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    const double lnm = 2.0 * s + 2.0 * s * (s2 * p);
    constexpr double kLn2Hi = 6.93147180369123816490e-01, kLn2Lo = 1.90821492927058770002e-10;
    constexpr double kInvLn10 = 0.43429448190325182765;
    const double de = static_cast<double>(e);
    return ((de * kLn2Hi + lnm) + de * kLn2Lo) * kInvLn10;
}

// Bilinear interpolation in (log10 rho, log10 T) over a (rho, T) table,
// arguments clamped to the table (src/opacity.cpp:307-346).
__host__ __device__ inline double col_lookup(const double* lut, const ColumnTables& t, double rho, double T) {
#pragma clang fp contract(off)
    rho = rho < 1e-18 ? 1e-18 : (rho > 1e-6 ? 1e-6 : rho);
    T = T < 3000.0 ? 3000.0 : (T > 1e8 ? 1e8 : T);
    double tr = (col_log10(rho) - t.lr0) / (t.lr1 - t.lr0) * (t.n_rho - 1);
    tr = tr < 0.0 ? 0.0 : (tr > t.n_rho - 2 ? static_cast<double>(t.n_rho - 2) : tr);
    double tt = (col_log10(T) - t.lt0) / (t.lt1 - t.lt0) * (t.n_T - 1);
    tt = tt < 0.0 ? 0.0 : (tt > t.n_T - 2 ? static_cast<double>(t.n_T - 2) : tt);
    const int ir = static_cast<int>(tr), it = static_cast<int>(tt);
    const double fr = tr - ir, ft = tt - it;
    const int n = t.n_T;
    return lut[ir * n + it] * (1 - fr) * (1 - ft) + lut[ir * n + it + 1] * (1 - fr) * ft +
           lut[(ir + 1) * n + it] * fr * (1 - ft) + lut[(ir + 1) * n + it + 1] * fr * ft;
}

__host__ __device__ inline double col_qroot(double x) { return sqrt(sqrt(x)); }

// Levermore & Pomraning (1981) flux limiter and Eddington factor.
__host__ __device__ inline double col_eddington(double R) {
#pragma clang fp contract(off)
    const double lam = (2.0 + R) / (6.0 + 3.0 * R + R * R);
    return lam + lam * lam * R * R;
}

// Density scale of the column: tau_mid over the opacity at 1e-10 g/cm^3.
__host__ __device__ inline double col_rho_ref(const ColumnTables& t, const ColumnIn& c, double tau_mid) {
#pragma clang fp contract(off)
    double kref = col_lookup(t.ross, t, 1e-10, c.Te) + col_lookup(t.es, t, 1e-10, c.Te);
    kref = kref < 1.0 ? 1.0 : kref;
    const double r = tau_mid / (kref * 3.0 * c.H);
    return r < 1e-18 ? 1e-18 : (r > 1e-6 ? 1e-6 : r);
}

// Half the optical-depth increment of node z's layer over a column of
// uniform density rho_ref and temperature T_mid: tau[z] = tau[z+1] + this.
__host__ __device__ inline double col_tau_step(const ColumnTables& t, double rho_ref, double T_mid, double dz) {
#pragma clang fp contract(off)
    const double rh = 1.0 * rho_ref, rn = 1.0 * rho_ref;
    const double kh = col_lookup(t.ross, t, rh, T_mid) + col_lookup(t.es, t, rh, T_mid);
    const double kn = col_lookup(t.ross, t, rn, T_mid) + col_lookup(t.es, t, rn, T_mid);
    return 0.5 * (kh * rh + kn * rn) * dz;
}

// Eddington temperature at optical depth tau.
__host__ __device__ inline double col_T(double Te4, double tau) {
#pragma clang fp contract(off)
    const double v = 0.75 * Te4 * (tau + 2.0 / 3.0);
    return col_qroot(v > 0.0 ? v : 0.0);
}

__host__ __device__ inline double col_E(double T) {
#pragma clang fp contract(off)
    return cgs::a_rad * T * T * T * T;
}

__host__ __device__ inline double col_mu(const ColumnTables& t, double rho_ref, double T) {
    const double m = col_lookup(t.mu, t, 1.0 * rho_ref, T);
    const bool finite = m == m && m - m == 0.0;   // not NaN, not infinite
    return (m <= 0.0 || !finite) ? 0.6 : m;
}

// Eddington factor at node z from the energy-density gradient.
__host__ __device__ inline double col_f(const ColumnTables& t, const double* E, const double* T, int z, int NZ,
                                        double dz, double rho_ref) {
#pragma clang fp contract(off)
    double dE = 0.0;
    if (z == NZ - 1) dE = (E[z] - E[z - 1]) / dz;
    else if (z > 0) dE = (E[z + 1] - E[z - 1]) / (2.0 * dz);
    const double rc = 1.0 * rho_ref;
    const double den = col_lookup(t.ross, t, rc, T[z]) * rc * E[z];
    return col_eddington(den < 1e-30 ? 1e30 : fabs(dE) / den);
}

__host__ __device__ inline double col_cs2(double T, double mu) {
#pragma clang fp contract(off)
    return cgs::k_B * T / (mu * cgs::m_p);
}

// Gradients of c_s^2 and f E at node z (0 at the midplane, one-sided at
// the top, central elsewhere).
__host__ __device__ inline void col_grads(const double* T, const double* mu, const double* f, const double* E, int z,
                                          int NZ, double dz, double& dcs2, double& dfE) {
#pragma clang fp contract(off)
    if (z == 0) {
        dcs2 = 0.0;
        dfE = 0.0;
    } else if (z == NZ - 1) {
        dcs2 = (col_cs2(T[z], mu[z]) - col_cs2(T[z - 1], mu[z - 1])) / dz;
        dfE = (f[z] * E[z] - f[z - 1] * E[z - 1]) / dz;
    } else {
        dcs2 = (col_cs2(T[z + 1], mu[z + 1]) - col_cs2(T[z - 1], mu[z - 1])) / (2.0 * dz);
        dfE = (f[z + 1] * E[z + 1] - f[z - 1] * E[z - 1]) / (2.0 * dz);
    }
}

// d(rho/rho_ref)/dz = [-rho Omz^2 z - rho d(cs^2)/dz - d(f E)/dz / rho_ref] / cs^2,
// geometric z with CGS thermodynamics converted by c^2; the coefficients
// are linear between nodes.  Everything but rho is fixed by the position,
// so the right-hand side splits into the position's terms (col_rhs_terms:
// Omz^2 z, d(cs^2)/dz / c^2, d(f E)/dz / (rho_ref c^2) and cs^2 / c^2, or a
// zero right-hand side where cs^2 < 1e-30) and the density's part
// (col_rhs_apply), which repeats the same operations on the same values.
struct RhsTerms {
    double a, b, c, d;   // d = 0: the right-hand side is 0
};

__host__ __device__ inline RhsTerms col_rhs_terms(const double* T, const double* mu, const double* dcs2,
                                                  const double* dfE, int NZ, double dz, double Omz2, double rho_ref,
                                                  double ze) {
#pragma clang fp contract(off)
    const double c2 = cgs::c * cgs::c;
    const double zf = ze / dz;
    int idx = static_cast<int>(zf);
    idx = idx < 0 ? 0 : (idx > NZ - 2 ? NZ - 2 : idx);
    const double t = zf - idx;
    const double cs = cgs::k_B * ((1.0 - t) * T[idx] + t * T[idx + 1]) /
                      (((1.0 - t) * mu[idx] + t * mu[idx + 1]) * cgs::m_p);
    const double dc = (1.0 - t) * dcs2[idx] + t * dcs2[idx + 1];
    const double df = (1.0 - t) * dfE[idx] + t * dfE[idx + 1];
    if (cs < 1e-30) return RhsTerms{0.0, 0.0, 0.0, 0.0};
    const double cc = cs / c2;
    return RhsTerms{Omz2 * ze, dc / c2, df / (rho_ref * c2), cc < 1e-30 ? 1e-30 : cc};
}

__host__ __device__ inline double col_rhs_apply(const RhsTerms& q, double re) {
#pragma clang fp contract(off)
    if (q.d == 0.0) return 0.0;
    return (-re * q.a - re * q.b - q.c) / q.d;
}

// Positions of RK4 step z: the node, the midpoint (stages 2 and 3) and the
// end, as the step forms them.
__host__ __device__ inline void col_step_terms(const double* T, const double* mu, const double* dcs2,
                                               const double* dfE, int NZ, double dz, double Omz2, double rho_ref,
                                               int z, RhsTerms& q0, RhsTerms& qh, RhsTerms& q1) {
#pragma clang fp contract(off)
    const double zh = z * dz;
    q0 = col_rhs_terms(T, mu, dcs2, dfE, NZ, dz, Omz2, rho_ref, zh);
    qh = col_rhs_terms(T, mu, dcs2, dfE, NZ, dz, Omz2, rho_ref, zh + 0.5 * dz);
    q1 = col_rhs_terms(T, mu, dcs2, dfE, NZ, dz, Omz2, rho_ref, zh + dz);
}

// The density profile by classic RK4 from 1 at the midplane, stage values
// and results kept above the floor; q0, qh, q1: each step's position terms.
__host__ __device__ inline void col_rk4(const RhsTerms* q0, const RhsTerms* qh, const RhsTerms* q1, int NZ,
                                        double dz, double* rho) {
#pragma clang fp contract(off)
    constexpr double floor_rho = 1e-15;
    auto fl = [](double v) { return v < floor_rho ? floor_rho : v; };
    rho[0] = 1.0;
    for (int z = 0; z < NZ - 1; ++z) {
        const double rh = rho[z];
        const double k1 = dz * col_rhs_apply(q0[z], rh);
        const double k2 = dz * col_rhs_apply(qh[z], fl(rh + 0.5 * k1));
        const double k3 = dz * col_rhs_apply(qh[z], fl(rh + 0.5 * k2));
        const double k4 = dz * col_rhs_apply(q1[z], fl(rh + k3));
        rho[z + 1] = fl(rh + (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0);
    }
}

// After a pass: raise the column top by H (up to 20 H) while the density
// at the top is above 1e-10; returns true when it did, so another pass
// runs.  A pass that leaves the top where it is ends the column.  (The
// reference then compares the profile with the previous pass's and, as the
// change is not below 1e-3 after a raise, runs the pass once more: with the
// same top and the same uniform start column that pass repeats this one
// exactly and stops on a zero change, so the profile is this pass's.)
__host__ __device__ inline bool col_raise_top(const double* rho, int NZ, double H, double& zmax) {
#pragma clang fp contract(off)
    if (rho[NZ - 1] > 1e-10 && zmax < 20.0 * H) {
        const double zn = zmax + H;
        zmax = zn < 20.0 * H ? zn : 20.0 * H;
        return true;
    }
    return false;
}

// Builds the device columns: n_r columns, one wavefront each.  in: per
// column inputs; outputs: z_max[n_r], rho[n_r][n_z], T[n_r][n_z].
hipError_t launch_vol_columns(const ColumnTables& t, const ColumnIn* in, int n_r, int n_z, double tau_mid,
                              double* z_max, double* rho, double* T, hipStream_t st);

}  // namespace grrt_amd
