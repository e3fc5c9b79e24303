// Kerr null-geodesic right-hand side in Boyer–Lindquist coordinates for the
// Hamiltonian H = g^{mu nu} p_mu p_nu / 2 with covariant momenta.
//
// Computes the same quantities as include/grrt/spacetime/kerr.h:31-108
// (dx^mu = g^{mu nu} p_nu, dp_r = -dH/dr, dp_theta = -dH/dtheta,
// dp_t = dp_phi = 0), rearranged with exact Kerr identities so that one
// reciprocal replaces the reference's seventeen divisions:
//   * g_tt g_phph - g_tph^2 = -Delta sin^2(theta)   (t-phi block determinant)
//   * dp_r and dp_theta regrouped into sums of squares of
//       u = dt - a s^2 dphi   and   v = a dt - (Sigma + a^2 s^2) dphi,
//     which also removes 1/Delta from the g_rr terms.
// The pole guard |sin theta| >= 1e-10 of kerr.h:41 is kept.
#pragma once
#include "device_math.hpp"

namespace grrt_amd {

struct KerrK {
    double M, a, a2, twoM;
};

// Five evolving components of the state (r, theta, phi, p_r, p_theta); p_t
// and p_phi are conserved exactly and carried as P0, P3.
struct Vec5 {
    double v[5];
};

// Returns (dr, dtheta, dphi, dp_r, dp_theta); dt is written to dt_out.
// With R = 1/(Sigma Delta s^2):
//   dt   = -[(r^2+a^2) Sigma p_t + 2Mar (a s^2 p_t + p_phi)] s^2 R
//   dphi =  [(Sigma - 2Mr) p_phi - 2Mar s^2 p_t] R
//   dr   = Delta p_r / Sigma,   dtheta = p_theta / Sigma
//   dp_r = [M(Sigma - 2r^2) u^2 + (r Delta - Sigma (r-M)) p_r^2 + r p_th^2] / Sigma^2 + r s^2 dphi^2
//   dp_th = s c [ (2Mr v^2 - a^2 (Delta p_r^2 + p_th^2)) / Sigma^2 + Delta dphi^2 ]
// FP64 operation tally (add/sub/mul = 1, fma = 2, reciprocal = 1):
//   sincos 36, trig products 3, Sigma/Delta 5, reciprocal block 5,
//   dt 10, dphi 8, dr/dtheta 3, u/v 8, dp_r 20, dp_theta 13  => 111.
// kerr_rhs_sc takes sin and cos of y.v[1] from the caller (sincos_cephes of
// that same angle), for callers that need them as well.
__device__ __forceinline__ Vec5 kerr_rhs_sc(const KerrK& K, double P0, double P3, const Vec5& y,
                                             double& dt_out, double s, double c) {
    const double r = y.v[0], pr = y.v[3], pth = y.v[4];
    // Pole guard |sin theta| >= 1e-10 (kerr.h:41); a sin of exactly -0 keeps its sign.
    s = __builtin_copysign(__builtin_fmax(__builtin_fabs(s), 1e-10), s);
    const double s2 = s * s;
    const double c2 = c * c;
    const double sc = s * c;                       // sin(2 theta) / 2
    const double r2 = r * r;
    const double Sig = __builtin_fma(K.a2, c2, r2);
    const double Del = __builtin_fma(r, r - K.twoM, K.a2);
    const double SD = Sig * Del;
    const double R = rcp(SD * s2);                 // 1 / (Sigma Delta s^2)
    const double iS = (Del * s2) * R;              // 1 / Sigma
    const double iS2 = iS * iS;
    const double twoMr = K.twoM * r;
    const double twoMar = twoMr * K.a;
    const double as2 = K.a * s2;
    const double nt = __builtin_fma(twoMar, __builtin_fma(as2, P0, P3), ((r2 + K.a2) * Sig) * P0);
    const double dt = -(nt * s2) * R;
    const double dph = __builtin_fma(-(twoMar * s2), P0, (Sig - twoMr) * P3) * R;
    Vec5 d;
    d.v[0] = (Del * iS) * pr;                      // g^rr p_r
    d.v[1] = iS * pth;                             // g^thth p_theta
    d.v[2] = dph;
    const double u = __builtin_fma(-as2, dph, dt);
    const double v = __builtin_fma(K.a, dt, -(__builtin_fma(K.a2, s2, Sig) * dph));
    const double pr2 = pr * pr;
    const double pth2 = pth * pth;
    const double dph2 = dph * dph;
    const double cu = K.M * __builtin_fma(-2.0, r2, Sig);
    const double cr = __builtin_fma(r, Del, -(Sig * (r - K.M)));
    const double br = __builtin_fma(cu, u * u, __builtin_fma(cr, pr2, r * pth2));
    d.v[3] = __builtin_fma(br, iS2, (r * s2) * dph2);
    const double bt = __builtin_fma(twoMr, v * v, -K.a2 * __builtin_fma(Del, pr2, pth2));
    d.v[4] = sc * __builtin_fma(bt, iS2, Del * dph2);
    dt_out = dt;
    return d;
}

__device__ __forceinline__ Vec5 kerr_rhs(const KerrK& K, double P0, double P3, const Vec5& y,
                                          double& dt_out) {
    double s, c;
    sincos_cephes(y.v[1], s, c);
    return kerr_rhs_sc(K, P0, P3, y, dt_out, s, c);
}

}  // namespace grrt_amd
