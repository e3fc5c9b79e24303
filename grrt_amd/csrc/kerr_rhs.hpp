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
// FP64 operation tally (add/sub/mul = 1, fma = 2, reciprocal = 1):
//   sincos 36, trig products 6, Sigma/Delta 5, reciprocal block 9,
//   dt 9, dphi 8, dr/dtheta 3, u/v 7, dp_r 22, dp_theta 15  => 120.
__device__ __forceinline__ Vec5 kerr_rhs(const KerrK& K, double P0, double P3, const Vec5& y,
                                          double& dt_out) {
    const double r = y.v[0], th = y.v[1], pr = y.v[3], pth = y.v[4];
    double s, c;
    sincos_cephes(th, s, c);
    if (__builtin_fabs(s) < 1e-10) s = (s >= 0.0) ? 1e-10 : -1e-10;
    const double s2 = s * s;
    const double c2 = c * c;
    const double sc = s * c;                       // sin(2 theta) / 2
    const double r2 = r * r;
    const double Sig = __builtin_fma(K.a2, c2, r2);
    const double Del = __builtin_fma(r, r - K.twoM, K.a2);
    const double Ds2 = Del * s2;
    const double SD = Sig * Del;
    const double R = rcp(SD * s2);                 // 1 / (Sigma Delta s^2)
    const double iS = Ds2 * R;                     // 1 / Sigma
    const double iD = (Sig * s2) * R;              // 1 / Delta
    const double is2 = SD * R;                     // 1 / s^2
    const double iS2 = iS * iS;
    const double twoMr = K.twoM * r;
    const double A = twoMr * K.a * iS;             // 2 M a r / Sigma
    // dt = g^tt P0 + g^tph P3,  dphi = g^tph P0 + g^phph P3
    const double dt = -iD * __builtin_fma(r2 + K.a2, P0, A * __builtin_fma(K.a * s2, P0, P3));
    const double dph = iD * __builtin_fma(-A, P0, __builtin_fma(-twoMr, iS, 1.0) * P3 * is2);
    Vec5 d;
    d.v[0] = (Del * iS) * pr;                      // g^rr p_r
    d.v[1] = iS * pth;                             // g^thth p_theta
    d.v[2] = dph;
    const double u = __builtin_fma(-K.a * s2, dph, dt);
    const double v = __builtin_fma(K.a, dt, -(__builtin_fma(K.a2, s2, Sig) * dph));
    const double pr2 = pr * pr;
    const double pth2 = pth * pth;
    const double dph2 = dph * dph;
    // dp_r = [M(Sigma - 2r^2) u^2 + (r Delta - Sigma (r - M)) p_r^2 + r p_th^2] / Sigma^2
    //        + r s^2 dphi^2
    const double cu = K.M * (Sig - 2.0 * r2);
    const double cr = __builtin_fma(r, Del, -Sig * (r - K.M));
    const double br = __builtin_fma(cu, u * u, __builtin_fma(cr, pr2, r * pth2));
    d.v[3] = __builtin_fma(br, iS2, r * s2 * dph2);
    // dp_theta = s c [ (2 M r v^2 - a^2 (Delta p_r^2 + p_th^2)) / Sigma^2 + Delta dphi^2 ]
    const double bt = __builtin_fma(twoMr, v * v, -K.a2 * __builtin_fma(Del, pr2, pth2));
    d.v[4] = sc * __builtin_fma(bt, iS2, Del * dph2);
    dt_out = dt;
    return d;
}

}  // namespace grrt_amd
