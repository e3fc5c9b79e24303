// Volumetric-disk tracer for gfx950: one lane per ray.
//
// Per-ray semantics follow the reference:
//   src/geodesic_tracer.cpp:23-181    outer loop, step clamp near the disk,
//                                     entry detection (RGB mode)
//   src/geodesic_tracer.cpp:183-337   RGB ray-march (450/550/650 nm)
//   src/geodesic_tracer.cpp:495-603   spectral outer loop
//   src/geodesic_tracer.cpp:605-746   N-bin spectral ray-march
// The geodesic is integrated in FP64 (the same Dormand–Prince step as the
// thin-disk kernel; classic RK4 inside the ray-march).  Radiative transfer
// runs in the template type RT: double for the all-FP64 path, float for the
// mixed-precision path, where each bin's intensity is accumulated with
// Neumaier compensated summation.  On request (far_f32) the mixed path also
// takes the outer Dormand–Prince steps in FP32 on far-field segments: outside
// r = 15 M, on the way out, away from the disk's step-clamp zone and not
// right after a theta turning point, with the error tolerance raised to
// 1e-6, the smallest the FP32 error estimate resolves.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_math.hpp"
#include "kerr_rhs.hpp"
#include "camera_device.hpp"
#include "star_device.hpp"
#include "trace_params.hpp"
#include "volumetric_device.hpp"

namespace grrt_amd {

namespace {

constexpr double A21 = 1.0 / 5.0;
constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                 A54 = -212.0 / 729.0;
constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                 A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
constexpr double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                 B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
constexpr double E1 = 71.0 / 57600.0, E3 = -71.0 / 16695.0, E4 = 71.0 / 1920.0,
                 E5 = -17253.0 / 339200.0, E6 = 22.0 / 525.0, E7 = -1.0 / 40.0;

struct GState {
    Vec5 y;         // r, theta, phi, p_r, p_theta
    double P0, P3;  // conserved p_t, p_phi
};

// 0.9 x^(1/5) of the step controller (src/rk4.cpp:113-125, pow in FP64):
// an FP32 estimate from the hardware log2/exp2, then two Newton steps on
// y^5 = x in FP64, which converge quadratically from ~1e-7 to rounding.
// The ray-march samples the volume at positions that follow the outer step
// lengths, so an FP32-rounded step length (~1e-7 of a step) would move its
// samples and show in the intensities at ~1e-5.
__device__ __forceinline__ double step_factor_v(double x) {
    const float lx = __builtin_amdgcn_logf(static_cast<float>(x));
    double y = static_cast<double>(__builtin_amdgcn_exp2f(0.2f * lx));
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double y2 = y * y, y4 = y2 * y2;
        y = __builtin_fma(-(__builtin_fma(y4, y, -x)), rcp1(5.0 * y4), y);
    }
    return 0.9 * y;
}

// One adaptive Dormand–Prince step (src/rk4.cpp:87-136); returns next dl,
// counts attempts.  k1 is the right-hand side at s on entry and at the new
// state on return: the accepted stage k7 is evaluated at y5, which becomes
// the state (first same as last), so the next step and the near-disk step
// clamp reuse it instead of evaluating it again.  sn, cn: sin and cos of the
// new state's theta, from the same k7 evaluation.
__device__ double adaptive_dp45(const KerrK& K, GState& s, Vec5& k1, double dl, double tol, int& attempts,
                                double& sn, double& cn) {
    double dtmp;
    for (int retry = 0; retry <= 20; ++retry) {
        const double h = dl;
        Vec5 st, k2, k3, k4, k5, k6, k7, y5, e;
        for (int i = 0; i < 5; ++i) st.v[i] = __builtin_fma(k1.v[i], h * A21, s.y.v[i]);
        k2 = kerr_rhs(K, s.P0, s.P3, st, dtmp);
        for (int i = 0; i < 5; ++i)
            st.v[i] = __builtin_fma(k2.v[i], h * A32, __builtin_fma(k1.v[i], h * A31, s.y.v[i]));
        k3 = kerr_rhs(K, s.P0, s.P3, st, dtmp);
        for (int i = 0; i < 5; ++i)
            st.v[i] = __builtin_fma(k3.v[i], h * A43, __builtin_fma(k2.v[i], h * A42, __builtin_fma(k1.v[i], h * A41, s.y.v[i])));
        k4 = kerr_rhs(K, s.P0, s.P3, st, dtmp);
        for (int i = 0; i < 5; ++i)
            st.v[i] = __builtin_fma(k4.v[i], h * A54, __builtin_fma(k3.v[i], h * A53,
                          __builtin_fma(k2.v[i], h * A52, __builtin_fma(k1.v[i], h * A51, s.y.v[i]))));
        k5 = kerr_rhs(K, s.P0, s.P3, st, dtmp);
        for (int i = 0; i < 5; ++i)
            st.v[i] = __builtin_fma(k5.v[i], h * A65, __builtin_fma(k4.v[i], h * A64,
                          __builtin_fma(k3.v[i], h * A63, __builtin_fma(k2.v[i], h * A62,
                              __builtin_fma(k1.v[i], h * A61, s.y.v[i])))));
        k6 = kerr_rhs(K, s.P0, s.P3, st, dtmp);
        for (int i = 0; i < 5; ++i) {
            y5.v[i] = __builtin_fma(k6.v[i], h * B6, __builtin_fma(k5.v[i], h * B5,
                          __builtin_fma(k4.v[i], h * B4, __builtin_fma(k3.v[i], h * B3,
                              __builtin_fma(k1.v[i], h * B1, s.y.v[i])))));
            e.v[i] = __builtin_fma(k6.v[i], h * E6, __builtin_fma(k5.v[i], h * E5,
                         __builtin_fma(k4.v[i], h * E4, __builtin_fma(k3.v[i], h * E3, k1.v[i] * (h * E1)))));
        }
        double s7, c7;
        sincos_cephes(y5.v[1], s7, c7);
        k7 = kerr_rhs_sc(K, s.P0, s.P3, y5, dtmp, s7, c7);
        double err = 0.0;
        for (int i = 0; i < 5; ++i)
            err = __builtin_fmax(err, __builtin_fabs(__builtin_fma(k7.v[i], h * E7, e.v[i])) *
                                          rcp1(__builtin_fabs(y5.v[i]) + 1e-10));
        ++attempts;
        if (retry == 20) {
            s.y = y5;
            k1 = k7;
            sn = s7;
            cn = c7;
            return h;
        }
        if (err <= tol) {
            double f = (err > 1e-30) ? step_factor_v(tol / err) : 5.0;
            f = sclamp(f, 1.0, 5.0);
            s.y = y5;
            k1 = k7;
            sn = s7;
            cn = c7;
            return smin(h * f, 5.0 * smax(__builtin_fabs(y5.v[0]), 1.0));
        }
        dl = h * sclamp(step_factor_v(tol / err), 0.2, 1.0);
        if (dl < 1e-6) {
            s.y = y5;
            k1 = k7;
            sn = s7;
            cn = c7;
            return smin(1e-6, 5.0 * smax(__builtin_fabs(y5.v[0]), 1.0));
        }
    }
    return dl;   // not reached
}

// Classic RK4 step (include/grrt/geodesic/rk4.h:34-52).
__device__ GState rk4_step(const KerrK& K, const GState& s, double h);

// Adaptive step-doubling RK4 (src/rk4.cpp:33-80): a step of dl against two
// of dl/2, error over r, theta, phi, p_r, p_theta; accept (doubling dl
// below 1% of tol) or halve, keeping the half steps once dl < 1e-6; after
// 20 halvings the two half steps unchecked.  Returns the next dl.
__device__ double adaptive_rk4_doubling(const KerrK& K, GState& s, double dl, double tol, int& attempts) {
    for (int retry = 0; retry < 20; ++retry) {
        const GState full = rk4_step(K, s, dl);
        const GState mid = rk4_step(K, s, dl * 0.5);
        const GState half = rk4_step(K, mid, dl * 0.5);
        ++attempts;
        double err = 0.0;
        for (int i = 0; i < 5; ++i)
            err = __builtin_fmax(err, __builtin_fabs(full.y.v[i] - half.y.v[i]) *
                                          rcp1(__builtin_fabs(half.y.v[i]) + 1e-10));
        if (err <= tol) {
            s = half;
            return smin(err < tol * 0.01 ? dl * 2.0 : dl, 5.0 * smax(__builtin_fabs(half.y.v[0]), 1.0));
        }
        dl *= 0.5;
        if (dl < 1e-6) {
            s = half;
            return smin(1e-6, 5.0 * smax(__builtin_fabs(half.y.v[0]), 1.0));
        }
    }
    const GState mid = rk4_step(K, s, dl * 0.5);
    s = rk4_step(K, mid, dl * 0.5);
    ++attempts;
    return dl;
}

// ---- FP32 far-field step ----
constexpr double kFarField = 15.0;       // in units of M
constexpr float kFarTol = 1e-6f;

struct Vec5f {
    float v[5];
};

// Only the FP32-transfer instantiation carries the FP32 far-field stepper.
template <class RT>
constexpr bool kFarFieldF32 = sizeof(RT) == sizeof(float);

// The Kerr right-hand side of kerr_rhs.hpp in FP32 (same regrouped
// expressions; library sincos).
__device__ __forceinline__ Vec5f kerr_rhs_f32(float M, float a, float a2, float twoM, float P0, float P3,
                                              const Vec5f& y) {
    const float r = y.v[0], th = y.v[1], pr = y.v[3], pth = y.v[4];
    float s, c;
    sincosf(th, &s, &c);
    s = copysignf(fmaxf(fabsf(s), 1e-10f), s);
    const float s2 = s * s, c2 = c * c, sc = s * c, r2 = r * r;
    const float Sig = fmaf(a2, c2, r2);
    const float Del = fmaf(r, r - twoM, a2);
    const float R = 1.0f / (Sig * Del * s2);
    const float iS = (Del * s2) * R;
    const float iS2 = iS * iS;
    const float twoMr = twoM * r, twoMar = twoMr * a, as2 = a * s2;
    const float nt = fmaf(twoMar, fmaf(as2, P0, P3), ((r2 + a2) * Sig) * P0);
    const float dt = -(nt * s2) * R;
    const float dph = fmaf(-(twoMar * s2), P0, (Sig - twoMr) * P3) * R;
    Vec5f d;
    d.v[0] = (Del * iS) * pr;
    d.v[1] = iS * pth;
    d.v[2] = dph;
    const float u = fmaf(-as2, dph, dt);
    const float v = fmaf(a, dt, -(fmaf(a2, s2, Sig) * dph));
    const float pr2 = pr * pr, pth2 = pth * pth, dph2 = dph * dph;
    const float cu = M * fmaf(-2.0f, r2, Sig);
    const float cr = fmaf(r, Del, -(Sig * (r - M)));
    d.v[3] = fmaf(fmaf(cu, u * u, fmaf(cr, pr2, r * pth2)), iS2, (r * s2) * dph2);
    d.v[4] = sc * fmaf(fmaf(twoMr, v * v, -a2 * fmaf(Del, pr2, pth2)), iS2, Del * dph2);
    return d;
}

// adaptive_dp45 in FP32: the same tableau, error norm and controller, the
// state converted from and back to FP64 around the step.
__device__ double adaptive_dp45_f32(const KerrK& K, GState& s, double dl, int& attempts) {
    const float M = static_cast<float>(K.M), a = static_cast<float>(K.a), a2 = static_cast<float>(K.a2),
                twoM = static_cast<float>(K.twoM);
    const float P0 = static_cast<float>(s.P0), P3 = static_cast<float>(s.P3);
    Vec5f y;
    for (int i = 0; i < 5; ++i) y.v[i] = static_cast<float>(s.y.v[i]);
    const Vec5f k1 = kerr_rhs_f32(M, a, a2, twoM, P0, P3, y);
    float h = static_cast<float>(dl);
    for (int retry = 0; retry <= 20; ++retry) {
        Vec5f st, k2, k3, k4, k5, k6, k7, y5, e;
        for (int i = 0; i < 5; ++i) st.v[i] = fmaf(k1.v[i], h * float(A21), y.v[i]);
        k2 = kerr_rhs_f32(M, a, a2, twoM, P0, P3, st);
        for (int i = 0; i < 5; ++i) st.v[i] = fmaf(k2.v[i], h * float(A32), fmaf(k1.v[i], h * float(A31), y.v[i]));
        k3 = kerr_rhs_f32(M, a, a2, twoM, P0, P3, st);
        for (int i = 0; i < 5; ++i)
            st.v[i] = fmaf(k3.v[i], h * float(A43), fmaf(k2.v[i], h * float(A42), fmaf(k1.v[i], h * float(A41), y.v[i])));
        k4 = kerr_rhs_f32(M, a, a2, twoM, P0, P3, st);
        for (int i = 0; i < 5; ++i)
            st.v[i] = fmaf(k4.v[i], h * float(A54), fmaf(k3.v[i], h * float(A53),
                          fmaf(k2.v[i], h * float(A52), fmaf(k1.v[i], h * float(A51), y.v[i]))));
        k5 = kerr_rhs_f32(M, a, a2, twoM, P0, P3, st);
        for (int i = 0; i < 5; ++i)
            st.v[i] = fmaf(k5.v[i], h * float(A65), fmaf(k4.v[i], h * float(A64),
                          fmaf(k3.v[i], h * float(A63), fmaf(k2.v[i], h * float(A62),
                              fmaf(k1.v[i], h * float(A61), y.v[i])))));
        k6 = kerr_rhs_f32(M, a, a2, twoM, P0, P3, st);
        for (int i = 0; i < 5; ++i) {
            y5.v[i] = fmaf(k6.v[i], h * float(B6), fmaf(k5.v[i], h * float(B5),
                          fmaf(k4.v[i], h * float(B4), fmaf(k3.v[i], h * float(B3),
                              fmaf(k1.v[i], h * float(B1), y.v[i])))));
            e.v[i] = fmaf(k6.v[i], h * float(E6), fmaf(k5.v[i], h * float(E5),
                         fmaf(k4.v[i], h * float(E4), fmaf(k3.v[i], h * float(E3), k1.v[i] * (h * float(E1))))));
        }
        k7 = kerr_rhs_f32(M, a, a2, twoM, P0, P3, y5);
        float err = 0.0f;
        for (int i = 0; i < 5; ++i) err = fmaxf(err, fabsf(fmaf(k7.v[i], h * float(E7), e.v[i])) / (fabsf(y5.v[i]) + 1e-10f));
        ++attempts;
        bool done = retry == 20;
        double next = h;
        if (!done && err <= kFarTol) {
            float f = (err > 1e-30f) ? static_cast<float>(step_factor_v(kFarTol / err)) : 5.0f;
            f = fminf(fmaxf(f, 1.0f), 5.0f);
            next = smin(static_cast<double>(h) * f, 5.0 * smax(fabs(static_cast<double>(y5.v[0])), 1.0));
            done = true;
        } else if (!done) {
            h = h * fminf(fmaxf(static_cast<float>(step_factor_v(kFarTol / err)), 0.2f), 1.0f);
            if (h < 1e-6f) {
                next = smin(1e-6, 5.0 * smax(fabs(static_cast<double>(y5.v[0])), 1.0));
                done = true;
            }
        }
        if (done) {
            for (int i = 0; i < 5; ++i) s.y.v[i] = static_cast<double>(y5.v[i]);
            return next;
        }
    }
    return h;   // not reached
}

// Classic RK4 step (include/grrt/geodesic/rk4.h:34-52) from s, whose
// right-hand side k1 is known.
__device__ GState rk4_step_from(const KerrK& K, const GState& s, const Vec5& k1, double h) {
    double dt;
    GState t = s, o = s;
    for (int i = 0; i < 5; ++i) t.y.v[i] = s.y.v[i] + k1.v[i] * (h * 0.5);
    const Vec5 k2 = kerr_rhs(K, s.P0, s.P3, t.y, dt);
    for (int i = 0; i < 5; ++i) t.y.v[i] = s.y.v[i] + k2.v[i] * (h * 0.5);
    const Vec5 k3 = kerr_rhs(K, s.P0, s.P3, t.y, dt);
    for (int i = 0; i < 5; ++i) t.y.v[i] = s.y.v[i] + k3.v[i] * h;
    const Vec5 k4 = kerr_rhs(K, s.P0, s.P3, t.y, dt);
    for (int i = 0; i < 5; ++i)
        o.y.v[i] = s.y.v[i] + (k1.v[i] + k2.v[i] * 2.0 + k3.v[i] * 2.0 + k4.v[i]) * (h / 6.0);
    return o;
}

__device__ GState rk4_step(const KerrK& K, const GState& s, double h) {
    double dt;
    const Vec5 k1 = kerr_rhs(K, s.P0, s.P3, s.y, dt);
    return rk4_step_from(K, s, k1, h);
}

// Per-ray radiative-transfer accumulators: J in units of 2h/c^2 (with
// Neumaier compensation Jc when RT is float), transmission T and
// accumulated optical depth tau, addressed by slot j of this lane.  The
// kernel kinds:
//   kBinsRGB     three channels (450/550/650 nm) over a group of kRgbGroup
//                lanes per ray, channel c on lane c of the group;
//   kBinsGlobal  any number of spectral bins in device scratch laid out
//                [quantity][bin][ray], one lane per ray;
//   kBinsGroup   up to 64 * kVolBinsPerLane spectral bins over a group of G
//                lanes (G a power of two dividing the wavefront): lane b of
//                the group holds bins b, b + G, b + 2G, ... in registers.
// Every lane of a group integrates the same geodesic, so the group's lanes
// always take the same branches; only the bin loop is split, and the
// quantities that couple bins (the step-control bin's opacity, the
// all-bins-opaque exits, the record's colour) are exchanged inside the
// group.  A few rays per wavefront keep the rays of one wavefront from
// waiting on one another's long ray-marches (the RGB frame of the Kerr
// a=0.9 512^2 volumetric scene: 172 ms with one lane per ray, 63 ms with 8
// lanes, 58 ms with 16).
enum : int { kBinsRGB = 0, kBinsGlobal = 1, kBinsGroup = 2 };
constexpr int kRgbGroup = 16;
static_assert(kRgbGroup >= 3 && kRgbGroup <= 64 && (kRgbGroup & (kRgbGroup - 1)) == 0,
              "each RGB channel is slot 0 of its own lane");

template <class RT, int kStore>
struct Bins;

// Storage of a kernel kind: scratch for kBinsGlobal, lane groups otherwise.
template <class RT, int kKind>
using BinsOf = Bins<RT, kKind == kBinsGlobal ? kBinsGlobal : kBinsGroup>;

template <class RT>
struct Bins<RT, kBinsGlobal> {
    RT *Jp, *Jcp, *Tp, *taup;
    long long stride;
    int nb;
    __device__ Bins(const TraceParams& P, long long id, int) {
        RT* base = static_cast<RT*>(P.rt_scratch);
        const long long plane = static_cast<long long>(P.nbins) * P.n_rays;
        stride = P.n_rays;
        nb = P.nbins;
        Jp = base + id;
        Jcp = base + plane + id;
        Tp = base + 2 * plane + id;
        taup = base + 3 * plane + id;
        for (int c = 0; c < P.nbins; ++c) {
            Jp[c * stride] = RT(0);
            Jcp[c * stride] = RT(0);
            Tp[c * stride] = RT(1);
            taup[c * stride] = RT(0);
        }
    }
    __device__ int slots() const { return nb; }
    __device__ int bin(int j) const { return j; }
    __device__ bool lead() const { return true; }
    template <class V>
    __device__ V share(V v, int) const { return v; }
    __device__ bool all(bool pred) const { return pred; }
    __device__ RT& J(int j) { return Jp[j * stride]; }
    __device__ RT& Jc(int j) { return Jcp[j * stride]; }
    __device__ RT& T(int j) { return Tp[j * stride]; }
    __device__ RT& tau(int j) { return taup[j * stride]; }
};

template <class RT>
struct Bins<RT, kBinsGroup> {
    RT Jv[kVolBinsPerLane], Jcv[kVolBinsPerLane], Tv[kVolBinsPerLane], tauv[kVolBinsPerLane];
    int b, G, nb;
    unsigned long long gbits;   // the group's lanes in a wavefront-wide mask
    __device__ Bins(const TraceParams& P, long long, int group) : G(group), nb(P.nbins) {
        const int lane = static_cast<int>(threadIdx.x & 63u);
        b = lane & (G - 1);
        gbits = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << (lane - b);
#pragma unroll kVolBinsPerLane
        for (int j = 0; j < kVolBinsPerLane; ++j) { Jv[j] = RT(0); Jcv[j] = RT(0); Tv[j] = RT(1); tauv[j] = RT(0); }
    }
    static constexpr int slots() { return kVolBinsPerLane; }
    __device__ int bin(int j) const {
        const int c = j * G + b;
        return c < nb ? c : -1;
    }
    __device__ bool lead() const { return b == 0; }
    // The value held by the lane that owns bin c.
    template <class V>
    __device__ V share(V v, int c) const { return __shfl(v, c & (G - 1), G); }
    // pred on every lane of the group.
    __device__ bool all(bool pred) const { return (__ballot(!pred) & gbits) == 0ull; }
    __device__ RT& J(int j) { return Jv[j]; }
    __device__ RT& Jc(int j) { return Jcv[j]; }
    __device__ RT& T(int j) { return Tv[j]; }
    __device__ RT& tau(int j) { return tauv[j]; }
};

__device__ __forceinline__ double rcp_rt(double x) { return rcp(x); }
__device__ __forceinline__ float rcp_rt(float x) { return __builtin_amdgcn_rcpf(x); }

// Compensated addition (Neumaier 1974): s + c carries the exact running sum.
__device__ __forceinline__ void neumaier(float& s, float& c, float y) {
    const float t = s + y;
    c += (fabsf(s) >= fabsf(y)) ? (s - t) + y : (y - t) + s;
    s = t;
}
__device__ __forceinline__ void neumaier(double& s, double&, double y) { s += y; }

struct MarchOut {
    int emitted;      // at least one emitting sample
    int steps;        // RK4 steps taken
};

// Backward ray-march through the volume from `st`; advances `st` to the exit
// state.  RGB mode: channels at 450/550/650 nm, step control on 550 nm and
// an opacity exit; spectral mode: step control on bin N/2 and an exit once
// every bin is past tau = 10.  Each lane updates the bins of its slots.
// sth, cth: sin and cos of st's theta on entry, and of the exit state's on
// return (every step evaluates them for the sample height; the next step's
// first right-hand side takes them instead of evaluating them again).
template <class RT, int kKind>
__device__ MarchOut raymarch(const TraceParams& P, const VolDev& V, const VolRadial& R, const int* perm,
                             const double* nu_tab, const double* lnu_tab, const KerrK& K, GState& st,
                             BinsOf<RT, kKind>& b, double lnu_rgb, double& sth, double& cth) {
    constexpr bool kSpectral = kKind != kBinsRGB;
    const int med = kSpectral ? P.nbins / 2 : 1;
    const double ut_obs = 1.0 / sqrt(1.0 - 2.0 / P.view.r_obs);
    MarchOut mo{0, 0};
    double r = st.y.v[0];
    const double z0 = r * cth;
    const double H0 = vol_H(V, R, r);
    bool been_inside = vol_inside(V, R, r, z0);
    double ds = been_inside ? H0 / 16.0 : smin(__builtin_fabs(z0) / 8.0, H0 * 2.0);
    // Exit tests use the bins' running extreme values.
    RT ext = kSpectral ? RT(1e30) : RT(0);
#pragma unroll kVolBinsPerLane
    for (int j = 0; j < b.slots(); ++j) {
        if (b.bin(j) < 0) continue;
        ext = kSpectral ? (b.tau(j) < ext ? b.tau(j) : ext) : (b.T(j) > ext ? b.T(j) : ext);
    }
    for (int n = 0; n < 4096; ++n) {
        double dtmp;
        const GState ns = rk4_step_from(K, st, kerr_rhs_sc(K, st.P0, st.P3, st.y, dtmp, sth, cth), ds);
        ++mo.steps;
        r = ns.y.v[0];
        const double th = ns.y.v[1], ph = ns.y.v[2];
        sincos_cephes(th, sth, cth);
        const double z = r * cth;
        if (r < V.r_horizon || r > V.r_outer) { st = ns; break; }
        if (b.all(kSpectral ? (ext > RT(10)) : (ext < RT(1e-6)))) { st = ns; break; }
        const double H = vol_H(V, R, r);
        if (!vol_inside(V, R, r, z)) {
            const double zm = vol_zmax(V, R, r);
            if (been_inside && __builtin_fabs(z) > zm + 3.0 * H) { st = ns; break; }
            if (!been_inside) ds = smax(smin(__builtin_fabs(z) / 8.0, H * 2.0), H / 64.0);
            else ds = sclamp(H / 4.0, H / 64.0, H);
            st = ns;
            continue;
        }
        been_inside = true;
        const double rho = vol_density(V, R, perm, r, z, ph);
        const double Tmp = vol_temperature(V, R, r, __builtin_fabs(z));
        if (rho <= 0.0 || Tmp <= 0.0) { st = ns; continue; }
        double ut, ur, up;
        vol_velocity(V, r, th, ut, ur, up);
        const double pue = ns.P0 * ut + ns.y.v[3] * ur + ns.P3 * up;
        const double g = pue * rcp(ns.P0 * ut_obs);
        const double ag = __builtin_fabs(g);
        const double lag = log10(ag);
        const RT dsp = static_cast<RT>(__builtin_fabs(pue) * __builtin_fabs(ds));
        const OpCell<RT> cell = vol_op_cell<RT>(V, rho, Tmp);
        const RT kes = vol_kes<RT>(V, cell);
        const RT rhoT = static_cast<RT>(rho), hkT = static_cast<RT>(vol_hk_over_T(Tmp));
        RT kmed = RT(0);
        ext = kSpectral ? RT(1e30) : RT(0);
#pragma unroll kVolBinsPerLane
        for (int j = 0; j < b.slots(); ++j) {
            const int c = b.bin(j);
            if (c < 0) continue;
            const double nu_o = kSpectral ? nu_tab[c] : (c == 0 ? 2.997924e10 / 450e-7 : (c == 1 ? 2.997924e10 / 550e-7 : 2.997924e10 / 650e-7));
            const double lnu_o = kSpectral ? lnu_tab[c] : lnu_rgb;   // RGB: slot 0 only
            const RT nue = static_cast<RT>(ag * nu_o);
            const RT kabs = vol_kabs<RT>(V, cell, lag + lnu_o);
            if (c == med) kmed = kabs;
            const RT ktot = kabs + kes;
            const RT eps = (ktot > RT(0)) ? kabs * rcp_rt(ktot) : RT(1);
            const RT dtau = ktot * rhoT * dsp;
            const RT src = eps * vol_source(nue, hkT);   // S = eps B_nu(nu_e) / nu_e^3 in units of 2h/c^2
            const RT ex = exp(-dtau);
            RT& Tc = b.T(j);
            neumaier(b.J(j), b.Jc(j), Tc * src * (RT(1) - ex));
            Tc *= ex;
            if (kSpectral) {
                RT& tc = b.tau(j);
                tc += dtau;
                ext = tc < ext ? tc : ext;
            } else {
                ext = Tc > ext ? Tc : ext;
            }
        }
        kmed = b.share(kmed, med);
        const double alpha_tot = static_cast<double>((kmed + kes) * rhoT);
        const double ds_tau = alpha_tot > 0.0 ? 0.05 * rcp(alpha_tot) : ds * 2.0;
        const double ds_geo = 0.1 * smax(r - V.r_horizon, 0.5);
        ds = sclamp(smin(ds_tau, ds_geo), H / 64.0, H);
        st = ns;
        mo.emitted = 1;
    }
    return mo;
}

// Launch order -> ray index (row-major samples of the window): whole rows
// centre-out, the middle row first and then alternately the rows above and
// below it, so the rays that cross the disk, the longest ones, start first
// instead of holding the launch open at its end.  Within a row, claim
// position p takes ray (p * stride) mod row_rays (stride coprime to the row
// length, about an eighth of it), so the lane groups of one wavefront, which
// claim consecutive positions, trace rays spread across the row.  Long rays
// cluster in a few neighbouring pixels; with row-major claims several of
// them shared a wavefront and ran their divergent paths one after another.
__device__ __forceinline__ long long centre_out_ray(const ViewParams& V, long long q, long long stride) {
    const long long row_rays = static_cast<long long>(V.tw) * V.nS;
    const long long k = q / row_rays;
    const long long mid = V.th / 2, d = (k + 1) / 2;
    const long long row = (k & 1) ? mid - d : mid + d;
    return row * row_rays + ((q - k * row_rays) * stride) % row_rays;
}

}  // namespace

// Persistent volumetric tracer.  A ray is owned by a group of `group`
// lanes (one lane for kBinsGlobal); each group claims its rays one at a
// time from the device-wide queue (the window's rows centre-out) and takes
// the next one as soon as its ray ends, so a group never waits for the
// other rays of its wavefront and the longest rays (around the disk image)
// start first.  Per ray: the outer Dormand–Prince loop with the near-disk
// step clamp and volume entry detection, ray-marching on every entry
// (src/geodesic_tracer.cpp:23-181, :495-603).  A cancel request, read with
// each claim, stops further claims; claimed rays always finish.
//
// kWaves: waves per SIMD the register budget is cut for.  Two halve each
// wave's issue share while another wave runs beside it; one gives every
// wavefront a SIMD (and up to 512 registers) but hides no latency.
template <class RT, int kKind, int kWaves>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWaves))) void trace_vol(const TraceParams* Pg,
                                                                                       int group) {
    constexpr bool kSpectral = kKind != kBinsRGB;
    // LDS: the noise permutation and the radial tables (scale height, z_max,
    // midplane density), which the step clamp, the entry test and every
    // ray-march sample read several times.
    __shared__ int s_perm[512];
    __shared__ double s_rad[3 * kVolMaxR];
    // spectral bins of the lane-group kernel (at most 64 * kVolBinsPerLane)
    __shared__ double s_nu[kKind == kBinsGroup ? 2 * 64 * kVolBinsPerLane : 1];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) s_perm[i] = Pg->vol.perm[i];
    if (kKind == kBinsGroup)
        for (int i = threadIdx.x; i < Pg->nbins; i += blockDim.x) {
            s_nu[i] = Pg->nu_obs[i];
            s_nu[64 * kVolBinsPerLane + i] = Pg->lognu_obs[i];
        }
    const double* const nu_tab = kKind == kBinsGroup ? s_nu : Pg->nu_obs;
    const double* const lnu_tab = kKind == kBinsGroup ? s_nu + 64 * kVolBinsPerLane : Pg->lognu_obs;
    const int n_r = Pg->vol.n_r;
    for (int i = threadIdx.x; i < n_r; i += blockDim.x) {
        s_rad[i] = Pg->vol.H[i];
        s_rad[kVolMaxR + i] = Pg->vol.z_max[i];
        s_rad[2 * kVolMaxR + i] = Pg->vol.rho_mid[i];
    }
    __syncthreads();
    const TraceParams& P = *Pg;
    const VolDev& V = P.vol;
    const VolRadial R{s_rad, s_rad + kVolMaxR, s_rad + 2 * kVolMaxR};
    const ViewParams& Vw = P.view;
    const KerrK K{Vw.M, Vw.a, Vw.a * Vw.a, 2.0 * Vw.M};
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int G = kKind == kBinsGlobal ? 1 : group;
    const int leader = lane & ~(G - 1);
    const unsigned long long n_queue = static_cast<unsigned long long>(P.n_rays);
    const volatile int* cancel = P.cancel;
    unsigned long long* const claimed = P.claimed;
    const int nb = kSpectral ? P.nbins : 3;
    const double half_pi = kPi / 2.0;
    while (true) {
        unsigned long long qv = 0;
        int stop = 0;
        if (lane == leader) {
            qv = atomicAdd(&P.counters->queue_head, 1ull);
            stop = cancel ? *cancel : 0;
            if (claimed && (qv & 63ull) == 0ull)
                __hip_atomic_store(claimed, qv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        qv = __shfl(qv, leader);
        stop = __shfl(stop, leader);
        if (qv >= n_queue || stop) break;
        const long long id = centre_out_ray(Vw, static_cast<long long>(qv), P.row_stride);
        GState s;
        SampleId sid;
        decode_ray(Vw, id, sid);      // row-major order: every id is a sample of the window
        camera_ray(Vw, sid, s.P0, s.P3, s.y);
        BinsOf<RT, kKind> b(P, id, group);
        // RGB: log10 of this lane's channel frequency, once per ray.
        double lnu_rgb = 0.0;
        if (!kSpectral && b.bin(0) >= 0) {
            const int c = b.bin(0);
            lnu_rgb = log10(c == 0 ? 2.997924e10 / 450e-7 : (c == 1 ? 2.997924e10 / 550e-7 : 2.997924e10 / 650e-7));
        }
        double dl = 0.01 * Vw.r_obs;
        int term = 2, steps = 0, attempts = 0, entries = 0, emitted = 0, march_steps = 0;
        double dtmp;
        double sth_s, cth_s;                              // sin, cos of s's theta
        sincos_cephes(s.y.v[1], sth_s, cth_s);
        Vec5 k1 = kerr_rhs_sc(K, s.P0, s.P3, s.y, dtmp, sth_s, cth_s);   // right-hand side at s
        bool turned = false;                              // p_theta changed sign in the last step
        for (int step = 0; step < P.max_steps; ++step) {
            const double r = s.y.v[0];
            if (r < Vw.r_term) { term = 0; break; }
            if (r > P.r_escape) { term = 1; break; }
            const GState prev = s;
            const double cth_p = cth_s;
            bool disk_zone = false;
            {
                // near-disk step clamp from dz/dlambda at s (k1 is the
                // derivative the reference evaluates there)
                const double sth = sth_s, cth = cth_s;
                const double z = r * cth;
                if (r >= V.r_horizon * 0.9 && r <= V.r_outer * 1.5) {
                    const double H = vol_H(V, R, r), zm = vol_zmax(V, R, r);
                    if (__builtin_fabs(z) < zm + 6.0 * H) {
                        disk_zone = true;
                        const double dz = __builtin_fabs(cth * k1.v[0] - r * sth * k1.v[1]);
                        if (dz > 1e-20) dl = smin(dl, smax(0.25 * H * rcp(dz), H / 64.0));
                    }
                }
            }
            // FP32 far field: outside r = 15 M on the way out (p_r > 0:
            // r grows along the traced ray), away from the disk's step-clamp
            // zone and not right after a theta turning point.  Inbound
            // segments stay FP64: their end point is where the ray meets the
            // disk, and the coarse ray-march (dtau ~ 0.05 per sample) turns
            // an FP32-level shift of it into percent-level intensity changes.
            const double pth_before = s.y.v[4];
            if (kFarFieldF32<RT> && P.far_f32 && r > kFarField * Vw.M && s.y.v[3] > 0.0 && !disk_zone && !turned) {
                dl = adaptive_dp45_f32(K, s, dl, attempts);
                sincos_cephes(s.y.v[1], sth_s, cth_s);
                k1 = kerr_rhs_sc(K, s.P0, s.P3, s.y, dtmp, sth_s, cth_s);
            } else if (P.integrator == 1) {
                dl = adaptive_rk4_doubling(K, s, dl, P.tol, attempts);
                sincos_cephes(s.y.v[1], sth_s, cth_s);
                k1 = kerr_rhs_sc(K, s.P0, s.P3, s.y, dtmp, sth_s, cth_s);
            } else {
                dl = adaptive_dp45(K, s, k1, dl, P.tol, attempts, sth_s, cth_s);
            }
            turned = (pth_before < 0.0) != (s.y.v[4] < 0.0);
            ++steps;
            bool opaque = false;
            if (!kSpectral) {
                bool mine = true;
#pragma unroll kVolBinsPerLane
                for (int j = 0; j < b.slots(); ++j)
                    if (b.bin(j) >= 0) mine = mine && (b.T(j) < RT(1e-6));
                opaque = b.all(mine);
            }
            if (!opaque) {
                const double tp = prev.y.v[1], tn = s.y.v[1];
                const double dp = tp - half_pi, dn = tn - half_pi;
                const double rn = s.y.v[0], rp = prev.y.v[0];
                const double zn = rn * cth_s, zp = rp * cth_p;
                const bool crossed = (dp * dn < 0.0) && __builtin_fabs(dp - dn) > 1e-12;
                const bool inside_now = vol_inside(V, R, rn, zn);
                const bool near = (__builtin_fabs(zn) < vol_zmax(V, R, rn) + 2.0 * vol_H(V, R, rn) ||
                                   __builtin_fabs(zp) < vol_zmax(V, R, rp) + 2.0 * vol_H(V, R, rp)) &&
                                  rn >= V.r_horizon && rn <= V.r_outer;
                if (crossed || inside_now || near) {
                    const double rlo = smin(rp, rn), rhi = smax(rp, rn);
                    if (rhi >= V.r_horizon && rlo <= V.r_outer && rp >= V.r_horizon * 0.9 && rp <= V.r_outer * 1.5) {
                        GState entry = prev;
                        sincos_cephes(entry.y.v[1], sth_s, cth_s);
                        const MarchOut mo = raymarch<RT, kKind>(P, V, R, s_perm, nu_tab, lnu_tab, K, entry, b, lnu_rgb,
                                                                sth_s, cth_s);
                        emitted |= mo.emitted;
                        march_steps += mo.steps;
                        ++entries;
                        s = entry;
                        k1 = kerr_rhs_sc(K, s.P0, s.P3, s.y, dtmp, sth_s, cth_s);
                        continue;
                    }
                }
            }
        }
        // I_obs = J nu_obs^3 (2h/c^2 restores CGS units).
        constexpr double two_h_c2 = 2.0 * 6.626070e-27 / (2.997924e10 * 2.997924e10);
        double col[3] = {0.0, 0.0, 0.0};
        double mine = 0.0;               // RGB: this lane's channel (slot 0 of lanes 0-2)
#pragma unroll kVolBinsPerLane
        for (int j = 0; j < b.slots(); ++j) {
            const int c = b.bin(j);
            if (c < 0) continue;
            const double nu_o = kSpectral ? P.nu_obs[c] : (c == 0 ? 2.997924e10 / 450e-7 : (c == 1 ? 2.997924e10 / 550e-7 : 2.997924e10 / 650e-7));
            const double Jt = static_cast<double>(b.J(j)) + static_cast<double>(b.Jc(j));
            const double I = Jt * two_h_c2 * nu_o * nu_o * nu_o;
            if (kSpectral) P.sample_spec[id * nb + c] = I;
            else mine = I;
        }
        if (!kSpectral) {
            const int c = b.bin(0);
            if (c >= 0) {
                // stars: the thin tracer's catalogue search (src/renderer.cpp:56-61)
                if (term == 1 && P.stars) mine += star_sample(P, s.y.v[1], s.y.v[2]);
                if (term == 1 && P.bg_tex) {      // texture background: this lane's channel
                    const TexRGB t = texture_sample(P, s.y.v[1], s.y.v[2]);
                    mine += c == 0 ? t.c[0] : (c == 1 ? t.c[1] : t.c[2]);
                }
                P.sample_rgb[3 * id + c] = mine;
            }
            for (int k = 0; k < 3; ++k) col[k] = b.share(mine, k);
        }
        if (b.lead()) {                 // per-ray outputs: one lane of the group
            const int cls = emitted ? GRRT_CLASS_DISK
                                    : (term == 0 ? GRRT_CLASS_CAPTURED
                                                 : (term == 1 ? GRRT_CLASS_ESCAPED : GRRT_CLASS_MAXSTEPS));
            P.sample_cls[id] = static_cast<uint8_t>(cls);
            if (P.records) {
                GRRTRayRecord& o = P.records[id];
                o.color[0] = col[0];
                o.color[1] = col[1];
                o.color[2] = col[2];
                o.x_final[0] = 0.0;
                o.x_final[1] = s.y.v[0];
                o.x_final[2] = s.y.v[1];
                o.x_final[3] = s.y.v[2];
                o.p_final[0] = s.P0;
                o.p_final[1] = s.y.v[3];
                o.p_final[2] = s.y.v[4];
                o.p_final[3] = s.P3;
                o.r_hit = o.g_hit = 0.0;
                o.h_max = o.h_final = 0.0;
                o.term = term;
                o.cls = cls;
                o.steps = steps;
                o.attempts = attempts;
                o.n_cross = entries;
                o.pad = march_steps;
            }
            atomicAdd(&P.counters->steps, static_cast<unsigned long long>(steps));
            atomicAdd(&P.counters->attempts, static_cast<unsigned long long>(attempts));
            atomicAdd(&P.counters->rays_done, 1ull);
            atomicAdd(&P.counters->march_steps, static_cast<unsigned long long>(march_steps));
        }
    }
}

// Average the N-bin spectra of each pixel (src/renderer.cpp:84-137); the
// output is double (ABI) or float (cube) per `out_f32`.
__global__ void finalize_spectral(const double* __restrict__ spec, const uint8_t* __restrict__ cls, void* out,
                                  long long npix, int nS, int nb, double inv, int out_f32) {
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= npix * nb) return;
    const long long p = q / nb;
    const int k = static_cast<int>(q - p * nb);
    bool done = true;                 // a cancelled render leaves untraced samples (class 0xFF): zero
    for (int s = 0; s < nS; ++s) done = done && cls[p * nS + s] != 0xFF;
    double acc = 0.0;
    if (done)
        for (int s = 0; s < nS; ++s) acc += spec[(p * nS + s) * nb + k];
    if (out_f32) static_cast<float*>(out)[q] = static_cast<float>(acc * inv);
    else static_cast<double*>(out)[q] = acc * inv;
}

int vol_group_size(int nbins) {
    if (nbins > 64 * kVolBinsPerLane) return 0;
    int g = 1;
    while (g * kVolBinsPerLane < nbins) g *= 2;
    return g;
}

template <int kWaves>
void launch_vol_kernel(const TraceParams* P, bool spectral, bool fp32_rt, int G, dim3 grid, hipStream_t st) {
    const dim3 blk(64);
    if (spectral && G > 0) {
        if (fp32_rt) hipLaunchKernelGGL((trace_vol<float, kBinsGroup, kWaves>), grid, blk, 0, st, P, G);
        else hipLaunchKernelGGL((trace_vol<double, kBinsGroup, kWaves>), grid, blk, 0, st, P, G);
    } else if (spectral) {
        if (fp32_rt) hipLaunchKernelGGL((trace_vol<float, kBinsGlobal, kWaves>), grid, blk, 0, st, P, 1);
        else hipLaunchKernelGGL((trace_vol<double, kBinsGlobal, kWaves>), grid, blk, 0, st, P, 1);
    } else {
        if (fp32_rt) hipLaunchKernelGGL((trace_vol<float, kBinsRGB, kWaves>), grid, blk, 0, st, P, G);
        else hipLaunchKernelGGL((trace_vol<double, kBinsRGB, kWaves>), grid, blk, 0, st, P, G);
    }
}

hipError_t launch_trace_vol(const TraceParams* P, long long n_rays, int nbins, bool spectral, bool fp32_rt,
                            int num_cus, int waves_per_simd, hipStream_t st) {
    const int G = spectral ? vol_group_size(nbins) : kRgbGroup;
    // Persistent one-wavefront blocks, one per wave slot the register
    // budget leaves (4 SIMDs per CU), never more than the queue can feed.
    const long long per_wave = 64 / (G > 0 ? G : 1);
    const long long need = (n_rays + per_wave - 1) / per_wave;
    const int waves = waves_per_simd >= 2 ? 2 : 1;
    const long long blocks = std::max<long long>(1, std::min<long long>(need, 4LL * num_cus * waves));
    const dim3 grid(static_cast<unsigned>(blocks));
    if (waves == 2) launch_vol_kernel<2>(P, spectral, fp32_rt, G, grid, st);
    else launch_vol_kernel<1>(P, spectral, fp32_rt, G, grid, st);
    return hipGetLastError();
}

hipError_t launch_finalize_spectral(const double* spec, const uint8_t* cls, void* out, long long npix, int nS,
                                    int nb, bool f32, hipStream_t st) {
    const int T = 256;
    const long long n = npix * nb;
    hipLaunchKernelGGL(finalize_spectral, dim3(static_cast<unsigned>((n + T - 1) / T)), dim3(T), 0, st, spec,
                       cls, out, npix, nS, nb, 1.0 / nS, f32 ? 1 : 0);
    return hipGetLastError();
}

}  // namespace grrt_amd
