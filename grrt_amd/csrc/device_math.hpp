// Device-side scalar helpers for gfx950 (CDNA4): reciprocal, sin/cos of the
// polar angle, and the std::min/max/clamp comparison semantics the
// reference relies on.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace grrt_amd {

constexpr double kPi = 3.141592653589793;   // std::numbers::pi

// 1/x from v_rcp_f64 refined by two Newton–Raphson steps (FMA form),
// accurate to about 1 ulp for the finite, non-zero arguments used here.
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// 1/x with one Newton–Raphson step (about 1e-15 relative), for ratios that
// only feed comparisons and the step controller.
__device__ __forceinline__ double rcp1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// Comparison semantics of std::min / std::max / std::clamp: a NaN in the
// second operand of min/max is ignored, clamp passes a NaN through.
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double sclamp(double v, double lo, double hi) {
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

// sin and cos of the polar angle by the Cephes method (S. L. Moshier,
// public domain), the method of include/grrt/math/fast_trig.h:21-82: octant
// j = even rounding of |x|*4/pi, three-term Cody–Waite reduction by pi/4,
// then the Cephes minimax polynomials for sin and cos on [-pi/4, pi/4].  The
// coefficient table is the published Cephes one, kept identical so both
// paths evaluate the same approximation.  Quadrant selection is branch-free.
__device__ __forceinline__ void sincos_cephes(double x, double& s, double& c) {
    constexpr double DP1 = 7.85398125648498535156e-1;
    constexpr double DP2 = 3.77489470793079817668e-8;
    constexpr double DP3 = 2.69515142907905952645e-15;
    constexpr double FOPI = 1.27323954473516268615;
    const double ax = __builtin_fabs(x);
    int j = static_cast<int>(ax * FOPI);
    j = (j + 1) & ~1;
    const double y = static_cast<double>(j);
    const double t = __builtin_fma(-y, DP3, __builtin_fma(-y, DP2, __builtin_fma(-y, DP1, ax)));
    const double z = t * t;
    double ps = __builtin_fma(1.58969099521155010221e-10, z, -2.50507602534068634195e-08);
    ps = __builtin_fma(ps, z, 2.75573137070700676789e-06);
    ps = __builtin_fma(ps, z, -1.98412698298579493134e-04);
    ps = __builtin_fma(ps, z, 8.33333333332248946124e-03);
    ps = __builtin_fma(ps, z, -1.66666666666666324348e-01);
    double pc = __builtin_fma(-1.13596475577881948265e-11, z, 2.08757232129817482790e-09);
    pc = __builtin_fma(pc, z, -2.75573143513906633035e-07);
    pc = __builtin_fma(pc, z, 2.48015872894767294178e-05);
    pc = __builtin_fma(pc, z, -1.38888888888741095749e-03);
    pc = __builtin_fma(pc, z, 4.16666666666666019037e-02);
    const double sr = __builtin_fma(t * z, ps, t);
    const double cr = __builtin_fma(z * z, pc, __builtin_fma(-0.5, z, 1.0));
    // Quadrant q = j/2: q&1 swaps sin/cos and negates cos, q&2 negates both.
    const int q = j >> 1;
    const bool sw = (q & 1) != 0;
    double so = sw ? cr : sr;
    double co = sw ? sr : cr;
    const bool neg_s = ((q & 2) != 0) != (x < 0.0);
    const bool neg_c = ((q & 2) != 0) != sw;
    s = neg_s ? -so : so;
    c = neg_c ? -co : co;
}

// SplitMix64 finaliser (Steele, Lea & Flood 2014), for seeded jitter.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace grrt_amd
