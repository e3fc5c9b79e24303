// Persistent thin-disk geodesic tracer for gfx950.
//
// One lane owns one ray.  A wavefront loops over Dormand–Prince attempts;
// every loop iteration is exactly one attempt for every live lane, so lanes
// whose rays take more or fewer steps never wait on each other.  When enough
// lanes have finished (refill_min), the wavefront claims that many new ray
// indices from a device-wide queue with a single atomic and hands them to
// its idle lanes in lane order (ballot + mbcnt), which keeps the wavefront
// full until the queue drains.
//
// Per-ray semantics follow the reference:
//   src/geodesic_tracer.cpp:23-181   outer loop, termination, disk crossings
//   src/rk4.cpp:87-136              adaptive step control
//   include/grrt/geodesic/rk4.h:73-177  Dormand–Prince stages and error norm
//   src/camera.cpp:69-96             pixel -> null covector
//   src/renderer.cpp:14-82           sample placement and averaging
//   src/accretion_disk.cpp:96-147    thin-disk emission
//   src/celestial_sphere.cpp:48-90   star background
// The accepted state's k7 (first-same-as-last) becomes the next step's k1,
// which is the same evaluation the reference repeats.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_math.hpp"
#include "kerr_rhs.hpp"
#include "camera_device.hpp"
#include "disk_device.hpp"
#include "star_device.hpp"
#include "trace_params.hpp"

namespace grrt_amd {

// Dormand–Prince 5(4) tableau (Dormand & Prince 1980).
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

// Algorithmic FP64 work per Dormand–Prince attempt as this kernel performs
// it: six right-hand-side evaluations (k1 is reused) of 120 flops each, 26
// coefficient products, stage/solution/error combinations over five
// components (5*(2+4+6+8+10) + 5*10 + 5*12 = 260), the error norm (5*5) and
// the step controller (12).  Used only for reporting.
constexpr double kFlopsPerAttempt = 6 * 120 + 26 + 260 + 25 + 12;

// 0.9 * x^(1/5) for the step controller, in FP32.
__device__ __forceinline__ double step_factor(double x) {
    const float lx = __builtin_amdgcn_logf(static_cast<float>(x));   // log2
    return static_cast<double>(0.9f * __builtin_amdgcn_exp2f(0.2f * lx));
}

struct Ray {
    Vec5 y;          // r, theta, phi, p_r, p_theta
    Vec5 k1;         // right-hand side at y
    double P0, P3;   // conserved p_t, p_phi
    double dl;
    double col[3];
    double r_hit, g_hit, h_max, h_last;
    int steps, attempts, retry, n_cross;
};

__device__ __forceinline__ double hamiltonian_from_rhs(const Ray& R, double dt) {
    // H = p_mu dx^mu / 2 with dx^mu = g^{mu nu} p_nu.
    return 0.5 * (R.P0 * dt + R.y.v[3] * R.k1.v[0] + R.y.v[4] * R.k1.v[1] + R.P3 * R.k1.v[2]);
}

// Classic RK4 step of length h from y whose right-hand side k1 is known
// (include/grrt/geodesic/rk4.h:34-52).
__device__ __forceinline__ Vec5 rk4_from(const KerrK& K, double P0, double P3, const Vec5& y, const Vec5& k1,
                                         double h) {
    Vec5 s, o;
    double dt;
#pragma unroll
    for (int i = 0; i < 5; ++i) s.v[i] = y.v[i] + k1.v[i] * (h * 0.5);
    const Vec5 k2 = kerr_rhs(K, P0, P3, s, dt);
#pragma unroll
    for (int i = 0; i < 5; ++i) s.v[i] = y.v[i] + k2.v[i] * (h * 0.5);
    const Vec5 k3 = kerr_rhs(K, P0, P3, s, dt);
#pragma unroll
    for (int i = 0; i < 5; ++i) s.v[i] = y.v[i] + k3.v[i] * h;
    const Vec5 k4 = kerr_rhs(K, P0, P3, s, dt);
#pragma unroll
    for (int i = 0; i < 5; ++i) o.v[i] = y.v[i] + (k1.v[i] + k2.v[i] * 2.0 + k3.v[i] * 2.0 + k4.v[i]) * (h / 6.0);
    return o;
}

// Ray start: camera covector (src/camera.cpp:69-96) and the first
// right-hand side.
__device__ __forceinline__ void init_ray(const ViewParams& V, const SampleId& d, const KerrK& K, Ray& R) {
    camera_ray(V, d, R.P0, R.P3, R.y);
    R.dl = 0.01 * V.r_obs;
    R.col[0] = R.col[1] = R.col[2] = 0.0;
    R.r_hit = R.g_hit = 0.0;
    R.steps = R.attempts = R.retry = R.n_cross = 0;
    double dt;
    R.k1 = kerr_rhs(K, R.P0, R.P3, R.y, dt);
    R.h_last = __builtin_fabs(hamiltonian_from_rhs(R, dt));
    R.h_max = R.h_last;
}

// Thin-disk emission at an in-range crossing (src/accretion_disk.cpp:96-147).
__device__ __forceinline__ void disk_emission(const ViewParams& V, const double* flux,
                                              const double* s_spec, double rc, double P0,
                                              double P3, Ray& R) {
    const double g = disk_redshift(V, rc, P0, P3);
    if (R.n_cross == 0) {
        R.r_hit = rc;
        R.g_hit = g;
    }
    ++R.n_cross;
    thin_disk_color(V, flux, s_spec, rc, g, R.col);
}

__device__ __forceinline__ void finish_ray(const TraceParams& P, long long oid, Ray& R, int term) {
    if (term == 1 && P.stars) {
        const double b = star_sample(P, R.y.v[1], R.y.v[2]);
        R.col[0] += b;
        R.col[1] += b;
        R.col[2] += b;
    } else if (term == 1 && P.bg_tex) {
        const TexRGB t = texture_sample(P, R.y.v[1], R.y.v[2]);
        R.col[0] += t.c[0];
        R.col[1] += t.c[1];
        R.col[2] += t.c[2];
    }
    const int cls = (R.n_cross > 0) ? GRRT_CLASS_DISK
                                    : (term == 0 ? GRRT_CLASS_CAPTURED
                                                 : (term == 1 ? GRRT_CLASS_ESCAPED : GRRT_CLASS_MAXSTEPS));
    P.sample_rgb[3 * oid + 0] = R.col[0];
    P.sample_rgb[3 * oid + 1] = R.col[1];
    P.sample_rgb[3 * oid + 2] = R.col[2];
    P.sample_cls[oid] = static_cast<uint8_t>(cls);
    if (P.records) {
        GRRTRayRecord& o = P.records[oid];
        o.color[0] = R.col[0];
        o.color[1] = R.col[1];
        o.color[2] = R.col[2];
        o.x_final[0] = 0.0;
        o.x_final[1] = R.y.v[0];
        o.x_final[2] = R.y.v[1];
        o.x_final[3] = R.y.v[2];
        o.p_final[0] = R.P0;
        o.p_final[1] = R.y.v[3];
        o.p_final[2] = R.y.v[4];
        o.p_final[3] = R.P3;
        o.r_hit = R.r_hit;
        o.g_hit = R.g_hit;
        o.h_max = R.h_max;
        o.h_final = R.h_last;
        o.term = term;
        o.cls = cls;
        o.steps = R.steps;
        o.attempts = R.attempts;
        o.n_cross = R.n_cross;
        o.pad = 0;
    }
}

// Termination test of src/geodesic_tracer.cpp:52-63 on the current state;
// -1 while the ray continues.
__device__ __forceinline__ int termination(int max_steps, double r_term, double r_escape, const Ray& R) {
    if (R.steps >= max_steps) return 2;
    const double r = R.y.v[0];
    if (r < r_term) return 0;
    if (r > r_escape) return 1;
    return -1;
}

// Compiler-only fence: keeps the parameter-block loads of the rare paths
// (ray start, disk crossing, ray end) where they are used instead of
// hoisting them out of the loop into registers.
__device__ __forceinline__ void keep_loads_here() { asm volatile("" ::: "memory"); }

__device__ __forceinline__ unsigned lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

// View owning global queue index m (views sorted by q_offset).
__device__ __forceinline__ int find_view(const ViewParams* views, int n, unsigned long long m) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (static_cast<unsigned long long>(views[mid].q_offset) <= m) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}


// The parameter block lives in device memory; the hot scalars are read once
// at entry, everything else where it is used.
//   kMode 0: one view, queue slot -> sample of the view's window.
//   kMode 1 (batch): several views (each with its own spin, camera and disk)
//     concatenated in queue order; every lane carries its own view's metric
//     constants.
//   kMode 2 (list): one view, the queue is the listed ray ids (the surrogate's
//     fallback rays, then its judged rays, whose results go after the frame's
//     samples); the list lengths are read from device memory.
//   kDoubling: the step-doubling RK4 controller (src/rk4.cpp:33-80) in place
//     of Dormand-Prince; a separate instantiation, so the Dormand-Prince
//     kernels carry none of its code.
// Register budget: the frame and batch kernels are held to 168 VGPRs,
// three waves per SIMD, with a few dwords spilled outside the attempt
// block (Kerr a=0.998 1024^2 x 4 frame: 17.1 ms against 17.4 ms at 189
// VGPRs and two waves; the 64-view batch: 128 ms against 133 ms).  The list
// kernel, launched one block per CU for the surrogate's fallback rays, keeps
// two waves and no spills.
template <int kMode, bool kDoubling>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kMode == 2 ? 2 : 3)))
void trace_thin(const TraceParams* Pg) {
    constexpr bool kBatch = kMode == 1;
    constexpr bool kList = kMode == 2;
    __shared__ double s_flux[kBatch ? 1 : kFluxN];
    __shared__ double s_spec[3 * kSpecN];
    const double r_escape = Pg->r_escape, tol = Pg->tol;
    const int max_steps = Pg->max_steps, refill_min = Pg->refill_min, chunk = Pg->chunk;
    constexpr bool rk4_doubling = kDoubling;
    unsigned long long n_fb = 0;
    long long n_queue = Pg->n_rays;
    if (kList) {
        n_fb = Pg->list_counts[0];
        const unsigned long long nj = Pg->list_counts[1];
        n_queue = static_cast<long long>(n_fb + (nj < static_cast<unsigned long long>(Pg->judge_cap)
                                                      ? nj : static_cast<unsigned long long>(Pg->judge_cap)));
    }
    DevCounters* const counters = Pg->counters;
    const int use_disk = kBatch ? 1 : Pg->view.disk;
    if (!kBatch && use_disk) {
        const double* fl = Pg->view.flux_lut;
        for (int i = threadIdx.x; i < kFluxN; i += blockDim.x) s_flux[i] = fl[i];
    }
    {
        const double* sl = Pg->spec_lut;
        for (int i = threadIdx.x; i < 3 * kSpecN; i += blockDim.x) s_spec[i] = sl[i];
    }
    __syncthreads();

    // Metric constants: uniform for a single view, per lane in a batch.
    KerrK K{Pg->view.M, Pg->view.a, Pg->view.a * Pg->view.a, 2.0 * Pg->view.M};
    double r_term = Pg->view.r_term;
    int vi = 0;
    const unsigned lane = lane_id();
    Ray R;
    long long oid = -1;
    bool live = false;
    unsigned long long n_steps = 0, n_att = 0, n_rays = 0;

    // Ray indices are claimed from the device-wide queue in chunks of CH
    // (>= 64) by lane 0.  The claim for the next chunk is issued as soon as
    // the previous one is taken, so the atomic's latency overlaps with the
    // integration; the chunk is read only when the current range runs dry.
    // A cancel request stops further claims; claimed rays always complete,
    // so the finished rays are exactly [0, queue_head).
    const unsigned long long nR = static_cast<unsigned long long>(n_queue);
    const unsigned long long CH = static_cast<unsigned long long>(chunk);
    const volatile int* cancel = Pg->cancel;
    unsigned long long* const claimed = Pg->claimed;
    unsigned long long cur = 0, cur_end = 0;     // wave-uniform claimed range
    unsigned long long pre = 0;                  // lane 0: base of the claim in flight
    int cancel_seen = 0;                         // lane 0: cancel flag read with that claim
    bool have_pre = true;
    bool queue_open = true;
    if (lane == 0) {
        pre = atomicAdd(&counters->queue_head, CH);
        cancel_seen = cancel ? *cancel : 0;
    }

    while (true) {
        // ---- refill idle lanes (wave-uniform branch) ----
        const unsigned long long idle = __ballot(!live);
        const int n_idle = __popcll(idle);
        if (queue_open && (n_idle >= refill_min || n_idle == 64)) {
            // cur never passes cur_end (see the advance below), so avail >= 0.
            const unsigned long long avail = cur_end > cur ? cur_end - cur : 0ull;
            unsigned long long nb = ~0ull;
            if (avail < static_cast<unsigned long long>(n_idle) && have_pre) {
                const unsigned long long got = __shfl(pre, 0);
                const int stop = __shfl(cancel_seen, 0);
                have_pre = false;
                // Progress for the host: a plain system-scope store of the
                // claim position (stores from different waves may land out
                // of order; the host keeps the largest value it has seen).
                if (lane == 0 && claimed)
                    __hip_atomic_store(claimed, got, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                if (got < nR) {
                    nb = got;
                    if (!stop) {
                        if (lane == 0) {
                            pre = atomicAdd(&counters->queue_head, CH);
                            cancel_seen = cancel ? *cancel : 0;
                        }
                        have_pre = true;
                    }
                }
            }
            if (!live) {
                const unsigned long long k = lanes_below(idle);
                unsigned long long mine = ~0ull;
                if (k < avail) mine = cur + k;
                else if (nb != ~0ull) mine = nb + (k - avail);
                if (mine < nR) {
                    keep_loads_here();
                    if (kBatch) vi = find_view(Pg->views, Pg->n_views, mine);
                    const ViewParams& V = kBatch ? Pg->views[vi] : Pg->view;
                    SampleId d;
                    long long q = static_cast<long long>(mine) - V.q_offset;
                    if (kList) q = mine < n_fb ? Pg->fb_list[mine] : Pg->judge_list[mine - n_fb];
                    if (decode_ray(V, q, d)) {
                        if (kBatch) {
                            K = KerrK{V.M, V.a, V.a * V.a, 2.0 * V.M};
                            r_term = V.r_term;
                        }
                        oid = V.ray_offset + d.local;
                        if (kList && mine >= n_fb) oid = Pg->judge_out + static_cast<long long>(mine - n_fb);
                        init_ray(V, d, K, R);
                        live = true;
                        const int term = termination(max_steps, r_term, r_escape, R);
                        if (term >= 0) {
                            finish_ray(*Pg, oid, R, term);
                            live = false;
                            ++n_rays;
                        }
                    }
                }
            }
            if (avail >= static_cast<unsigned long long>(n_idle)) {
                cur += n_idle;
            } else if (nb != ~0ull) {
                cur_end = nb + CH < nR ? nb + CH : nR;
                const unsigned long long next = nb + (n_idle - avail);
                cur = next < cur_end ? next : cur_end;   // the last chunk may be short
            } else {
                cur = cur_end;
            }
            if (cur >= cur_end && !have_pre) queue_open = false;
        }
        if (__ballot(live) == 0ull) {
            if (!queue_open) break;
            continue;
        }
        if (!live) continue;

        const double h = R.dl;
        Vec5 y5, k7;
        double dt7 = 0.0, err = 0.0;
        if (rk4_doubling) {
            // ---- one step-doubling RK4 attempt (src/rk4.cpp:33-80): a step of
            // h against two of h/2, error over r, theta, phi, p_r, p_theta
            // (p_t and p_phi are exact in both) ----
            const Vec5 full = rk4_from(K, R.P0, R.P3, R.y, R.k1, h);
            const Vec5 mid = rk4_from(K, R.P0, R.P3, R.y, R.k1, h * 0.5);
            double dtm;
            const Vec5 km = kerr_rhs(K, R.P0, R.P3, mid, dtm);
            y5 = rk4_from(K, R.P0, R.P3, mid, km, h * 0.5);
#pragma unroll
            for (int i = 0; i < 5; ++i)
                err = __builtin_fmax(err, __builtin_fabs(full.v[i] - y5.v[i]) * rcp1(__builtin_fabs(y5.v[i]) + 1e-10));
        } else {
        // ---- one Dormand–Prince attempt (include/grrt/geodesic/rk4.h:73-177) ----
        const double h21 = h * A21;
        const double h31 = h * A31, h32 = h * A32;
        const double h41 = h * A41, h42 = h * A42, h43 = h * A43;
        const double h51 = h * A51, h52 = h * A52, h53 = h * A53, h54 = h * A54;
        const double h61 = h * A61, h62 = h * A62, h63 = h * A63, h64 = h * A64, h65 = h * A65;
        Vec5 s, k2, k3, k4, k5, k6, e;
#pragma unroll
        for (int i = 0; i < 5; ++i) s.v[i] = __builtin_fma(R.k1.v[i], h21, R.y.v[i]);
        k2 = kerr_rhs(K, R.P0, R.P3, s, dt7);
#pragma unroll
        for (int i = 0; i < 5; ++i)
            s.v[i] = __builtin_fma(k2.v[i], h32, __builtin_fma(R.k1.v[i], h31, R.y.v[i]));
        k3 = kerr_rhs(K, R.P0, R.P3, s, dt7);
#pragma unroll
        for (int i = 0; i < 5; ++i)
            s.v[i] = __builtin_fma(k3.v[i], h43, __builtin_fma(k2.v[i], h42,
                         __builtin_fma(R.k1.v[i], h41, R.y.v[i])));
        k4 = kerr_rhs(K, R.P0, R.P3, s, dt7);
#pragma unroll
        for (int i = 0; i < 5; ++i)
            s.v[i] = __builtin_fma(k4.v[i], h54, __builtin_fma(k3.v[i], h53,
                         __builtin_fma(k2.v[i], h52, __builtin_fma(R.k1.v[i], h51, R.y.v[i]))));
        k5 = kerr_rhs(K, R.P0, R.P3, s, dt7);
#pragma unroll
        for (int i = 0; i < 5; ++i)
            s.v[i] = __builtin_fma(k5.v[i], h65, __builtin_fma(k4.v[i], h64,
                         __builtin_fma(k3.v[i], h63, __builtin_fma(k2.v[i], h62,
                             __builtin_fma(R.k1.v[i], h61, R.y.v[i])))));
        k6 = kerr_rhs(K, R.P0, R.P3, s, dt7);
        const double b1 = h * B1, b3 = h * B3, b4 = h * B4, b5 = h * B5, b6 = h * B6;
        const double e1 = h * E1, e3 = h * E3, e4 = h * E4, e5 = h * E5, e6 = h * E6, e7 = h * E7;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            y5.v[i] = __builtin_fma(k6.v[i], b6, __builtin_fma(k5.v[i], b5,
                          __builtin_fma(k4.v[i], b4, __builtin_fma(k3.v[i], b3,
                              __builtin_fma(R.k1.v[i], b1, R.y.v[i])))));
            e.v[i] = __builtin_fma(k6.v[i], e6, __builtin_fma(k5.v[i], e5,
                         __builtin_fma(k4.v[i], e4, __builtin_fma(k3.v[i], e3, R.k1.v[i] * e1))));
        }
        k7 = kerr_rhs(K, R.P0, R.P3, y5, dt7);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const double ei = __builtin_fma(k7.v[i], e7, e.v[i]);
            err = __builtin_fmax(err, __builtin_fabs(ei) * rcp1(__builtin_fabs(y5.v[i]) + 1e-10));
        }
        }
        ++R.attempts;

        // ---- step control (src/rk4.cpp:87-136) ----
        // The growth/shrink factor 0.9 (tol/err)^(1/5) is evaluated in FP32
        // (hardware log2/exp2).  It only sets the next trial step length; a
        // relative change d in that length moves the solution by about
        // d * (local error), far below the FP64 rounding of the state.
        bool accept;
        double next;
        if (R.retry >= 20) {
            accept = true;           // retries exhausted: take the last attempt
            next = h;
        } else if (rk4_doubling) {
            // src/rk4.cpp:60-75: keep h (double it below 1% of tol), or halve
            if (err <= tol) {
                accept = true;
                next = smin(err < tol * 0.01 ? h * 2.0 : h, 5.0 * smax(__builtin_fabs(y5.v[0]), 1.0));
            } else if (h * 0.5 < 1e-6) {
                accept = true;       // step too small: keep the two half steps
                next = smin(1e-6, 5.0 * smax(__builtin_fabs(y5.v[0]), 1.0));
            } else {
                accept = false;
                next = h * 0.5;
            }
        } else if (err <= tol) {
            accept = true;
            double f = (err > 1e-30) ? step_factor(tol / err) : 5.0;
            f = sclamp(f, 1.0, 5.0);
            next = smin(h * f, 5.0 * smax(__builtin_fabs(y5.v[0]), 1.0));
        } else {
            double f = step_factor(tol / err);
            f = sclamp(f, 0.2, 1.0);
            const double dl = h * f;
            if (dl < 1e-6) {
                accept = true;       // step too small: keep the rejected y5
                next = smin(1e-6, 5.0 * smax(__builtin_fabs(y5.v[0]), 1.0));
            } else {
                accept = false;
                next = dl;
            }
        }
        R.dl = next;
        if (!accept) {
            ++R.retry;
            continue;
        }
        R.retry = 0;
        if (rk4_doubling) k7 = kerr_rhs(K, R.P0, R.P3, y5, dt7);   // right-hand side at the new state

        // ---- thin-disk crossing between the previous and new state ----
        if (use_disk) {
            const double half_pi = kPi / 2.0;
            const double dp = R.y.v[1] - half_pi, dn = y5.v[1] - half_pi;
            if (dp * dn < 0.0 && __builtin_fabs(dp - dn) > 1e-12) {
                const double frac = -dp / (dn - dp);
                const double rc = R.y.v[0] + frac * (y5.v[0] - R.y.v[0]);
                keep_loads_here();
                const ViewParams& V = kBatch ? Pg->views[vi] : Pg->view;
                if (V.disk && rc >= V.r_in && rc <= V.r_out)
                    disk_emission(V, kBatch ? V.flux_lut : s_flux, s_spec, rc, R.P0, R.P3, R);
            }
        }
        R.y = y5;
        R.k1 = k7;
        ++R.steps;
        R.h_last = __builtin_fabs(hamiltonian_from_rhs(R, dt7));
        R.h_max = smax(R.h_max, R.h_last);
        const int term = termination(max_steps, r_term, r_escape, R);
        if (term >= 0) {
            n_steps += R.steps;
            n_att += R.attempts;
            ++n_rays;
            keep_loads_here();
            finish_ray(*Pg, oid, R, term);
            live = false;
        }
    }

    // Per-wave statistics, one atomic per counter per wave.
    for (int off = 32; off > 0; off >>= 1) {
        n_steps += __shfl_xor(n_steps, off);
        n_att += __shfl_xor(n_att, off);
        n_rays += __shfl_xor(n_rays, off);
    }
    if (lane == 0) {
        atomicAdd(&counters->steps, n_steps);
        atomicAdd(&counters->attempts, n_att);
        atomicAdd(&counters->rays_done, n_rays);
    }
}

// Average the samples of each pixel in sample order (src/renderer.cpp:56-72)
// and write RGBA float32.  A pixel with a sample that was never traced (a
// cancelled render leaves the 0xFF class sentinel) is written as zero.
__global__ void finalize_pixels(const double* __restrict__ rgb, const uint8_t* __restrict__ cls,
                                float* __restrict__ out, long long npix, int nS, double inv) {
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    bool done = true;
    for (int s = 0; s < nS; ++s) done = done && cls[p * nS + s] != 0xFF;
    float4 o;
    if (!done) {
        o.x = o.y = o.z = 0.0f;
    } else {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int s = 0; s < nS; ++s) {
            const double* c = rgb + 3 * (p * nS + s);
            a0 = a0 + c[0];
            a1 = a1 + c[1];
            a2 = a2 + c[2];
        }
        o.x = static_cast<float>(a0 * inv);
        o.y = static_cast<float>(a1 * inv);
        o.z = static_cast<float>(a2 * inv);
    }
    o.w = 1.0f;
    reinterpret_cast<float4*>(out)[p] = o;
}

// Zero the counters and mark every sample untraced.
__global__ void reset_launch(DevCounters* c, uint8_t* __restrict__ cls, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < static_cast<long long>(sizeof(DevCounters) / 8)) reinterpret_cast<unsigned long long*>(c)[i] = 0ull;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long k = i; k < n; k += stride) cls[k] = 0xFF;
}

}  // namespace grrt_amd

// Host-callable launchers (C++ linkage inside the library).
namespace grrt_amd {

hipError_t launch_reset(DevCounters* c, uint8_t* cls, long long n_rays, hipStream_t st) {
    const long long blocks = std::min<long long>(std::max<long long>((n_rays + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(reset_launch, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, c, cls, n_rays);
    return hipGetLastError();
}

hipError_t launch_trace_thin(const TraceParams* P, int mode, int blocks, int threads, hipStream_t st,
                             int integrator) {
    const dim3 g(static_cast<unsigned>(blocks)), b(static_cast<unsigned>(threads));
    if (integrator == 1) {
        if (mode == 1) hipLaunchKernelGGL((trace_thin<1, true>), g, b, 0, st, P);
        else if (mode == 2) hipLaunchKernelGGL((trace_thin<2, true>), g, b, 0, st, P);
        else hipLaunchKernelGGL((trace_thin<0, true>), g, b, 0, st, P);
    } else {
        if (mode == 1) hipLaunchKernelGGL((trace_thin<1, false>), g, b, 0, st, P);
        else if (mode == 2) hipLaunchKernelGGL((trace_thin<2, false>), g, b, 0, st, P);
        else hipLaunchKernelGGL((trace_thin<0, false>), g, b, 0, st, P);
    }
    return hipGetLastError();
}

hipError_t launch_finalize(const double* rgb, const uint8_t* cls, float* out, long long npix, int nS,
                           hipStream_t st) {
    const int T = 256;
    const long long B = (npix + T - 1) / T;
    hipLaunchKernelGGL(finalize_pixels, dim3(static_cast<unsigned>(B)), dim3(T), 0, st, rgb, cls, out,
                       npix, nS, 1.0 / nS);
    return hipGetLastError();
}

double flops_per_attempt() { return kFlopsPerAttempt; }

}  // namespace grrt_amd
