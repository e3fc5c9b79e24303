// Surrogate MLP kernels for gfx950: forward (training, plain, and the frame
// inference pass with fallback queuing and shading), backward, Adam.
//
// Dense layers run as v_mfma_f32_32x32x16_bf16 with the network's neurons
// as the MFMA rows (A operand = weights) and rays as the columns (B operand
// = activations).  A 32x32 accumulator tile holds row (v/4)*8 + g*4 + v%4
// in register v of lane group g = lane/32, for ray lane%32.  The weights are
// stored pre-permuted (kslot_neuron) so that one layer's accumulators,
// passed through SiLU and rounded to BF16, are the next layer's B operand in
// the registers that already hold them: no LDS round trip between layers.
//
// The 5 inputs enter layer 0 as BF16 hi/lo pairs (x = hi + lo, 16 of the
// 16 K slots: 10 used), so the first layer sees them to ~2^-16 relative
// while its weights are BF16 like every other layer's.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "camera_device.hpp"
#include "disk_device.hpp"
#include "star_device.hpp"
#include "surrogate.hpp"

namespace grrt_amd::sg {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ __forceinline__ f32x16 mfma(const bf16x8& a, const bf16x8& b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Logistic function with the hardware exp2 and reciprocal (about 1 ulp each;
// the network is trained and evaluated with this same function).
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float bf16r(float x) { return static_cast<float>(static_cast<__bf16>(x)); }

__device__ __forceinline__ long long data_row(const MlpArgs& A, long long i) {
    if (A.perm_a == 0) return i;
    const uint64_t k = A.perm_a * static_cast<uint64_t>(A.perm_base + i) + A.perm_c;
    return static_cast<long long>(k & static_cast<uint64_t>(A.n_data - 1));
}

// Accumulator tile initialised with the bias of its 32 rows.
__device__ __forceinline__ f32x16 bias_tile(const float* b, int mt, int g, int rows) {
    f32x16 o;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int m = mt * 32 + acc_row(g, v);
        o[v] = m < rows ? b[m] : 0.0f;
    }
    return o;
}

// SiLU of a finished hidden layer into the next layer's B operands; in
// training also the pre-activations Z^T and the BF16 activations H^T (both
// neuron-major: lanes store consecutive rays) for the backward pass.
template <int NT, bool kTrain>
__device__ __forceinline__ void finish_hidden(const MlpArgs& A, int layer, long long ray0, f32x16 (&acc)[NT][8],
                                              bf16x8 (&act)[NT][16]) {
    const int lane = threadIdx.x & 63, g = lane >> 5, c = lane & 31;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const long long ray = ray0 + t * 32 + c;
        const bool ok = ray < A.n;
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const float z = acc[t][mt][v];
                const __bf16 h = static_cast<__bf16>(silu(z));
                act[t][2 * mt + (v >> 3)][v & 7] = h;
                if (kTrain && ok) {
                    const int m = mt * 32 + acc_row(g, v);
                    A.Z[(static_cast<long long>(layer) * kHid + m) * A.n + ray] = z;
                    A.HT[(static_cast<long long>(layer) * kHid + m) * A.n + ray] = h;
                }
            }
        }
    }
}

// Last hidden layer: SiLU in place in FP32 (the output layer reads it as
// BF16 hi + lo pairs, so its input keeps ~16 significant bits); training
// stores Z^T and the BF16 H^T as for the other layers.
template <int NT, bool kTrain>
__device__ __forceinline__ void finish_last_hidden(const MlpArgs& A, long long ray0, f32x16 (&acc)[NT][8]) {
    const int lane = threadIdx.x & 63, g = lane >> 5, c = lane & 31;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const long long ray = ray0 + t * 32 + c;
        const bool ok = ray < A.n;
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const float z = acc[t][mt][v];
                const float h = silu(z);
                acc[t][mt][v] = h;
                if (kTrain && ok) {
                    const int m = mt * 32 + acc_row(g, v);
                    A.Z[(3LL * kHid + m) * A.n + ray] = z;
                    A.HT[(3LL * kHid + m) * A.n + ray] = static_cast<__bf16>(h);
                }
            }
        }
    }
}

// B operands (hi, lo) of k-step s from FP32 activations held in
// accumulator layout.
template <int NT>
__device__ __forceinline__ void hilo_operand(const f32x16 (&h)[NT][8], int t, int s, bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float x = h[t][s >> 1][(s & 1) * 8 + j];
        const __bf16 xh = static_cast<__bf16>(x);
        hi[j] = xh;
        lo[j] = static_cast<__bf16>(x - static_cast<float>(xh));
    }
}

// Layer 0 of NT ray tiles: inputs as BF16 hi/lo pairs, one k-step.
template <int NT>
__device__ __forceinline__ void input_layer(const float (&x)[NT][kIn], const __bf16* pack0, const float* b0,
                                            f32x16 (&acc)[NT][8]) {
    const int lane = threadIdx.x & 63, g = lane >> 5;
    bf16x8 in0[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int f = g * 8 + j;
            float v = 0.0f;
            if (f < 2 * kIn) {
                const float xi = x[t][f >> 1];
                const float hi = bf16r(xi);
                v = (f & 1) ? xi - hi : hi;
            }
            in0[t][j] = static_cast<__bf16>(v);
        }
    }
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        const f32x16 b = bias_tile(b0, mt, g, kHid);
        const bf16x8 w = ld8(pack0 + (mt * 64 + lane) * 8);
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t][mt] = mfma(w, in0[t], b);
    }
}

// The whole network for NT tiles of 32 rays (lane c of tile t = ray
// ray0 + 32 t + c), weights read from global memory.  x: that ray's inputs.
// Writes the 9 outputs per ray into outs[32 NT][9] (the calling wave's LDS
// slice).
template <int NT, bool kTrain>
__device__ __forceinline__ void mlp_run(const MlpArgs& A, const float (&x)[NT][kIn], long long ray0,
                                        float (*outs)[kOut]) {
    const int lane = threadIdx.x & 63, g = lane >> 5, c = lane & 31;
    bf16x8 act[NT][16];
    f32x16 acc[NT][8];
    input_layer<NT>(x, A.pack0, A.params + oB0, acc);
    finish_hidden<NT, kTrain>(A, 0, ray0, acc, act);
#pragma unroll 1
    for (int layer = 1; layer <= 3; ++layer) {
        const __bf16* P = A.packH + static_cast<long long>(layer - 1) * kPackH;
        const float* bias = A.params + bias_offset(layer);
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
            const f32x16 b = bias_tile(bias, mt, g, kHid);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][mt] = b;
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const bf16x8 w = ld8(P + ((mt * 16 + s) * 64 + lane) * 8);
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t][mt] = mfma(w, act[t][s], acc[t][mt]);
            }
        }
        if (layer < 3) finish_hidden<NT, kTrain>(A, layer, ray0, acc, act);
    }
    finish_last_hidden<NT, kTrain>(A, ray0, acc);
    f32x16 o[NT];
    {
        const f32x16 b = bias_tile(A.params + oB4, 0, g, kOut);
#pragma unroll
        for (int t = 0; t < NT; ++t) o[t] = b;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const bf16x8 w = ld8(A.pack4 + (s * 64 + lane) * 8);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                bf16x8 hi, lo;
                hilo_operand<NT>(acc, t, s, hi, lo);
                o[t] = mfma(w, hi, o[t]);
                o[t] = mfma(w, lo, o[t]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int m = acc_row(g, v);
            if (m < kOut) outs[t * 32 + c][m] = o[t][v];
        }
    }
}

// Plain and training forward over rows of X; kNT rays tiles per wave.
template <int kNT, bool kTrain>
__global__ __launch_bounds__(256) void mlp_forward_kernel(MlpArgs A) {
    __shared__ float outs[4][32 * kNT][kOut];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 31;
    const long long ray0 = (static_cast<long long>(blockIdx.x) * 4 + wave) * (32 * kNT);
    float x[kNT][kIn];
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
        long long i = ray0 + t * 32 + c;
        if (i >= A.n) i = A.n - 1;
        const long long row = data_row(A, i);
#pragma unroll
        for (int k = 0; k < kIn; ++k) x[t][k] = A.X[row * kIn + k];
    }
    mlp_run<kNT, kTrain>(A, x, ray0, outs[wave]);
    __syncthreads();
    for (int r = lane; r < 32 * kNT; r += 64) {
        const long long i = ray0 + r;
        if (i < A.n)
            for (int m = 0; m < kOut; ++m) A.Y[i * kOut + m] = outs[wave][r][m];
    }
}

__device__ __forceinline__ uint32_t hash32(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return static_cast<uint32_t>(x);
}

__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

// Wave-aggregated append: one atomic per wave.
__device__ __forceinline__ long long wave_append(bool want, unsigned long long* counter) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0ull) return -1;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned long long base = 0;
    if (static_cast<int>(threadIdx.x & 63) == leader) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader);
    return want ? static_cast<long long>(base + lanes_below(mask)) : -1;
}

// Surrogate frame pass.  Persistent blocks of 4 waves (two blocks per CU);
// each wave owns 64 rays of a 256-ray tile: their inputs are computed one
// ray per lane, the network runs as two passes of 32 rays (the MFMA tile
// width), and the per-ray work after it (gates, judge selection, shading)
// is done one ray per lane again.  Hidden-layer weights stream
// through a double-buffered 16 KB LDS window, one 32-neuron block (16
// k-steps) at a time, shared by the block's four waves; the layer-0 and
// output-layer weights and all biases stay in LDS.  Every finished 32-neuron
// block goes through SiLU while the MFMAs of the next block run (the two
// are independent, so the SIMD overlaps them), and the output layer is
// accumulated block by block inside the last hidden layer.  Per ray: inputs
// from the view's sample position, the network, then either shading from
// the prediction or a place in the fallback list; a hashed fraction of the
// shaded rays is listed for the FP64 judge.
constexpr int kInferTile = 256;              // rays per block tile: 64 per wave, two 32-ray passes
constexpr int kChunkVec = 16 * 64;          // uint4 (8 BF16) per 16 KB weight window
constexpr int kChunks = 3 * 8;              // hidden (layer, 32-neuron block) windows

__device__ __forceinline__ const uint4* chunk_src(const MlpArgs& A, int c) {
    return reinterpret_cast<const uint4*>(A.packH + static_cast<long long>(c) * kChunkVec * 8);
}

// Copy a 16 KB weight window global -> LDS with the direct-to-LDS loads
// (no staging registers): wave w of the block moves four 1 KB rows, lane
// l's 16 bytes landing at row base + 16 l.  Completion is awaited with
// s_waitcnt before the block barrier that publishes the window.
__device__ __forceinline__ void load_window(const uint4* src, uint4* dst) {
    // The window's base address stays an opaque scalar here, so the
    // compiler does not precompute (and spill) every window's per-lane
    // addresses ahead of the network.
    asm volatile("" : "+s"(src));
    const int tid = threadIdx.x, wbase = tid & ~63;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)(src + tid + 256 * k),
                                         (__attribute__((address_space(3))) void*)(dst + 256 * k + wbase), 16, 0, 0);
}

// Scheduling pattern for one weight window: each MFMA is followed by the LDS
// read of a later k-step's weights and a share of the previous block's SiLU
// (VALU), so LDS latency and SiLU work hide under the MFMA chain.
template <int valu_per_mfma>
__device__ __forceinline__ void interleave_window() {
    __builtin_amdgcn_sched_group_barrier(0x100, 7, 0);       // DS read: 4 bias + 3 weights (prefetch distance 3)
#pragma unroll
    for (int s = 0; s < 13; ++s) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);   // DS read
        __builtin_amdgcn_sched_group_barrier(0x002, valu_per_mfma, 0);   // VALU
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, valu_per_mfma, 0);
    }
}

// Ray visited at position l of an inference launch: in order, or row by
// row from the image centre outwards (H/2, H/2-1, H/2+1, ...) so the rays
// around the shadow, whose fallback traces are the longest, are queued for
// the integrator first.
//
// The visit positions of a wavefront's lanes are p0 + off with p0 the same
// for the whole wavefront and off < 64: one 64-bit division (on the scalar
// unit) per wavefront, then each lane steps across row ends by subtraction.
__device__ __forceinline__ long long visit_ray(const InferArgs& IA, long long p0, int off) {
    const long long p0u = static_cast<long long>(
        (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(static_cast<int>(p0 >> 32))) << 32) |
        static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(p0))));
    if (IA.row_rays <= 0) return IA.ray_base + p0u + off;
    const long long rr = IA.row_rays;
    long long k = p0u / rr;
    long long col = p0u - k * rr + off;
    while (col >= rr) {
        col -= rr;
        ++k;
    }
    const long long half = IA.rows / 2;
    const long long row = (k & 1) ? half - (k + 1) / 2 : half + k / 2;
    return IA.ray_base + row * rr + col;
}

// decode_ray for a whole-view window (no tile order) in 32-bit arithmetic
// while the ray index fits; the general 64-bit form otherwise.
__device__ __forceinline__ void decode_view_ray(const ViewParams& V, long long q, SampleId& d) {
    if (V.tile <= 0 && q >= 0 && q < 0x7fffffffLL) {
        const unsigned uq = static_cast<unsigned>(q), nS = static_cast<unsigned>(V.nS);
        const unsigned tw = static_cast<unsigned>(V.tw);
        const unsigned pl = uq / nS;
        const unsigned pj = pl / tw;
        d.s = static_cast<int>(uq - pl * nS);
        d.i = V.x0 + static_cast<int>(pl - pj * tw);
        d.j = V.y0 + static_cast<int>(pj);
        d.local = q;
    } else {
        decode_ray(V, q, d);
    }
}

// One 32-neuron block of a layer from a weight window: the bias, then the
// 16 k-step MFMA chain with the window's reads issued three k-steps ahead.
// Distance of a null ray from the critical curve (the edge of the shadow).
// With E = -p_t, L = p_phi, b = L/E and q = Q/E^2 (Carter's constant
// Q = p_theta^2 + cos^2 theta (L^2 / sin^2 theta - a^2 E^2), conserved, so
// the camera's covector fixes it), the radial potential of Kerr null
// geodesics in units of E^2 is
//   R(r) = (r^2 + a^2 - a b)^2 - Delta(r) k,  k = q + (b - a)^2.
// A ray on the critical curve has a double root of R at a spherical photon
// orbit, whose radii run from the prograde to the retrograde equatorial
// photon orbit (r0..r1).  The score is the minimum of R(r) / r^4 over that
// range: 0 on the critical curve, > 0 for rays with no turning point there
// (captured), < 0 for rays that turn outside it.  In x = 1/r,
//   R / r^4 = 1 + c2 x^2 + c1 x^3 + c0 x^4,
//   c2 = 2 A - k, c1 = 2 M k, c0 = A^2 - a^2 k, A = a^2 - a b,
// a quartic whose stationary points other than x = 0 solve
// 4 c0 x^2 + 3 c1 x + 2 c2 = 0: the minimum is the least of the quartic at
// the range's ends and at those roots that fall inside it.
// sth, cth: sin and cos of the ray's polar angle theta.
__device__ double critical_score(double a, double M, double P0, double P3, double pth, double sth, double cth,
                                 double r0, double r1) {
    const double b = -P3 / P0;
    const double s = fmax(fabs(sth), 1e-10), c = cth;
    const double pe = pth / P0, a2 = a * a;
    const double k = pe * pe + c * c * (b * b / (s * s) - a2) + (b - a) * (b - a);
    const double A = a2 - a * b;
    const double c2 = 2.0 * A - k, c1 = 2.0 * M * k, c0 = A * A - a2 * k;
    auto f = [&](double x) { return 1.0 + x * x * (c2 + x * (c1 + x * c0)); };
    const double xl = 1.0 / r1, xh = 1.0 / r0;
    double lo = fmin(f(xl), f(xh));
    const double qa = 4.0 * c0, qb = 3.0 * c1, qc = 2.0 * c2;
    const double disc = qb * qb - 4.0 * qa * qc;
    if (disc >= 0.0) {
        // the two roots without cancellation: qq / qa and qc / qq
        const double qq = -0.5 * (qb + copysign(sqrt(disc), qb));
        const double xs[2] = {qa != 0.0 ? qq / qa : -1.0, qq != 0.0 ? qc / qq : -1.0};
        for (double x : xs)
            if (x >= xl && x <= xh) lo = fmin(lo, f(x));
    }
    return lo;
}

// Critical-curve flags, one lane per sample: the camera covector of
// camera_ray (its two pixel angles' sin/cos by the branch-free Cephes form;
// the gate needs ~1e-9, not the integrator's last bit), then the score.
__global__ __launch_bounds__(256) void critical_flags_kernel(const TraceParams* __restrict__ Pg, long long n,
                                                              double score_max, double r0, double r1, double sth,
                                                              double cth, uint8_t* __restrict__ flags) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ViewParams& V = Pg->view;
    SampleId d;
    decode_view_ray(V, i, d);
    double px, py, sa, ca_, sb, cb;
    sample_position(V, d, px, py);
    sincos_cephes((px - V.W / 2.0) * V.fov / V.W, sa, ca_);
    sincos_cephes((py - V.H / 2.0) * V.fov / V.W, sb, cb);
    const double ca = cb * sa, cc = cb * ca_;
    const double p0 = -V.e0t + (-ca * V.e1[0] - sb * V.e2[0] + cc * V.e3[0]);
    const double p2 = -ca * V.e1[2] - sb * V.e2[2] + cc * V.e3[2];
    const double p3 = -ca * V.e1[3] - sb * V.e2[3] + cc * V.e3[3];
    const double P0 = V.g00 * p0 + V.g03 * p3, P3 = V.g03 * p0 + V.g33 * p3;
    flags[i] = fabs(critical_score(V.a, V.M, P0, P3, V.g22 * p2, sth, cth, r0, r1)) < score_max ? 1 : 0;
}

__device__ __forceinline__ f32x16 window_block(const bf16x8* W, const bf16x8 (&in)[16], const float* bias, int mt,
                                               int g, int lane) {
    f32x16 a = bias_tile(bias, mt, g, kHid);
    const bf16x8* Wl = W + lane;
    bf16x8 w0 = Wl[0], w1 = Wl[64], w2 = Wl[128], w3;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        if (s + 3 < 16) w3 = Wl[(s + 3) * 64];
        a = mfma(w0, in[s], a);
        w0 = w1;
        w1 = w2;
        w2 = w3;
    }
    return a;
}

// SiLU of one 32-neuron accumulator block into B-operand slots 2mt, 2mt+1.
__device__ __forceinline__ void silu_block(const f32x16& a, bf16x8& lo_slot, bf16x8& hi_slot) {
#pragma unroll
    for (int j = 0; j < 8; ++j) lo_slot[j] = static_cast<__bf16>(silu(a[j]));
#pragma unroll
    for (int j = 0; j < 8; ++j) hi_slot[j] = static_cast<__bf16>(silu(a[8 + j]));
}

__global__ __launch_bounds__(256, 2) void mlp_infer_kernel(InferArgs IA) {
    // The two weight-window buffers are separate LDS variables: the compiler
    // gives each its own alias scope, so reads of the current window are not
    // made to wait for the direct-to-LDS loads filling the other one (with a
    // single [2][n] array indexed by a run-time parity it waits, and every
    // window paid the full L2 load latency).  Window c (0..23 in a pass, an
    // even count) reads buffer c % 2 and fills the other.
    __shared__ uint4 wbuf0[kChunkVec];
    __shared__ uint4 wbuf1[kChunkVec];
    __shared__ uint4 s_pack0[kPack0 / 8];
    __shared__ uint4 s_pack4[kPack4 / 8];
    __shared__ float s_bias[4 * kHid + 16];
    __shared__ float outs[4][64][kOut];
    __shared__ float s_x[4][64][kIn];
    __shared__ double s_flux[kFluxN];
    const MlpArgs& A = IA.mlp;
    const TraceParams& P = *IA.P;
    const ViewParams& V = P.view;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, g = lane >> 5, c = lane & 31;
    for (int i = tid; i < kFluxN; i += 256) s_flux[i] = V.flux_lut[i];
    for (int i = tid; i < kPack0 / 8; i += 256) s_pack0[i] = reinterpret_cast<const uint4*>(A.pack0)[i];
    for (int i = tid; i < kPack4 / 8; i += 256) s_pack4[i] = reinterpret_cast<const uint4*>(A.pack4)[i];
    for (int i = tid; i < 4 * kHid; i += 256) s_bias[i] = A.params[bias_offset(i / kHid) + i % kHid];
    if (tid < 16) s_bias[4 * kHid + tid] = tid < kOut ? A.params[oB4 + tid] : 0.0f;
    const long long n = A.n;
    const long long n_tiles = (n + kInferTile - 1) / kInferTile;
    // Tiles are claimed dynamically (one claim ahead), so blocks that start
    // late, while other kernels hold part of the chip, simply take fewer.
    __shared__ long long s_claim;
    if (tid == 0) s_claim = static_cast<long long>(atomicAdd(IA.tile_counter, 1ull));
    __syncthreads();
    long long tile = s_claim;
    if (tile < n_tiles) load_window(chunk_src(A, 0), wbuf0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    while (tile < n_tiles) {
        if (tid == 0) s_claim = static_cast<long long>(atomicAdd(IA.tile_counter, 1ull));
        bool more_tiles = true;         // read from s_claim after the first window's barrier
        const long long ray0 = tile * kInferTile + wave * 64;
        {
            // Inputs of this lane's ray (ray0 + lane), once per 64 rays.
            // View constants are re-read per tile (cached) rather than held
            // in registers across the network.
            asm volatile("" ::: "memory");
            // Lanes past the end repeat the last ray.
            const long long p0 = ray0 < n ? ray0 : n - 1;
            const int off = static_cast<int>((ray0 + lane < n ? ray0 + lane : n - 1) - p0);
            const long long vi = visit_ray(IA, p0, off);
            float xl[kIn];
            if (IA.x_in) {
#pragma unroll
                for (int k = 0; k < kIn; ++k) xl[k] = IA.x_in[vi * kIn + k];
            } else {
                SampleId d;
                decode_view_ray(V, vi, d);
                double px, py;
                sample_position(V, d, px, py);
                xl[0] = static_cast<float>((px - 0.5 * V.W) * IA.feat_scale);
                xl[1] = static_cast<float>((py - 0.5 * V.H) * IA.feat_scale);
                xl[2] = IA.feat_view[0];
                xl[3] = IA.feat_view[1];
                xl[4] = IA.feat_view[2];
            }
#pragma unroll
            for (int k = 0; k < kIn; ++k) s_x[wave][lane][k] = xl[k];
            __builtin_amdgcn_wave_barrier();
        }
        // Two passes of the network, rays ray0 + 32 t + c; the window
        // stream continues from one pass into the next.
#pragma unroll 1
        for (int t = 0; t < 2; ++t) {
        float x[1][kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) x[0][k] = s_x[wave][t * 32 + c][k];
        bf16x8 actA[16], actB[16];
        // Layer 0: one k-step per 32-neuron block.
        {
            bf16x8 in0;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int f = g * 8 + j;
                float v = 0.0f;
                if (f < 2 * kIn) {
                    const float xi = x[0][f >> 1];
                    const float hi = bf16r(xi);
                    v = (f & 1) ? xi - hi : hi;
                }
                in0[j] = static_cast<__bf16>(v);
            }
            const bf16x8* w0 = reinterpret_cast<const bf16x8*>(s_pack0);
#pragma unroll
            for (int mt = 0; mt < 8; ++mt) {
                const f32x16 a = mfma(w0[mt * 64 + lane], in0, bias_tile(s_bias, mt, g, kHid));
                silu_block(a, actA[2 * mt], actA[2 * mt + 1]);
            }
        }
        // One 16 KB weight window: start the copy of the next window into
        // the other buffer, run this window's block (and the SiLU of the
        // previous block), then wait for the copy and publish it.
        auto window = [&](int cidx, auto&& body) {
            if (cidx == kChunks - 1 && t == 1) more_tiles = s_claim < n_tiles;
            const bool has_next = cidx + 1 < kChunks || t == 0 || more_tiles;
            static_assert(kChunks % 2 == 0, "window parity must repeat from pass to pass");
            uint4* const cur = (cidx & 1) ? wbuf1 : wbuf0;
            uint4* const nxt = (cidx & 1) ? wbuf0 : wbuf1;
            if (has_next) load_window(chunk_src(A, cidx + 1 < kChunks ? cidx + 1 : 0), nxt);
            body(reinterpret_cast<const bf16x8*>(cur));
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        };
        auto hidden = [&](int layer, const bf16x8 (&in)[16], bf16x8 (&out)[16]) {
            f32x16 prev;
#pragma unroll
            for (int mt = 0; mt < 8; ++mt) {
                window((layer - 1) * 8 + mt, [&](const bf16x8* W) {
                    const f32x16 a = window_block(W, in, s_bias + layer * kHid, mt, g, lane);
                    if (mt > 0) silu_block(prev, out[2 * mt - 2], out[2 * mt - 1]);
                    prev = a;
                    interleave_window<8>();
                });
            }
            silu_block(prev, out[14], out[15]);
        };
        hidden(1, actA, actB);
        hidden(2, actB, actA);
        // Layer 3 with the output layer folded in: block mt's FP32
        // activations enter output k-steps 2mt, 2mt+1 as BF16 hi + lo.
        f32x16 o = bias_tile(s_bias + 4 * kHid, 0, g, kOut);
        {
            const bf16x8* w4 = reinterpret_cast<const bf16x8*>(s_pack4);
            auto out_block = [&](const f32x16& a, int mt) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    bf16x8 hi, lo;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const float h = silu(a[q * 8 + j]);
                        const __bf16 hh = static_cast<__bf16>(h);
                        hi[j] = hh;
                        lo[j] = static_cast<__bf16>(h - static_cast<float>(hh));
                    }
                    const bf16x8 w = w4[(2 * mt + q) * 64 + lane];
                    o = mfma(w, hi, o);
                    o = mfma(w, lo, o);
                }
            };
            f32x16 prev;
#pragma unroll
            for (int mt = 0; mt < 8; ++mt) {
                window(16 + mt, [&](const bf16x8* W) {
                    const f32x16 a = window_block(W, actA, s_bias + 3 * kHid, mt, g, lane);
                    if (mt > 0) out_block(prev, mt - 1);
                    prev = a;
                    interleave_window<10>();
                });
            }
            out_block(prev, 7);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int m = acc_row(g, v);
            if (m < kOut) outs[wave][t * 32 + c][m] = o[v];
        }
        }   // the two network passes
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        // One lane per ray: ray ray0 + lane.
        const bool valid = ray0 + lane < n;
        const long long i = visit_ray(IA, ray0 < n ? ray0 : 0, valid ? lane : 0);
        float y[kOut];
#pragma unroll
        for (int m = 0; m < kOut; ++m) y[m] = outs[wave][lane][m];
        __builtin_amdgcn_wave_barrier();
        if (IA.y_raw) {
            if (valid)
#pragma unroll
                for (int m = 0; m < kOut; ++m) IA.y_raw[i * kOut + m] = y[m];
            const long long nxt = s_claim;
            __syncthreads();
            tile = nxt;
            continue;
        }
        const float mx = fmaxf(y[0], fmaxf(y[1], y[2]));
        const float e0 = __expf(y[0] - mx), e1 = __expf(y[1] - mx), e2 = __expf(y[2] - mx);
        const float inv = 1.0f / (e0 + e1 + e2);
        const float p0 = e0 * inv, p1 = e1 * inv, p2 = e2 * inv;
        auto plogp = [](float p) { return p > 0.0f ? p * __logf(p) : 0.0f; };
        const float H = -(plogp(p0) + plogp(p1) + plogp(p2));
        const int cls = (p2 >= p1 && p2 >= p0) ? GRRT_CLASS_DISK : (p1 >= p0 ? GRRT_CLASS_ESCAPED : GRRT_CLASS_CAPTURED);
        // A disk ray's redshift follows exactly from the crossing radius and
        // the conserved p_t, p_phi of its camera ray (src/accretion_disk.cpp:
        // 113-130): shade with that g, and send the ray to the integrator when
        // the network's own ln g disagrees with it by more than lng_tol.
        // Rays near the critical curve (the shadow's edge) go to the
        // integrator whatever the network predicts (flags written before
        // the pass).
        double rc = 0.0, g_exact = 0.0;
        bool inconsistent = false;
        const bool near_critical = valid && IA.crit_flags && IA.crit_flags[i] != 0;
        if (valid && cls == GRRT_CLASS_DISK) {
            SampleId d;
            decode_view_ray(V, i, d);
            double P0, P3;
            Vec5 y0;
            camera_ray(V, d, P0, P3, y0);
            rc = exp(static_cast<double>(y[3]));
            g_exact = disk_redshift(V, rc, P0, P3);
            inconsistent = IA.lng_tol > 0.0 &&
                           !(fabs(log(fmax(g_exact, 1e-300)) - static_cast<double>(y[4])) <= IA.lng_tol);
        }
        const bool fb = valid && (static_cast<double>(H) > IA.entropy_max || inconsistent || near_critical);
        const long long pf = wave_append(fb, &IA.counts[0]);
        if (fb) IA.fb_list[pf] = i;
        const bool shade = valid && !fb;
        const bool judged = shade && hash32(static_cast<uint64_t>(i) ^ (static_cast<uint64_t>(IA.frame_seed) << 40)) <
                                         IA.judge_threshold;
        const long long pj = wave_append(judged, &IA.counts[1]);
        if (judged && pj < IA.judge_cap) IA.judge_list[pj] = i;
        if (shade) {
            double col[3] = {0.0, 0.0, 0.0};
            if (cls == GRRT_CLASS_DISK) {
                thin_disk_color(V, s_flux, P.spec_lut, rc, g_exact, col);
            } else if (cls == GRRT_CLASS_ESCAPED && (P.stars || P.bg_tex)) {
                const double th = atan2(static_cast<double>(y[5]), static_cast<double>(y[6]));
                const double ph = atan2(static_cast<double>(y[7]), static_cast<double>(y[8])) + V.ph_obs;
                if (P.stars) {
                    const double bb = star_sample(P, th, ph);
                    col[0] += bb;
                    col[1] += bb;
                    col[2] += bb;
                } else {
                    const TexRGB t = texture_sample(P, th, ph);
                    col[0] += t.c[0];
                    col[1] += t.c[1];
                    col[2] += t.c[2];
                }
            }
            P.sample_rgb[3 * i + 0] = col[0];
            P.sample_rgb[3 * i + 1] = col[1];
            P.sample_rgb[3 * i + 2] = col[2];
            P.sample_cls[i] = static_cast<uint8_t>(cls);
        }
        const long long next = s_claim;
        __syncthreads();                // everyone has read the claim before it is replaced
        tile = next;
    }
}

// BF16 operand images and transposed hidden weights from the FP32 master.
__global__ void pack_kernel(const float* __restrict__ w, __bf16* pack0, __bf16* packH, __bf16* pack4,
                            __bf16* WT) {
    long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < kPack0) {
        const int j = idx & 7, lane = (idx >> 3) & 63, mt = static_cast<int>(idx >> 9);
        const int m = mt * 32 + (lane & 31), f = (lane >> 5) * 8 + j;
        pack0[idx] = static_cast<__bf16>(f < 2 * kIn ? w[oW0 + m * kIn + (f >> 1)] : 0.0f);
        return;
    }
    idx -= kPack0;
    if (idx < 3 * kPackH) {
        const int layer = 1 + static_cast<int>(idx / kPackH);
        const int e = static_cast<int>(idx % kPackH);
        const int j = e & 7, lane = (e >> 3) & 63, s = (e >> 9) & 15, mt = e >> 13;
        const int m = mt * 32 + (lane & 31), k = kslot_neuron(s, lane >> 5, j);
        packH[idx] = static_cast<__bf16>(w[weight_offset(layer) + m * kHid + k]);
        return;
    }
    idx -= 3 * kPackH;
    if (idx < kPack4) {
        const int e = static_cast<int>(idx);
        const int j = e & 7, lane = (e >> 3) & 63, s = e >> 9;
        const int m = lane & 31, k = kslot_neuron(s, lane >> 5, j);
        pack4[idx] = static_cast<__bf16>(m < kOut ? w[oW4 + m * kHid + k] : 0.0f);
        return;
    }
    idx -= kPack4;
    if (idx < 3 * kHid * kHid) {
        const int layer = 1 + static_cast<int>(idx / (kHid * kHid));
        const int e = static_cast<int>(idx % (kHid * kHid));
        const int k = e / kHid, nn = e % kHid;     // WT[k][n] = W[n][k]
        WT[idx] = static_cast<__bf16>(w[weight_offset(layer) + nn * kHid + k]);
    }
}

// Loss and its gradient with respect to the 9 outputs (mean over the batch):
// softmax cross-entropy on the class logits; squared error on ln r, ln g of
// disk rays and on the escape direction of escaped rays.
// acc: [0] cross-entropy, [1] correct classes, [2] sum sq ln r, [3] sum sq
// ln g, [4] disk rays, [5] sum sq direction, [6] escaped rays, [7] rays.
__global__ __launch_bounds__(256) void loss_grad_kernel(const float* __restrict__ Y, const Target* __restrict__ T,
                                                        MlpArgs rows, LossWeights w, float* __restrict__ dY,
                                                        double* acc) {
    const long long b = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long B = rows.n;
    double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (b < B) {
        const Target t = T[data_row(rows, b)];
        const float* y = Y + b * kOut;
        float* d = dY + b * kOut;
        const float invB = 1.0f / static_cast<float>(B);
        for (int m = 0; m < kOut; ++m) d[m] = 0.0f;
        const int cls = static_cast<int>(t.cls);
        if (cls >= 0) {
            const float mx = fmaxf(y[0], fmaxf(y[1], y[2]));
            const float e[3] = {expf(y[0] - mx), expf(y[1] - mx), expf(y[2] - mx)};
            const float S = e[0] + e[1] + e[2];
            for (int k = 0; k < 3; ++k) d[k] = (e[k] / S - (k == cls ? 1.0f : 0.0f)) * invB;
            st[0] = -(static_cast<double>(y[cls] - mx) - log(static_cast<double>(S)));
            const int am = (y[2] >= y[1] && y[2] >= y[0]) ? 2 : (y[1] >= y[0] ? 1 : 0);
            st[1] = am == cls ? 1.0 : 0.0;
            st[7] = 1.0;
            if (cls == 2) {
                const float er = y[3] - t.lnr, eg = y[4] - t.lng;
                d[3] = 2.0f * w.w_r * er * invB;
                d[4] = 2.0f * w.w_g * eg * invB;
                st[2] = static_cast<double>(er) * er;
                st[3] = static_cast<double>(eg) * eg;
                st[4] = 1.0;
            }
            if (t.esc > 0.5f) {
                double s = 0.0;
                for (int k = 0; k < 4; ++k) {
                    const float e2 = y[5 + k] - t.dir[k];
                    d[5 + k] = 2.0f * w.w_dir * e2 * invB;
                    s += static_cast<double>(e2) * e2;
                }
                st[5] = s;
                st[6] = 1.0;
            }
        }
    }
    for (int k = 0; k < 8; ++k) {
        double v = st[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(&acc[k], v);
    }
}

// dZ = dH * silu'(Z) for a 64-ray x 64-neuron tile (inputs neuron-major):
// BF16 copies in neuron- and batch-major order for the two backward GEMMs,
// FP32 per-block sums over the rays for the bias gradient (dbp[block][256]).
// With dY given, dH = dY W4 (BF16 weights as in the forward pass) is formed
// here.
__global__ __launch_bounds__(256) void make_dz_kernel(const float* __restrict__ dHT, const float* __restrict__ dY,
                                                      const float* __restrict__ w, const float* __restrict__ ZT,
                                                      long long B, __bf16* __restrict__ dZbm,
                                                      __bf16* __restrict__ dZnm, float* __restrict__ dbp) {
    __shared__ __bf16 tile[64][66];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const long long b = static_cast<long long>(blockIdx.x) * 64 + tx;
    const int nb = blockIdx.y * 64;
    float dy[kOut];
    if (dY)
        for (int o = 0; o < kOut; ++o) dy[o] = dY[b * kOut + o];
    for (int nn = ty; nn < 64; nn += 4) {
        const int n = nb + nn;
        float dh;
        if (dY) {
            dh = 0.0f;
            for (int o = 0; o < kOut; ++o) dh += dy[o] * bf16r(w[oW4 + o * kHid + n]);
        } else {
            dh = dHT[static_cast<long long>(n) * B + b];
        }
        const float dz = dh * dsilu(ZT[static_cast<long long>(n) * B + b]);
        const __bf16 q = static_cast<__bf16>(dz);
        dZnm[static_cast<long long>(n) * B + b] = q;
        tile[nn][tx] = q;
        float sum = dz;
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
        if (tx == 0) dbp[static_cast<long long>(blockIdx.x) * kHid + n] = sum;
    }
    __syncthreads();
    for (int rr = ty; rr < 64; rr += 4)
        dZbm[(static_cast<long long>(blockIdx.x) * 64 + rr) * kHid + nb + tx] = tile[tx][rr];
}

// Output-layer gradients: dW4[o][k] = sum_b dY[b][o] H3[b][k] (block k < 256)
// and db4[o] = sum_b dY[b][o] (block 256).
__global__ __launch_bounds__(256) void out_layer_grads_kernel(const float* __restrict__ dY,
                                                              const __bf16* __restrict__ H3T, long long B,
                                                              float* __restrict__ grad) {
    __shared__ float red[kOut][256];
    const int k = blockIdx.x;
    float s[kOut];
    for (int o = 0; o < kOut; ++o) s[o] = 0.0f;
    for (long long b = threadIdx.x; b < B; b += 256) {
        const float h = k < kHid ? static_cast<float>(H3T[static_cast<long long>(k) * B + b]) : 1.0f;
        for (int o = 0; o < kOut; ++o) s[o] += dY[b * kOut + o] * h;
    }
    for (int o = 0; o < kOut; ++o) red[o][threadIdx.x] = s[o];
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if (static_cast<int>(threadIdx.x) < stp)
            for (int o = 0; o < kOut; ++o) red[o][threadIdx.x] += red[o][threadIdx.x + stp];
        __syncthreads();
    }
    if (threadIdx.x < kOut) {
        const int o = threadIdx.x;
        if (k < kHid) grad[oW4 + o * kHid + k] = red[o][0];
        else grad[oB4 + o] = red[o][0];
    }
}

// C[m][n] (+ split z at C + z*M*ldc) = sum_k A[m][k] Bm[n][k] over the
// split's K range; one 32x32 tile per wave, 2x2 tiles per block.
__global__ __launch_bounds__(256) void gemm_nt_kernel(int M, int N, int K, const __bf16* __restrict__ A,
                                                      long long lda, const __bf16* __restrict__ Bm, long long ldb,
                                                      float* __restrict__ C, long long ldc, int splits) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 5, r = lane & 31;
    const int m0 = blockIdx.y * 64 + (wave >> 1) * 32, n0 = blockIdx.x * 64 + (wave & 1) * 32;
    if (m0 >= M || n0 >= N) return;
    const int Ks = K / splits, k0 = blockIdx.z * Ks;
    const __bf16* pa = A + static_cast<long long>(m0 + r) * lda + k0 + g * 8;
    const __bf16* pb = Bm + static_cast<long long>(n0 + r) * ldb + k0 + g * 8;
    f32x16 acc = {};
    for (int k = 0; k < Ks; k += 16) acc = mfma(ld8(pa + k), ld8(pb + k), acc);
    float* out = C + static_cast<long long>(blockIdx.z) * M * ldc;
#pragma unroll
    for (int v = 0; v < 16; ++v) out[static_cast<long long>(m0 + acc_row(g, v)) * ldc + n0 + r] = acc[v];
}

__global__ void reduce_splits_kernel(const float* __restrict__ part, int splits, long long count,
                                     float* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[static_cast<long long>(z) * count + i];
    out[i] = s;
}

// Few outputs, many partials: one block per output, threads over the
// partials, then a tree in LDS.
__global__ __launch_bounds__(256) void reduce_splits_wide_kernel(const float* __restrict__ part, int splits,
                                                                 long long count, float* __restrict__ out) {
    __shared__ float red[256];
    const long long i = blockIdx.x;
    float s = 0.0f;
    for (int z = threadIdx.x; z < splits; z += 256) s += part[static_cast<long long>(z) * count + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int stp = 128; stp > 0; stp >>= 1) {
        if (static_cast<int>(threadIdx.x) < stp) red[threadIdx.x] += red[threadIdx.x + stp];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = red[0];
}

// Layer-0 gradients in FP32 for a block of 64 rays: part[block][m*5+i] =
// sum_b dz0[b][m] x[b][i] and part[block][1280+m] = sum_b dz0[b][m] (the
// parameter order of W0 then b0).  dH0T, Z0T neuron-major.
__global__ __launch_bounds__(256) void layer0_grads_kernel(const float* __restrict__ dH0T, const float* __restrict__ Z0T,
                                                           MlpArgs rows, float* __restrict__ part) {
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const long long B = rows.n;
    const long long b = static_cast<long long>(blockIdx.x) * 64 + tx;
    const long long row = data_row(rows, b);
    float x[kIn];
    for (int i = 0; i < kIn; ++i) x[i] = rows.X[row * kIn + i];
    float* out = part + static_cast<long long>(blockIdx.x) * (oB0 + kHid - oW0);
    for (int m = ty; m < kHid; m += 4) {
        const float dz = dH0T[static_cast<long long>(m) * B + b] * dsilu(Z0T[static_cast<long long>(m) * B + b]);
        float v[kIn + 1];
        for (int i = 0; i < kIn; ++i) v[i] = dz * x[i];
        v[kIn] = dz;
        for (int off = 32; off > 0; off >>= 1)
            for (int i = 0; i <= kIn; ++i) v[i] += __shfl_xor(v[i], off);
        if (tx == 0) {
            for (int i = 0; i < kIn; ++i) out[m * kIn + i] = v[i];
            out[kHid * kIn + m] = v[kIn];
        }
    }
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float bc1,
                            float bc2) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// Training rows from the FP64 integrator's per-ray records of a batch of
// views (records in each view's pixel-major sample order).
__global__ void targets_kernel(const GRRTRayRecord* __restrict__ rec, const ViewParams* __restrict__ views,
                               int n_views, long long n_rays, float* __restrict__ X, Target* __restrict__ T) {
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n_rays) return;
    int lo = 0, hi = n_views - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].ray_offset <= k) lo = mid;
        else hi = mid - 1;
    }
    const ViewParams& V = views[lo];
    const long long local = k - V.ray_offset;
    const long long p = local / V.nS;
    SampleId d;
    d.s = static_cast<int>(local - p * V.nS);
    d.i = V.x0 + static_cast<int>(p % V.tw);
    d.j = V.y0 + static_cast<int>(p / V.tw);
    d.local = local;
    double px, py;
    sample_position(V, d, px, py);
    const double alpha = (px - V.W / 2.0) * V.fov / V.W;
    const double beta = (py - V.H / 2.0) * V.fov / V.W;
    float x[kIn];
    features(alpha, beta, V.a / V.M, cos(V.th_obs), V.r_obs / V.M, x);
    for (int i = 0; i < kIn; ++i) X[k * kIn + i] = x[i];
    const GRRTRayRecord& r = rec[k];
    Target t;
    t.cls = r.cls == GRRT_CLASS_MAXSTEPS ? -1.0f : static_cast<float>(r.cls);
    t.lnr = 0.0f;
    t.lng = 0.0f;
    if (r.cls == GRRT_CLASS_DISK) {
        if (r.r_hit > 0.0 && r.g_hit > 0.0) {
            t.lnr = static_cast<float>(log(r.r_hit));
            t.lng = static_cast<float>(log(r.g_hit));
        } else {
            t.cls = -1.0f;
        }
    }
    t.esc = r.term == 1 ? 1.0f : 0.0f;
    const double th = r.x_final[2], ph = r.x_final[3] - V.ph_obs;
    t.dir[0] = static_cast<float>(sin(th));
    t.dir[1] = static_cast<float>(cos(th));
    t.dir[2] = static_cast<float>(sin(ph));
    t.dir[3] = static_cast<float>(cos(ph));
    T[k] = t;
}

// Judge: the listed shaded rays against their FP64 integrator result
// (written after the frame's n_rays samples).  rel_err: relative error of
// the summed RGB, -1 where both are dark.
__global__ void judge_kernel(const long long* __restrict__ judge_list, const unsigned long long* __restrict__ counts,
                             long long judge_cap, const double* __restrict__ rgb, const uint8_t* __restrict__ cls,
                             long long n_rays, float* __restrict__ rel_err, unsigned long long* agree) {
    const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long nj = static_cast<long long>(counts[1] < static_cast<unsigned long long>(judge_cap)
                                                    ? counts[1] : static_cast<unsigned long long>(judge_cap));
    bool same = false;
    if (j < nj) {
        const long long id = judge_list[j];
        const long long ref = n_rays + j;
        same = cls[id] == cls[ref];
        const double ls = rgb[3 * id] + rgb[3 * id + 1] + rgb[3 * id + 2];
        const double lr = rgb[3 * ref] + rgb[3 * ref + 1] + rgb[3 * ref + 2];
        float e = -1.0f;
        if (lr > 0.0) e = static_cast<float>(fabs(ls - lr) / lr);
        else if (ls > 0.0) e = 1.0f;
        rel_err[j] = e;
    }
    const unsigned long long m = __ballot(same);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(agree, static_cast<unsigned long long>(__popcll(m)));
}

__global__ void reset_counts_kernel(unsigned long long* c, int n) {
    if (static_cast<int>(threadIdx.x) < n) c[threadIdx.x] = 0ull;
}

unsigned blocks_for(long long n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

hipError_t launch_pack(const float* params, __bf16* pack0, __bf16* packH, __bf16* pack4, __bf16* WT,
                       hipStream_t st) {
    const long long total = kPack0 + 3LL * kPackH + kPack4 + 3LL * kHid * kHid;
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, params, pack0, packH, pack4, WT);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const MlpArgs& a, bool train, hipStream_t st) {
    const dim3 grid(blocks_for(a.n, 4 * 32));
    if (train) hipLaunchKernelGGL((mlp_forward_kernel<1, true>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((mlp_forward_kernel<1, false>), grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_critical_flags(const TraceParams* P, long long n, double score_max, double r0, double r1,
                                 double sth, double cth, uint8_t* flags, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(critical_flags_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, P, n, score_max, r0, r1,
                       sth, cth, flags);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const InferArgs& a, int num_cus, hipStream_t st) {
    if (a.mlp.n <= 0) return hipSuccess;
    const long long tiles = (a.mlp.n + kInferTile - 1) / kInferTile;
    const long long grid = std::min<long long>(tiles, 2LL * num_cus);
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(static_cast<unsigned>(std::max<long long>(grid, 1))), dim3(256), 0, st,
                       a);
    return hipGetLastError();
}

hipError_t launch_loss_grad(const float* Y, const Target* T, const MlpArgs& rows, LossWeights w, float* dY,
                            double* loss_acc, hipStream_t st) {
    hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks_for(rows.n, 256)), dim3(256), 0, st, Y, T, rows, w, dY,
                       loss_acc);
    return hipGetLastError();
}

hipError_t launch_make_dz(const float* dHT, const float* dY, const float* params, const float* ZT, long long B,
                          __bf16* dZbm, __bf16* dZnm, float* dbp, hipStream_t st) {
    if (B % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(make_dz_kernel, dim3(static_cast<unsigned>(B / 64), kHid / 64), dim3(256), 0, st, dHT, dY,
                       params, ZT, B, dZbm, dZnm, dbp);
    return hipGetLastError();
}

hipError_t launch_out_layer_grads(const float* dY, const __bf16* H3T, long long B, float* grad, hipStream_t st) {
    hipLaunchKernelGGL(out_layer_grads_kernel, dim3(kHid + 1), dim3(256), 0, st, dY, H3T, B, grad);
    return hipGetLastError();
}

hipError_t launch_gemm_nt(int M, int N, int K, const __bf16* A, long long lda, const __bf16* Bm, long long ldb,
                          float* C, long long ldc, int splits, hipStream_t st) {
    if (M % 64 || N % 64 || splits < 1 || K % (16 * splits)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gemm_nt_kernel, dim3(N / 64, M / 64, splits), dim3(256), 0, st, M, N, K, A, lda, Bm, ldb, C,
                       ldc, splits);
    return hipGetLastError();
}

hipError_t launch_reduce_splits(const float* part, int splits, long long count, float* out, hipStream_t st) {
    if (count <= 4096 && splits >= 64) {
        hipLaunchKernelGGL(reduce_splits_wide_kernel, dim3(static_cast<unsigned>(count)), dim3(256), 0, st, part,
                           splits, count, out);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(reduce_splits_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, part, splits, count, out);
    return hipGetLastError();
}

hipError_t launch_layer0_grads(const float* dH0T, const float* Z0T, const MlpArgs& rows, float* part,
                               hipStream_t st) {
    if (rows.n % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(layer0_grads_kernel, dim3(static_cast<unsigned>(rows.n / 64)), dim3(256), 0, st, dH0T, Z0T,
                       rows, part);
    return hipGetLastError();
}

hipError_t launch_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t st) {
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, w, g, m, v, n, lr, b1, b2, eps, bc1,
                       bc2);
    return hipGetLastError();
}

hipError_t launch_targets(const GRRTRayRecord* rec, const ViewParams* views, int n_views, long long n_rays, float* X,
                          Target* T, hipStream_t st) {
    hipLaunchKernelGGL(targets_kernel, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st, rec, views, n_views, n_rays,
                       X, T);
    return hipGetLastError();
}

hipError_t launch_judge(const long long* judge_list, const unsigned long long* counts, long long judge_cap,
                        const double* rgb, const uint8_t* cls, long long n_rays, float* rel_err,
                        unsigned long long* agree, hipStream_t st) {
    hipLaunchKernelGGL(judge_kernel, dim3(blocks_for(judge_cap, 256)), dim3(256), 0, st, judge_list, counts, judge_cap,
                       rgb, cls, n_rays, rel_err, agree);
    return hipGetLastError();
}

hipError_t launch_reset_counts(unsigned long long* c, int n, hipStream_t st) {
    hipLaunchKernelGGL(reset_counts_kernel, dim3(1), dim3(64), 0, st, c, n);
    return hipGetLastError();
}

}  // namespace grrt_amd::sg
