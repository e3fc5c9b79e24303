// Learned ray-transfer surrogate: a 5 -> 4x256 (SiLU) -> 9 MLP whose dense
// layers run on BF16 MFMA (v_mfma_f32_32x32x16_bf16) with FP32
// accumulation.  Shared by the host orchestration and the HIP kernels.
//
// Inputs (scaled to about [-1, 1]): pixel direction alpha, beta in the
// camera plane (src/camera.cpp:69-96 angles), spin a, cos(theta_obs),
// log r_obs.  Outputs: 3 class logits (captured, escaped, disk), ln r of the
// first disk crossing, ln g there, sin/cos of the escape theta and phi.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "trace_params.hpp"

namespace grrt_amd::sg {

constexpr int kIn = 5, kHid = 256, kOut = 9;
// FP32 master parameter vector: W0[256][5] b0 W1[256][256] b1 W2 b2 W3 b3 W4[9][256] b4.
constexpr int oW0 = 0, oB0 = oW0 + kHid * kIn;
constexpr int oW1 = oB0 + kHid, oB1 = oW1 + kHid * kHid;
constexpr int oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid, oB3 = oW3 + kHid * kHid;
constexpr int oW4 = oB3 + kHid, oB4 = oW4 + kOut * kHid;
constexpr int kNumParams = oB4 + kOut;
static_assert(kNumParams == 201225, "5-256-256-256-256-9 parameter count");
__host__ __device__ constexpr int weight_offset(int layer) {
    return layer == 0 ? oW0 : layer == 1 ? oW1 : layer == 2 ? oW2 : layer == 3 ? oW3 : oW4;
}
__host__ __device__ constexpr int bias_offset(int layer) {
    return layer == 0 ? oB0 : layer == 1 ? oB1 : layer == 2 ? oB2 : layer == 3 ? oB3 : oB4;
}

// BF16 A-operand images, in the order the kernels load them (one 16-byte
// vector per lane per MFMA): layer 0 [mt 8][lane 64][8], hidden layers
// [mt 8][kstep 16][lane 64][8], output layer [kstep 16][lane 64][8].
constexpr int kPack0 = 8 * 64 * 8;
constexpr int kPackH = 8 * 16 * 64 * 8;
constexpr int kPack4 = 16 * 64 * 8;

// Neuron feeding slot (lane group g, element j) of k-step s: the order in
// which an MFMA accumulator tile holds its rows, so one layer's outputs are
// the next layer's B operand without any data movement.
__host__ __device__ constexpr int kslot_neuron(int s, int g, int j) {
    return (s >> 1) * 32 + (2 * (s & 1) + (j >> 2)) * 8 + g * 4 + (j & 3);
}
// Row of accumulator element v held by lane group g (MFMA 32x32 layout).
__host__ __device__ constexpr int acc_row(int g, int v) { return (v >> 2) * 8 + g * 4 + (v & 3); }

// Training-distribution box (a in [0, 0.998], theta in [5, 89] deg, r in
// [20, 100], directions over a 90 deg field) mapped to [-1, 1].
struct FeatureBox {
    static constexpr double kHalfFov = 0.78539816339744831;          // pi/4
    static constexpr double kSpinMax = 0.998;
    static constexpr double kCosLo = 0.017452406437283512;          // cos 89 deg
    static constexpr double kCosHi = 0.99619469809174555;           // cos 5 deg
    static constexpr double kLnR0 = 2.9957322735539909;             // ln 20
    static constexpr double kLnRSpan = 1.6094379124341003;          // ln 5
};
__host__ __device__ inline void features(double alpha, double beta, double a, double cos_th, double r,
                                         float x[kIn]) {
    x[0] = static_cast<float>(alpha / FeatureBox::kHalfFov);
    x[1] = static_cast<float>(beta / FeatureBox::kHalfFov);
    x[2] = static_cast<float>(2.0 * a / FeatureBox::kSpinMax - 1.0);
    x[3] = static_cast<float>(2.0 * (cos_th - FeatureBox::kCosLo) / (FeatureBox::kCosHi - FeatureBox::kCosLo) - 1.0);
    x[4] = static_cast<float>(2.0 * (log(r) - FeatureBox::kLnR0) / FeatureBox::kLnRSpan - 1.0);
}

// Training target of one ray.  cls: 0 captured, 1 escaped, 2 disk, -1 = no
// target (the step limit was hit).  lnr/lng valid for cls 2; dir (sin, cos of
// theta_inf, sin, cos of phi_inf) valid when esc = 1 (the ray left through
// r_escape, which disk rays may also do).
struct Target {
    float cls, lnr, lng, dir[4], esc;
};

struct LossWeights {
    float w_r, w_g, w_dir, pad;
};

// Arguments of the MLP forward kernel.
struct MlpArgs {
    const __bf16* pack0;                // kPack0
    const __bf16* packH;                // 3 * kPackH
    const __bf16* pack4;                // kPack4
    const float* params;                // biases read from the FP32 master vector
    long long n;                        // rays in the launch
    // Training / plain forward: inputs X[n][5] (row order idx(i) below).
    const float* X;
    uint64_t perm_a, perm_c;            // batch row i reads X[(perm_a * (base + i) + perm_c) mod n_data]
    long long perm_base, n_data;        // n_data a power of two; perm_a == 0 reads X[i] directly
    float* Y;                           // [n][9] outputs
    float* Z;                           // training: [4][256][n] pre-activations (neuron-major)
    __bf16* HT;                         // training: [4][256][n] activations (neuron-major)
};

// Surrogate frame: the inference kernel computes every ray's inputs from the
// view, runs the network, and either shades the ray from the prediction or
// queues it for the FP64 integrator.
struct InferArgs {
    MlpArgs mlp;
    const TraceParams* P;               // the frame's trace parameters (device memory)
    double entropy_max;                 // nats; above it the ray falls back
    double lng_tol;                     // |ln g_net - ln g(r_net)| above it falls back (<= 0: off)
    // Critical-curve gate: sample i falls back when crit_flags[i] != 0
    // (written by launch_critical_flags; nullptr: no gate).
    const uint8_t* crit_flags;
    uint32_t judge_threshold;           // hash < threshold (of 2^32) selects a judged ray
    uint32_t frame_seed;
    long long ray_base;                 // first ray of the launch (rays ray_base .. ray_base + mlp.n)
    long long row_rays;                 // > 0: visit whole rows centre-out (rays per row); 0: in order
    int rows;                           // image rows (centre-out order)
    long long* fb_list;                 // fallback ray ids
    long long* judge_list;              // judged ray ids (shaded by the surrogate)
    unsigned long long* counts;         // [0] fallbacks, [1] judged
    long long judge_cap;
    unsigned long long* tile_counter;   // dynamic 256-ray tile claims (zeroed before the launch)
    // Inspection of the network pass itself: with x_in the inputs of ray i
    // are x_in[i][5] instead of the ones computed from its sample position;
    // with y_raw the pass writes the network's 9 outputs of ray i to
    // y_raw[i][9] and neither shades nor lists any ray.
    const float* x_in;
    float* y_raw;
    // Per-view inputs, formed once on the host with features(): the pixel
    // scale fov / (W * kHalfFov) that turns (px - W/2) into input 0 (and
    // (py - H/2) into input 1), and inputs 2..4 (spin, cos theta, ln r).
    double feat_scale;
    float feat_view[3];
};

hipError_t launch_pack(const float* params, __bf16* pack0, __bf16* packH, __bf16* pack4, __bf16* WT,
                       hipStream_t st);
hipError_t launch_mlp_forward(const MlpArgs& a, bool train, hipStream_t st);
hipError_t launch_mlp_infer(const InferArgs& a, int num_cus, hipStream_t st);
// Critical-curve flags of the frame's n samples (P: the frame's trace
// parameters in device memory): flags[i] = 1 when |critical score| of
// sample i's camera ray is below score_max.  r0, r1: the view's prograde
// and retrograde equatorial photon-orbit radii; sth, cth: sin and cos of
// the observer's theta.
hipError_t launch_critical_flags(const TraceParams* P, long long n, double score_max, double r0, double r1,
                                 double sth, double cth, uint8_t* flags, hipStream_t st);
hipError_t launch_loss_grad(const float* Y, const Target* T, const MlpArgs& rows, LossWeights w, float* dY,
                            double* loss_acc, hipStream_t st);
hipError_t launch_make_dz(const float* dHT, const float* dY, const float* params, const float* ZT, long long B,
                          __bf16* dZbm, __bf16* dZnm, float* dbp, hipStream_t st);
hipError_t launch_out_layer_grads(const float* dY, const __bf16* H3T, long long B, float* grad, hipStream_t st);
hipError_t launch_gemm_nt(int M, int N, int K, const __bf16* A, long long lda, const __bf16* Bm, long long ldb,
                          float* C, long long ldc, int splits, hipStream_t st);
hipError_t launch_reduce_splits(const float* part, int splits, long long count, float* out, hipStream_t st);
hipError_t launch_layer0_grads(const float* dH0T, const float* Z0T, const MlpArgs& rows, float* part,
                               hipStream_t st);
hipError_t launch_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t st);
hipError_t launch_targets(const GRRTRayRecord* rec, const ViewParams* views, int n_views, long long n_rays,
                          float* X, Target* T, hipStream_t st);
hipError_t launch_judge(const long long* judge_list, const unsigned long long* counts, long long judge_cap,
                        const double* rgb, const uint8_t* cls, long long n_rays, float* rel_err,
                        unsigned long long* agree, hipStream_t st);
hipError_t launch_reset_counts(unsigned long long* c, int n, hipStream_t st);

}  // namespace grrt_amd::sg
