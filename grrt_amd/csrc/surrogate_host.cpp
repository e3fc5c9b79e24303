// Host side of the learned surrogate: weights and Adam state, training-set
// generation with the FP64 integrator, the training loop, and the surrogate
// frame (network pass, FP64 fallback and judge).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "api_internal.hpp"
#include "surrogate.hpp"

namespace grrt_amd::sg {

struct SurrogateState {
    DevBuf<float> w, m, v, grad;
    DevBuf<__bf16> pack0, packH, pack4, WT;
    long long adam_t = 0;
    // Training and validation rows.
    DevBuf<float> X, Xv;
    DevBuf<Target> T, Tv;
    long long n_data = 0, n_val = 0;
    // Work buffers of one batch.
    long long B = 0;
    DevBuf<float> Y, ZT, dY, dHT, part, dbp;
    DevBuf<__bf16> HT, dZbm, dZnm;
    DevBuf<double> loss_acc;
    // Surrogate frame.
    DevBuf<long long> fb_list, judge_list;
    DevBuf<unsigned long long> counts;
    DevBuf<float> judge_err;
    DevBuf<uint8_t> crit_flags;
    uint32_t frame_no = 0;
    unsigned long long last_fb = 0, last_judged = 0;   // list lengths of the last surrogate frame
    // Critical-curve gate of the surrogate frame: rays with |critical_score|
    // below this fall back (0: off).  0.03 flags 0.07% of the Kerr a=0.998
    // 1024^2 frame's rays, whose integrations average 2.8x the frame's mean.
    double crit_max = 0.03;

    ~SurrogateState() {
        for (auto* b : {&w, &m, &v, &grad, &X, &Xv, &Y, &ZT, &dY, &dHT, &part, &dbp, &judge_err}) b->release();
        for (auto* b : {&pack0, &packH, &pack4, &WT, &HT, &dZbm, &dZnm}) b->release();
        T.release();
        Tv.release();
        loss_acc.release();
        fb_list.release();
        judge_list.release();
        counts.release();
        crit_flags.release();
    }
};

void destroy_state(SurrogateState* s) { delete s; }

}  // namespace grrt_amd::sg

namespace {

using namespace grrt_amd;
using namespace grrt_amd::sg;
using grrt_api::check;
using grrt_api::HipFailure;

SurrogateState& state(GRRTContext* ctx) {
    if (!ctx->sg) throw HipFailure("surrogate not initialised (grrt_ext_surrogate_init)");
    return *ctx->sg;
}

MlpArgs net_args(SurrogateState& S) {
    MlpArgs A{};
    A.pack0 = S.pack0.ptr;
    A.packH = S.packH.ptr;
    A.pack4 = S.pack4.ptr;
    A.params = S.w.ptr;
    return A;
}

void repack(GRRTContext* ctx, SurrogateState& S) {
    check(launch_pack(S.w.ptr, S.pack0.ptr, S.packH.ptr, S.pack4.ptr, S.WT.ptr, ctx->stream), "pack weights");
}

// Work buffers for batches of B rows.
void ensure_batch(SurrogateState& S, long long B) {
    if (B <= 0 || B % 256) throw HipFailure("surrogate batch must be a positive multiple of 256 rays");
    if (S.B >= B) return;
    const size_t nb = static_cast<size_t>(B);
    const int splits = (B % 512 == 0) ? static_cast<int>(B / 512) : 1;
    check(S.Y.ensure(nb * kOut), "hipMalloc(Y)");
    check(S.dY.ensure(nb * kOut), "hipMalloc(dY)");
    check(S.ZT.ensure(4 * kHid * nb), "hipMalloc(Z)");
    check(S.HT.ensure(4 * kHid * nb), "hipMalloc(H)");
    check(S.dHT.ensure(kHid * nb), "hipMalloc(dH)");
    check(S.dZbm.ensure(kHid * nb), "hipMalloc(dZ)");
    check(S.dZnm.ensure(kHid * nb), "hipMalloc(dZ)");
    check(S.dbp.ensure((nb / 64) * kHid), "hipMalloc(bias partials)");
    check(S.part.ensure(std::max<size_t>(static_cast<size_t>(splits) * kHid * kHid, (nb / 64) * (oB0 + kHid))),
          "hipMalloc(partials)");
    check(S.loss_acc.ensure(8), "hipMalloc(loss)");
    S.B = B;
}

// One forward + backward over B rows (rows: X/T with the permutation);
// leaves the gradient in S.grad and adds the loss statistics to S.loss_acc.
void forward_backward(GRRTContext* ctx, SurrogateState& S, MlpArgs rows, const Target* T, LossWeights lw) {
    const long long B = rows.n;
    const hipStream_t st = ctx->stream;
    MlpArgs A = net_args(S);
    A.n = B;
    A.X = rows.X;
    A.perm_a = rows.perm_a;
    A.perm_c = rows.perm_c;
    A.perm_base = rows.perm_base;
    A.n_data = rows.n_data;
    A.Y = S.Y.ptr;
    A.Z = S.ZT.ptr;
    A.HT = S.HT.ptr;
    check(launch_mlp_forward(A, true, st), "forward");
    check(launch_loss_grad(S.Y.ptr, T, A, lw, S.dY.ptr, S.loss_acc.ptr, st), "loss");
    const long long LB = static_cast<long long>(kHid) * B;      // one layer's activations
    check(launch_out_layer_grads(S.dY.ptr, S.HT.ptr + 3 * LB, B, S.grad.ptr, st), "output-layer gradients");
    check(launch_make_dz(nullptr, S.dY.ptr, S.w.ptr, S.ZT.ptr + 3 * LB, B, S.dZbm.ptr, S.dZnm.ptr, S.dbp.ptr, st),
          "dZ3");
    check(launch_reduce_splits(S.dbp.ptr, static_cast<int>(B / 64), kHid, S.grad.ptr + oB3, st), "db3");
    const int splits = (B % 512 == 0) ? static_cast<int>(B / 512) : 1;
    for (int l = 3; l >= 1; --l) {
        // dW_l[n][k] = sum_b dZ_l[b][n] H_{l-1}[b][k]
        check(launch_gemm_nt(kHid, kHid, static_cast<int>(B), S.dZnm.ptr, B, S.HT.ptr + (l - 1) * LB, B, S.part.ptr,
                             kHid, splits, st),
              "dW gemm");
        check(launch_reduce_splits(S.part.ptr, splits, static_cast<long long>(kHid) * kHid,
                                   S.grad.ptr + weight_offset(l), st),
              "dW reduce");
        // dH_{l-1}^T[k][b] = sum_n W_l[n][k] dZ_l[b][n]
        check(launch_gemm_nt(kHid, static_cast<int>(B), kHid, S.WT.ptr + static_cast<long long>(l - 1) * kHid * kHid,
                             kHid, S.dZbm.ptr, kHid, S.dHT.ptr, B, 1, st),
              "dH gemm");
        if (l - 1 >= 1) {
            check(launch_make_dz(S.dHT.ptr, nullptr, S.w.ptr, S.ZT.ptr + (l - 1) * LB, B, S.dZbm.ptr, S.dZnm.ptr,
                                 S.dbp.ptr, st),
                  "dZ");
            check(launch_reduce_splits(S.dbp.ptr, static_cast<int>(B / 64), kHid, S.grad.ptr + bias_offset(l - 1), st),
                  "db");
        } else {
            check(launch_layer0_grads(S.dHT.ptr, S.ZT.ptr, A, S.part.ptr, st), "layer-0 gradients");
            check(launch_reduce_splits(S.part.ptr, static_cast<int>(B / 64), oB0 + kHid - oW0, S.grad.ptr + oW0, st),
                  "layer-0 reduce");
        }
    }
}

// Means of the loss statistics accumulated in acc (see loss_grad_kernel).
void fill_stats(const double* acc, LossWeights lw, GRRTTrainStats* out) {
    const double n = std::max(acc[7], 1.0);
    out->rays = static_cast<int64_t>(acc[7]);
    out->ce = acc[0] / n;
    out->accuracy = acc[1] / n;
    out->rmse_lnr = std::sqrt(acc[2] / std::max(acc[4], 1.0));
    out->rmse_lng = std::sqrt(acc[3] / std::max(acc[4], 1.0));
    out->rmse_dir = std::sqrt(acc[5] / std::max(4.0 * acc[6], 1.0));
    out->loss = (acc[0] + lw.w_r * acc[2] + lw.w_g * acc[3] + lw.w_dir * acc[5]) / n;
}

// The frame's trace parameters (one view: the context's camera, disk and
// sampling) for the network pass and the fallback trace.
TraceParams frame_params(GRRTContext* ctx, long long n_rays) {
    const Resolved& r = ctx->res;
    TraceParams P{};
    grrt_api::fill_view(ctx, P.view, r.a, r.r_obs, r.th_obs, r.ph_obs, r.fov, r.W, r.H, ctx->cam, ctx->flux.ptr,
                        ctx->flux_max, r.r_in, 0, 0, r.W, r.H, 0);
    P.r_escape = 1000.0;
    P.tol = r.tol;
    P.max_steps = r.max_steps;
    P.n_rays = n_rays;
    P.view.n_rays = n_rays;
    P.spec_lut = ctx->spec.ptr;
    P.stars = r.stars;
    P.star_start = ctx->star_start.ptr;
    P.star_tpb = ctx->star_tpb.ptr;
    grrt_api::set_background(ctx, P);
    P.sample_rgb = ctx->sample_rgb.ptr;
    P.sample_cls = ctx->sample_cls.ptr;
    P.records = nullptr;
    P.cancel = nullptr;
    P.refill_min = std::clamp(ctx->tune_refill, 1, 64);
    P.integrator = ctx->integrator;
    return P;
}

// The pass's per-view inputs (features() of the view's spin, inclination
// and distance) and the pixel scale of inputs 0 and 1, formed once here
// instead of per ray.
void set_view_features(InferArgs& IA, const ViewParams& V) {
    float x[kIn];
    features(0.0, 0.0, V.a / V.M, std::cos(V.th_obs), V.r_obs / V.M, x);
    IA.feat_view[0] = x[2];
    IA.feat_view[1] = x[3];
    IA.feat_view[2] = x[4];
    IA.feat_scale = V.fov / (static_cast<double>(V.W) * FeatureBox::kHalfFov);
}

}  // namespace

extern "C" {

int grrt_ext_surrogate_infer_raw(GRRTContext* ctx, const float* x, int64_t n, float* y) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        if (!y || n <= 0) throw HipFailure("surrogate raw pass: no rays");
        const Resolved& r = ctx->res;
        int grid = 1;
        const int nS = grrt_api::samples_per_pixel(ctx->sampling_mode, ctx->sampling_spp, grid);
        const long long frame_rays = static_cast<long long>(r.W) * r.H * nS;
        if (!x && n != frame_rays) throw HipFailure("surrogate raw pass: without inputs n must be the frame's ray count");
        check(ctx->params_dev.ensure(1), "hipMalloc(params)");
        check(S.counts.ensure(4), "hipMalloc(counts)");
        DevBuf<float> dx, dy;
        if (x) {
            check(dx.ensure(static_cast<size_t>(n) * kIn), "hipMalloc(x)");
            check(hipMemcpy(dx.ptr, x, sizeof(float) * kIn * n, hipMemcpyHostToDevice), "copy x");
        }
        check(dy.ensure(static_cast<size_t>(n) * kOut), "hipMalloc(y)");
        const TraceParams P = frame_params(ctx, x ? n : frame_rays);
        check(hipMemcpy(ctx->params_dev.ptr, &P, sizeof(P), hipMemcpyHostToDevice), "upload params");
        check(launch_reset_counts(S.counts.ptr, 4, ctx->stream), "reset counts");
        InferArgs IA{};
        IA.mlp = net_args(S);
        IA.mlp.n = n;
        IA.P = ctx->params_dev.ptr;
        IA.entropy_max = 1e30;
        IA.ray_base = 0;
        IA.row_rays = 0;
        IA.rows = r.H;
        IA.counts = S.counts.ptr;
        IA.tile_counter = S.counts.ptr + 3;
        IA.x_in = x ? dx.ptr : nullptr;
        IA.y_raw = dy.ptr;
        set_view_features(IA, P.view);
        check(launch_mlp_infer(IA, ctx->num_cus, ctx->stream), "surrogate pass");
        check(hipStreamSynchronize(ctx->stream), "surrogate pass");
        check(hipMemcpy(y, dy.ptr, sizeof(float) * kOut * n, hipMemcpyDeviceToHost), "copy y");
        dx.release();
        dy.release();
        return 0;
    });
}

int grrt_ext_surrogate_adam_step(GRRTContext* ctx, float* w, const float* g, float* m, float* v, int64_t n,
                                 double lr, double b1, double b2, double eps, int64_t t) {
    return grrt_api::guarded(ctx, [&] {
        if (!w || !g || !m || !v || n <= 0 || t <= 0) throw HipFailure("adam step: bad arguments");
        DevBuf<float> dw, dg, dm, dv;
        const size_t bytes = sizeof(float) * static_cast<size_t>(n);
        for (auto* b : {&dw, &dg, &dm, &dv}) check(b->ensure(static_cast<size_t>(n)), "hipMalloc(adam)");
        check(hipMemcpy(dw.ptr, w, bytes, hipMemcpyHostToDevice), "copy w");
        check(hipMemcpy(dg.ptr, g, bytes, hipMemcpyHostToDevice), "copy g");
        check(hipMemcpy(dm.ptr, m, bytes, hipMemcpyHostToDevice), "copy m");
        check(hipMemcpy(dv.ptr, v, bytes, hipMemcpyHostToDevice), "copy v");
        // The bias corrections exactly as the training loop forms them.
        check(launch_adam(dw.ptr, dg.ptr, dm.ptr, dv.ptr, n, static_cast<float>(lr), static_cast<float>(b1),
                          static_cast<float>(b2), static_cast<float>(eps),
                          static_cast<float>(1.0 - std::pow(b1, static_cast<double>(t))),
                          static_cast<float>(1.0 - std::pow(b2, static_cast<double>(t))), ctx->stream),
              "adam");
        check(hipStreamSynchronize(ctx->stream), "adam");
        check(hipMemcpy(w, dw.ptr, bytes, hipMemcpyDeviceToHost), "copy w");
        check(hipMemcpy(m, dm.ptr, bytes, hipMemcpyDeviceToHost), "copy m");
        check(hipMemcpy(v, dv.ptr, bytes, hipMemcpyDeviceToHost), "copy v");
        for (auto* b : {&dw, &dg, &dm, &dv}) b->release();
        return 0;
    });
}

int grrt_ext_surrogate_lists(GRRTContext* ctx, int64_t* fallback, int64_t n_fallback, int64_t* judged,
                             int64_t n_judged) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        if (n_fallback < 0 || n_judged < 0 || n_fallback > static_cast<int64_t>(S.last_fb) ||
            n_judged > static_cast<int64_t>(S.last_judged))
            throw HipFailure("surrogate lists: count out of range");
        static_assert(sizeof(long long) == sizeof(int64_t), "ray ids are 64-bit");
        if (fallback && n_fallback)
            check(hipMemcpy(fallback, S.fb_list.ptr, sizeof(int64_t) * n_fallback, hipMemcpyDeviceToHost), "fallback ids");
        if (judged && n_judged)
            check(hipMemcpy(judged, S.judge_list.ptr, sizeof(int64_t) * n_judged, hipMemcpyDeviceToHost), "judged ids");
        return 0;
    });
}

int grrt_ext_surrogate_init(GRRTContext* ctx, uint64_t seed) {
    return grrt_api::guarded(ctx, [&] {
        if (!ctx->sg) ctx->sg = new SurrogateState();
        SurrogateState& S = *ctx->sg;
        std::vector<float> w(kNumParams, 0.0f);
        std::mt19937_64 rng(seed);
        // He-uniform hidden layers, a smaller output layer, zero biases.
        auto fill = [&](int off, int rows, int cols, double scale) {
            std::uniform_real_distribution<double> U(-scale, scale);
            for (int i = 0; i < rows * cols; ++i) w[off + i] = static_cast<float>(U(rng));
        };
        fill(oW0, kHid, kIn, std::sqrt(6.0 / kIn));
        fill(oW1, kHid, kHid, std::sqrt(6.0 / kHid));
        fill(oW2, kHid, kHid, std::sqrt(6.0 / kHid));
        fill(oW3, kHid, kHid, std::sqrt(6.0 / kHid));
        fill(oW4, kOut, kHid, std::sqrt(1.0 / kHid));
        grrt_api::upload(S.w, w, ctx->stream);
        std::vector<float> zero(kNumParams, 0.0f);
        grrt_api::upload(S.m, zero, ctx->stream);
        grrt_api::upload(S.v, zero, ctx->stream);
        check(S.grad.ensure(kNumParams), "hipMalloc(grad)");
        check(S.pack0.ensure(kPack0), "hipMalloc(pack)");
        check(S.packH.ensure(3 * kPackH), "hipMalloc(pack)");
        check(S.pack4.ensure(kPack4), "hipMalloc(pack)");
        check(S.WT.ensure(3 * kHid * kHid), "hipMalloc(WT)");
        S.adam_t = 0;
        repack(ctx, S);
        check(hipStreamSynchronize(ctx->stream), "surrogate init");
        return 0;
    });
}

int grrt_ext_surrogate_params(GRRTContext* ctx, float* params, int64_t n, int set) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        if (!params || n != kNumParams) throw HipFailure("surrogate parameter vector must hold 201225 floats");
        if (set) {
            check(hipMemcpy(S.w.ptr, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice), "set params");
            repack(ctx, S);
            check(hipStreamSynchronize(ctx->stream), "pack");
        } else {
            check(hipMemcpy(params, S.w.ptr, sizeof(float) * kNumParams, hipMemcpyDeviceToHost), "get params");
        }
        return 0;
    });
}

int grrt_ext_surrogate_forward(GRRTContext* ctx, const float* x, int64_t n, float* y) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        if (!x || !y || n <= 0) throw HipFailure("surrogate forward: empty input");
        DevBuf<float> dx, dy;
        check(dx.ensure(static_cast<size_t>(n) * kIn), "hipMalloc(x)");
        check(dy.ensure(static_cast<size_t>(n) * kOut), "hipMalloc(y)");
        check(hipMemcpy(dx.ptr, x, sizeof(float) * kIn * n, hipMemcpyHostToDevice), "copy x");
        MlpArgs A = net_args(S);
        A.n = n;
        A.X = dx.ptr;
        A.Y = dy.ptr;
        check(launch_mlp_forward(A, false, ctx->stream), "forward");
        check(hipStreamSynchronize(ctx->stream), "forward");
        check(hipMemcpy(y, dy.ptr, sizeof(float) * kOut * n, hipMemcpyDeviceToHost), "copy y");
        dx.release();
        dy.release();
        return 0;
    });
}

int grrt_ext_surrogate_gradients(GRRTContext* ctx, const float* x, const float* t, int64_t n, const float w[3],
                                 float* grad, double* loss) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        if (!x || !t || !w || !grad) throw HipFailure("surrogate gradients: null argument");
        ensure_batch(S, n);
        DevBuf<float> dx;
        DevBuf<Target> dt;
        check(dx.ensure(static_cast<size_t>(n) * kIn), "hipMalloc(x)");
        check(dt.ensure(static_cast<size_t>(n)), "hipMalloc(t)");
        static_assert(sizeof(Target) == 8 * sizeof(float), "Target is 8 floats");
        check(hipMemcpy(dx.ptr, x, sizeof(float) * kIn * n, hipMemcpyHostToDevice), "copy x");
        check(hipMemcpy(dt.ptr, t, sizeof(Target) * n, hipMemcpyHostToDevice), "copy t");
        check(hipMemsetAsync(S.loss_acc.ptr, 0, 8 * sizeof(double), ctx->stream), "clear loss");
        MlpArgs rows{};
        rows.n = n;
        rows.X = dx.ptr;
        const LossWeights lw{w[0], w[1], w[2], 0.0f};
        forward_backward(ctx, S, rows, dt.ptr, lw);
        check(hipStreamSynchronize(ctx->stream), "gradients");
        check(hipMemcpy(grad, S.grad.ptr, sizeof(float) * kNumParams, hipMemcpyDeviceToHost), "copy grad");
        double acc[8];
        check(hipMemcpy(acc, S.loss_acc.ptr, sizeof(acc), hipMemcpyDeviceToHost), "read loss");
        if (loss) *loss = (acc[0] + lw.w_r * acc[2] + lw.w_g * acc[3] + lw.w_dir * acc[5]) / static_cast<double>(n);
        dx.release();
        dt.release();
        return 0;
    });
}

int grrt_ext_surrogate_train(GRRTContext* ctx, const GRRTTrainConfig* cfg, GRRTTrainStats* out) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        if (!cfg || cfg->steps <= 0) throw HipFailure("surrogate train: no steps");
        if (S.n_data <= 0) throw HipFailure("surrogate train: no training set (grrt_ext_surrogate_generate)");
        const long long B = cfg->batch;
        ensure_batch(S, B);
        const LossWeights lw{static_cast<float>(cfg->w_r), static_cast<float>(cfg->w_g),
                             static_cast<float>(cfg->w_dir), 0.0f};
        const int tail = std::max(1, cfg->steps / 10);
        check(hipEventRecord(ctx->ev[0], ctx->stream), "event");
        for (int k = 0; k < cfg->steps; ++k) {
            if (k == cfg->steps - tail)
                check(hipMemsetAsync(S.loss_acc.ptr, 0, 8 * sizeof(double), ctx->stream), "clear loss");
            const long long start = (S.adam_t) * B;
            const long long epoch = start / S.n_data;
            std::mt19937_64 er(cfg->seed * 0x9E3779B97F4A7C15ull + static_cast<uint64_t>(epoch));
            MlpArgs rows{};
            rows.n = B;
            rows.X = S.X.ptr;
            rows.perm_a = er() | 1ull;
            rows.perm_c = er();
            rows.perm_base = start % S.n_data;
            rows.n_data = S.n_data;
            forward_backward(ctx, S, rows, S.T.ptr, lw);
            // Learning rate: linear warm-up, then cosine decay over the call.
            double lr;
            if (k < cfg->warmup) {
                lr = cfg->lr_max * (k + 1) / cfg->warmup;
            } else {
                const double u = static_cast<double>(k - cfg->warmup) / std::max(1, cfg->steps - cfg->warmup);
                lr = cfg->lr_min + 0.5 * (cfg->lr_max - cfg->lr_min) * (1.0 + std::cos(3.141592653589793 * u));
            }
            ++S.adam_t;
            const double b1 = 0.9, b2 = 0.999;
            check(launch_adam(S.w.ptr, S.grad.ptr, S.m.ptr, S.v.ptr, kNumParams, static_cast<float>(lr),
                              static_cast<float>(b1), static_cast<float>(b2), 1e-8f,
                              static_cast<float>(1.0 - std::pow(b1, static_cast<double>(S.adam_t))),
                              static_cast<float>(1.0 - std::pow(b2, static_cast<double>(S.adam_t))), ctx->stream),
                  "adam");
            repack(ctx, S);
        }
        check(hipEventRecord(ctx->ev[1], ctx->stream), "event");
        check(hipEventSynchronize(ctx->ev[1]), "train");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "event time");
        double acc[8];
        check(hipMemcpy(acc, S.loss_acc.ptr, sizeof(acc), hipMemcpyDeviceToHost), "read loss");
        if (out) {
            *out = GRRTTrainStats{};
            fill_stats(acc, lw, out);
            out->steps = S.adam_t;
            out->ms = ms;
        }
        return 0;
    });
}

int grrt_ext_surrogate_evaluate(GRRTContext* ctx, GRRTTrainStats* out) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        if (S.n_val <= 0) throw HipFailure("surrogate evaluate: no validation set");
        const long long chunk = 65536;
        check(S.Y.ensure(static_cast<size_t>(chunk) * kOut), "hipMalloc(Y)");
        check(S.dY.ensure(static_cast<size_t>(chunk) * kOut), "hipMalloc(dY)");
        check(S.loss_acc.ensure(8), "hipMalloc(loss)");
        check(hipMemsetAsync(S.loss_acc.ptr, 0, 8 * sizeof(double), ctx->stream), "clear loss");
        check(hipEventRecord(ctx->ev[0], ctx->stream), "event");
        for (long long o = 0; o < S.n_val; o += chunk) {
            const long long n = std::min(chunk, S.n_val - o);
            MlpArgs A = net_args(S);
            A.n = n;
            A.X = S.Xv.ptr + o * kIn;
            A.Y = S.Y.ptr;
            check(launch_mlp_forward(A, false, ctx->stream), "forward");
            check(launch_loss_grad(S.Y.ptr, S.Tv.ptr + o, A, LossWeights{1, 1, 1, 0}, S.dY.ptr, S.loss_acc.ptr,
                                   ctx->stream),
                  "metrics");
        }
        check(hipEventRecord(ctx->ev[1], ctx->stream), "event");
        check(hipEventSynchronize(ctx->ev[1]), "evaluate");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "event time");
        double acc[8];
        check(hipMemcpy(acc, S.loss_acc.ptr, sizeof(acc), hipMemcpyDeviceToHost), "read loss");
        if (out) {
            *out = GRRTTrainStats{};
            fill_stats(acc, LossWeights{1, 1, 1, 0}, out);
            out->steps = S.adam_t;
            out->ms = ms;
        }
        return 0;
    });
}

int grrt_ext_surrogate_dataset(GRRTContext* ctx, int validation, float* x, float* t, int64_t n) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        const long long have = validation ? S.n_val : S.n_data;
        if (n < 0 || n > have) throw HipFailure("surrogate dataset: row count out of range");
        if (x) check(hipMemcpy(x, validation ? S.Xv.ptr : S.X.ptr, sizeof(float) * kIn * n, hipMemcpyDeviceToHost), "x");
        if (t) check(hipMemcpy(t, validation ? S.Tv.ptr : S.T.ptr, sizeof(Target) * n, hipMemcpyDeviceToHost), "t");
        return 0;
    });
}

int grrt_ext_surrogate_generate(GRRTContext* ctx, int64_t n_rays, uint64_t seed, int validation) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        if (n_rays <= 0) throw HipFailure("surrogate generate: no rays");
        if (!validation && (n_rays & (n_rays - 1))) throw HipFailure("training set size must be a power of two");
        if (ctx->res.volumetric) throw HipFailure("the surrogate covers the thin disk");
        constexpr int kSide = 32;                   // 32x32 views, 1 sample per pixel
        constexpr long long kPerView = kSide * kSide;
        constexpr int kChunkViews = 4096;
        const long long n_views = (n_rays + kPerView - 1) / kPerView;
        DevBuf<float>& X = validation ? S.Xv : S.X;
        DevBuf<Target>& T = validation ? S.Tv : S.T;
        check(X.ensure(static_cast<size_t>(n_views * kPerView) * kIn), "hipMalloc(X)");
        check(T.ensure(static_cast<size_t>(n_views * kPerView)), "hipMalloc(T)");
        // Keep the caller's sampling and records settings.
        const int mode = ctx->sampling_mode, spp = ctx->sampling_spp;
        const uint64_t sseed = ctx->sampling_seed;
        const bool keep = ctx->keep_records;
        ctx->sampling_mode = GRRT_SAMPLING_SEEDED;
        ctx->sampling_spp = 1;
        ctx->sampling_seed = seed;
        ctx->keep_records = true;
        std::mt19937_64 rng(seed ^ (validation ? 0x5A5A5A5A12345678ull : 0ull));
        std::uniform_real_distribution<double> U(0.0, 1.0);
        const double deg = 3.141592653589793 / 180.0;
        try {
            for (long long v0 = 0; v0 < n_views; v0 += kChunkViews) {
                const int nv = static_cast<int>(std::min<long long>(kChunkViews, n_views - v0));
                std::vector<GRRTView> views(nv);
                for (auto& w : views) {
                    w.spin = 0.998 * U(rng);
                    w.observer_theta = (5.0 + 84.0 * U(rng)) * deg;
                    w.observer_r = 20.0 * std::pow(5.0, U(rng));
                    w.observer_phi = 0.0;
                    w.fov = 90.0 * deg;
                    w.width = w.height = kSide;
                }
                // A distinct jitter key per view (and per chunk) so the rays
                // cover the field of view instead of repeating one pattern.
                ctx->sampling_seed = seed * 0x9E3779B97F4A7C15ull + static_cast<uint64_t>(v0) * 0xD1B54A32D192ED03ull +
                                     (validation ? 0x632BE59BD9B4E019ull : 0ull);
                grrt_api::prepare_batch(ctx, views.data(), nv, 0xBF58476D1CE4E5B9ull);
                grrt_api::run_batch(ctx);
                const long long rays = static_cast<long long>(nv) * kPerView;
                check(launch_targets(ctx->batch_records.ptr, ctx->batch_params.ptr, nv, rays,
                                     X.ptr + v0 * kPerView * kIn, T.ptr + v0 * kPerView, ctx->stream),
                      "training targets");
                check(hipStreamSynchronize(ctx->stream), "training targets");
            }
        } catch (...) {
            ctx->sampling_mode = mode;
            ctx->sampling_spp = spp;
            ctx->sampling_seed = sseed;
            ctx->keep_records = keep;
            ctx->batch_views.clear();
            throw;
        }
        ctx->sampling_mode = mode;
        ctx->sampling_spp = spp;
        ctx->sampling_seed = sseed;
        ctx->keep_records = keep;
        ctx->batch_views.clear();
        ctx->batch_records.release();
        if (validation) S.n_val = n_rays;
        else S.n_data = n_rays;
        return 0;
    });
}

int grrt_ext_surrogate_set_critical_gate(GRRTContext* ctx, double score_max) {
    return grrt_api::guarded(ctx, [&] {
        if (!(score_max >= 0.0)) throw HipFailure("critical gate: score_max must be >= 0");
        state(ctx).crit_max = score_max;
        return 0;
    });
}

int grrt_ext_render_surrogate(GRRTContext* ctx, float* out, int device_out, double entropy_max, double lng_tol,
                              double judge_fraction, GRRTSurrogateStats* stats) {
    return grrt_api::guarded(ctx, [&] {
        SurrogateState& S = state(ctx);
        const Resolved& r = ctx->res;
        if (r.volumetric) throw HipFailure("the surrogate covers the thin disk");
        if (!out && !device_out) throw HipFailure("null output");
        const auto t0 = std::chrono::steady_clock::now();
        const hipStream_t st = ctx->stream;
        int grid = 1;
        const int nS = grrt_api::samples_per_pixel(ctx->sampling_mode, ctx->sampling_spp, grid);
        const long long npix = static_cast<long long>(r.W) * r.H;
        const long long n_rays = npix * nS;
        judge_fraction = std::clamp(judge_fraction, 0.0, 1.0);
        // One band of all rays, visited row by row from the image centre
        // outwards: the rays around the shadow and the inner disk, whose
        // fallback traces are the longest, lead the fallback list and start
        // first in the integrator launch that follows the network pass.
        struct Band {
            long long r0, r1, cap, joff;
        };
        std::vector<Band> bands;
        const long long row_rays = static_cast<long long>(r.W) * nS;
        bands.push_back(Band{0, n_rays, static_cast<long long>(judge_fraction * 1.25 * static_cast<double>(n_rays)) + 4096, 0});
        const int nb = static_cast<int>(bands.size());
        const long long cap = bands.back().joff + bands.back().cap;
        check(ctx->sample_rgb.ensure(static_cast<size_t>(3 * (n_rays + cap))), "hipMalloc(samples)");
        check(ctx->sample_cls.ensure(static_cast<size_t>(n_rays + cap)), "hipMalloc(classes)");
        check(ctx->frame.ensure(static_cast<size_t>(4 * npix)), "hipMalloc(frame)");
        check(ctx->counters.ensure(static_cast<size_t>(nb)), "hipMalloc(counters)");
        check(ctx->params_dev.ensure(static_cast<size_t>(nb)), "hipMalloc(params)");
        check(S.fb_list.ensure(static_cast<size_t>(n_rays)), "hipMalloc(fallback list)");
        check(S.judge_list.ensure(static_cast<size_t>(cap)), "hipMalloc(judge list)");
        check(S.judge_err.ensure(static_cast<size_t>(cap)), "hipMalloc(judge errors)");
        check(S.counts.ensure(static_cast<size_t>(4 * nb)), "hipMalloc(counts)");

        TraceParams P = frame_params(ctx, n_rays);
        // One wavefront's worth of listed rays per queue claim: every lane of
        // the launch starts on a ray at once, so the longest integrations at
        // the head of the list are not left waiting for a refill (3.2 ms
        // after the network pass against 3.6 ms with the plain frame's
        // 128-ray claims).
        P.chunk = 64;
        P.integrator = ctx->integrator;
        std::vector<TraceParams> bp(nb, P);
        for (int k = 0; k < nb; ++k) {
            bp[k].counters = ctx->counters.ptr + k;
            bp[k].fb_list = S.fb_list.ptr + bands[k].r0;
            bp[k].judge_list = S.judge_list.ptr + bands[k].joff;
            bp[k].list_counts = S.counts.ptr + 4 * k;
            bp[k].judge_cap = bands[k].cap;
            bp[k].judge_out = n_rays + bands[k].joff;
        }
        check(hipMemcpy(ctx->params_dev.ptr, bp.data(), sizeof(TraceParams) * nb, hipMemcpyHostToDevice),
              "upload params");

        const uint32_t thr = static_cast<uint32_t>(std::min(judge_fraction * 4294967296.0, 4294967295.0));
        const uint32_t frame = ++S.frame_no;
        const hipStream_t s2 = ctx->stream2;
        check(launch_reset(ctx->counters.ptr, ctx->sample_cls.ptr, n_rays + cap, st), "reset");
        for (int k = 1; k < nb; ++k) check(launch_reset(ctx->counters.ptr + k, nullptr, 0, st), "reset");
        check(launch_reset_counts(S.counts.ptr, 4 * nb, st), "reset counts");
        check(hipEventRecord(ctx->ev[0], st), "event");
        if (S.crit_max > 0.0) {
            // prograde and retrograde equatorial photon orbits (Bardeen, Press & Teukolsky 1972)
            const double M = P.view.M, x = std::clamp(P.view.a / M, -1.0, 1.0);
            const double rp = 2.0 * M * (1.0 + std::cos(2.0 / 3.0 * std::acos(-x)));
            const double rm = 2.0 * M * (1.0 + std::cos(2.0 / 3.0 * std::acos(x)));
            check(S.crit_flags.ensure(static_cast<size_t>(n_rays)), "hipMalloc(critical flags)");
            check(launch_critical_flags(ctx->params_dev.ptr, n_rays, S.crit_max, std::min(rp, rm), std::max(rp, rm),
                                        std::sin(P.view.th_obs), std::cos(P.view.th_obs), S.crit_flags.ptr, st),
                  "critical flags");
        }
        // The fallback launch takes the plain frame's launch shape (three
        // blocks per CU by default).  On the 2048^2 x 4 frame (~0.77 M listed
        // rays) that traces them in 5.7 ms against 6.5 ms with one block per
        // CU; on the 1024^2 x 4 frame (~0.19 M) the two are within run-to-run
        // spread (3.2-3.5 ms).
        const int blocks = std::max(1, ctx->num_cus * ctx->tune_blocks_per_cu);
        for (int k = 0; k < nb; ++k) {
            InferArgs IA{};
            IA.mlp = net_args(S);
            IA.mlp.n = bands[k].r1 - bands[k].r0;
            IA.P = ctx->params_dev.ptr + k;
            IA.entropy_max = entropy_max;
            IA.lng_tol = lng_tol;
            IA.crit_flags = S.crit_max > 0.0 ? S.crit_flags.ptr : nullptr;
            IA.judge_threshold = thr;
            IA.frame_seed = frame;
            IA.ray_base = bands[k].r0;
            IA.row_rays = row_rays;
            IA.rows = r.H;
            IA.fb_list = S.fb_list.ptr + bands[k].r0;
            IA.judge_list = S.judge_list.ptr + bands[k].joff;
            IA.counts = S.counts.ptr + 4 * k;
            IA.judge_cap = bands[k].cap;
            IA.tile_counter = S.counts.ptr + 4 * k + 3;
            set_view_features(IA, P.view);
            check(launch_mlp_infer(IA, ctx->num_cus, st), "surrogate pass");
            // The end of the pass is recorded before any work goes to the
            // second stream: where the two streams share one hardware queue,
            // an event recorded later would land after that work.
            if (k == nb - 1) check(hipEventRecord(ctx->ev[1], st), "event");
            check(hipEventRecord(ctx->band_ev[k], st), "event");
            check(hipStreamWaitEvent(s2, ctx->band_ev[k], 0), "stream wait");
            check(launch_trace_thin(ctx->params_dev.ptr + k, 2, blocks, ctx->tune_block, s2, ctx->integrator),
                  "fallback trace");
            check(launch_judge(S.judge_list.ptr + bands[k].joff, S.counts.ptr + 4 * k, bands[k].cap,
                               ctx->sample_rgb.ptr, ctx->sample_cls.ptr, n_rays + bands[k].joff,
                               S.judge_err.ptr + bands[k].joff, S.counts.ptr + 4 * k + 2, s2),
                  "judge");
        }
        check(launch_finalize(ctx->sample_rgb.ptr, ctx->sample_cls.ptr, ctx->frame.ptr, npix, nS, s2), "finalize");
        check(hipEventRecord(ctx->ev[2], s2), "event");
        check(hipEventSynchronize(ctx->ev[2]), "surrogate frame");
        std::vector<unsigned long long> cnt(static_cast<size_t>(4 * nb));
        check(hipMemcpy(cnt.data(), S.counts.ptr, sizeof(unsigned long long) * cnt.size(), hipMemcpyDeviceToHost),
              "read counts");
        std::vector<DevCounters> cs(static_cast<size_t>(nb));
        check(hipMemcpy(cs.data(), ctx->counters.ptr, sizeof(DevCounters) * nb, hipMemcpyDeviceToHost),
              "read counters");
        unsigned long long n_fb = 0, nj = 0, n_agree = 0, attempts = 0;
        std::vector<float> errs;
        for (int k = 0; k < nb; ++k) {
            const unsigned long long jk =
                std::min<unsigned long long>(cnt[4 * k + 1], static_cast<unsigned long long>(bands[k].cap));
            n_fb += cnt[4 * k];
            nj += jk;
            n_agree += cnt[4 * k + 2];
            attempts += cs[k].attempts;
            const size_t o = errs.size();
            errs.resize(o + jk);
            if (jk)
                check(hipMemcpy(errs.data() + o, S.judge_err.ptr + bands[k].joff, sizeof(float) * jk,
                                hipMemcpyDeviceToHost),
                      "judge");
        }
        S.last_fb = n_fb;
        S.last_judged = nj;
        if (out)
            check(hipMemcpy(out, ctx->frame.ptr, sizeof(float) * 4 * npix,
                            device_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost),
                  "copy frame");
        float ms[2] = {0.0f, 0.0f};
        check(hipEventElapsedTime(&ms[0], ctx->ev[0], ctx->ev[1]), "event time");
        check(hipEventElapsedTime(&ms[1], ctx->ev[0], ctx->ev[2]), "event time");
        std::vector<float> lit;
        lit.reserve(errs.size());
        for (float e : errs)
            if (e >= 0.0f) lit.push_back(e);
        double med = 0.0;
        if (!lit.empty()) {
            std::nth_element(lit.begin(), lit.begin() + lit.size() / 2, lit.end());
            med = lit[lit.size() / 2];
        }
        if (stats) {
            GRRTSurrogateStats& s = *stats;
            s = GRRTSurrogateStats{};
            s.infer_ms = ms[0];
            s.trace_ms = std::max(0.0f, ms[1] - ms[0]);
            s.gpu_ms = ms[1];
            s.rays = n_rays;
            s.fallback_rays = static_cast<int64_t>(n_fb);
            s.judged_rays = static_cast<int64_t>(nj);
            s.trace_attempts = static_cast<int64_t>(attempts);
            s.judge_class_agreement = nj ? static_cast<double>(n_agree) / nj : 1.0;
            s.judge_median_rel_err = med;
            s.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        ctx->last_rays = n_rays;
        ctx->last_batch = false;
        ctx->last_cls = ctx->sample_cls.ptr;
        ctx->last_rgb = ctx->sample_rgb.ptr;
        ctx->records_valid = false;
        return 0;
    });
}

}  // extern "C"
