// Tone mapping of an RGBA float32 frame in device memory: the log-average
// auto-exposure, extended Reinhard curve with the frame's white point and
// gamma 2.2 of src/tonemapper.cpp:7-71, as a two-pass HBM-bound operation.
//   pass 1: per-block partial sums of ln(1e-6 + L) over lit pixels
//           (L > 1e-10), lit counts and maximum L; a second single-block
//           kernel folds the partials in a fixed order (the result is the
//           same on every run) and derives exposure and white point;
//   pass 2: per pixel, the curve in FP64, written back as float (alpha is
//           kept).
// Bytes per pixel of the algorithm: 16 read in pass 1, 16 read and 12
// written in pass 2.
#include <hip/hip_runtime.h>

#include "tonemap.hpp"

namespace grrt_amd {

namespace {

constexpr int kTmThreads = 256;

__device__ __forceinline__ double luminance(const float4& p, double& R, double& G, double& B) {
    R = fmax(0.0, static_cast<double>(p.x));
    G = fmax(0.0, static_cast<double>(p.y));
    B = fmax(0.0, static_cast<double>(p.z));
    return 0.2126 * R + 0.7152 * G + 0.0722 * B;
}

// Tree reduction over the block in shared memory, fixed order.
template <class T, class Op>
__device__ T block_reduce(T v, T* sh, Op op) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = kTmThreads / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = op(sh[t], sh[t + s]);
        __syncthreads();
    }
    const T r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kTmThreads) void tonemap_reduce(const float4* __restrict__ fb, long long n,
                                                             TonePartial* __restrict__ part) {
    __shared__ double shd[kTmThreads];
    __shared__ unsigned long long shu[kTmThreads];
    double log_sum = 0.0, lmax = 0.0;
    unsigned long long lit = 0;
    const long long stride = static_cast<long long>(gridDim.x) * kTmThreads;
    for (long long i = static_cast<long long>(blockIdx.x) * kTmThreads + threadIdx.x; i < n; i += stride) {
        double R, G, B;
        const double L = luminance(fb[i], R, G, B);
        if (L > 1e-10) {
            log_sum += log(1e-6 + L);
            ++lit;
        }
        if (L > lmax) lmax = L;
    }
    auto add = [](double a, double b) { return a + b; };
    auto mx = [](double a, double b) { return a > b ? a : b; };
    auto addu = [](unsigned long long a, unsigned long long b) { return a + b; };
    log_sum = block_reduce(log_sum, shd, add);
    lmax = block_reduce(lmax, shd, mx);
    lit = block_reduce(lit, shu, addu);
    if (threadIdx.x == 0) part[blockIdx.x] = TonePartial{log_sum, lmax, lit};
}

__global__ __launch_bounds__(kTmThreads) void tonemap_final(const TonePartial* __restrict__ part, int nb,
                                                            ToneParams* __restrict__ tp) {
    __shared__ double shd[kTmThreads];
    __shared__ unsigned long long shu[kTmThreads];
    double log_sum = 0.0, lmax = 0.0;
    unsigned long long lit = 0;
    for (int b = threadIdx.x; b < nb; b += kTmThreads) {
        log_sum += part[b].log_sum;
        lmax = part[b].lmax > lmax ? part[b].lmax : lmax;
        lit += part[b].lit;
    }
    auto add = [](double a, double b) { return a + b; };
    auto mx = [](double a, double b) { return a > b ? a : b; };
    auto addu = [](unsigned long long a, unsigned long long b) { return a + b; };
    log_sum = block_reduce(log_sum, shd, add);
    lmax = block_reduce(lmax, shd, mx);
    lit = block_reduce(lit, shu, addu);
    if (threadIdx.x == 0) {
        ToneParams p{};
        p.active = (lmax > 0.0 && lit > 0) ? 1 : 0;     // an all-black frame is left as it is
        if (p.active) {
            const double L_avg = exp(log_sum / static_cast<double>(lit));
            p.exposure = 0.18 / L_avg;
            const double lw = lmax * p.exposure;
            p.lw2 = lw * lw;
        }
        *tp = p;
    }
}

__device__ __forceinline__ float gamma22(double v) {
    return static_cast<float>(pow(fmin(fmax(v, 0.0), 1.0), 1.0 / 2.2));
}

__global__ __launch_bounds__(kTmThreads) void tonemap_apply(float4* __restrict__ fb, long long n,
                                                            const ToneParams* __restrict__ tp) {
    const ToneParams p = *tp;
    if (!p.active) return;
    const long long stride = static_cast<long long>(gridDim.x) * kTmThreads;
    for (long long i = static_cast<long long>(blockIdx.x) * kTmThreads + threadIdx.x; i < n; i += stride) {
        float4 v = fb[i];
        double R, G, B;
        const double L = luminance(v, R, G, B);
        if (L <= 0.0) {
            v.x = v.y = v.z = 0.0f;
        } else {
            const double ls = L * p.exposure;
            const double lm = fmin(fmax(ls * (1.0 + ls / p.lw2) / (1.0 + ls), 0.0), 1.0);
            const double sc = lm / L;
            v.x = gamma22(R * sc);
            v.y = gamma22(G * sc);
            v.z = gamma22(B * sc);
        }
        fb[i] = v;
    }
}

}  // namespace

int tonemap_blocks(long long n_pixels) {
    const long long b = (n_pixels + kTmThreads - 1) / kTmThreads;
    return static_cast<int>(b < 1 ? 1 : (b > kToneMaxBlocks ? kToneMaxBlocks : b));
}

hipError_t launch_tonemap(float* d_rgba, long long n_pixels, TonePartial* part, ToneParams* tp, hipStream_t st) {
    if (n_pixels <= 0) return hipSuccess;
    const int nb = tonemap_blocks(n_pixels);
    float4* fb = reinterpret_cast<float4*>(d_rgba);
    hipLaunchKernelGGL(tonemap_reduce, dim3(nb), dim3(kTmThreads), 0, st, fb, n_pixels, part);
    hipLaunchKernelGGL(tonemap_final, dim3(1), dim3(kTmThreads), 0, st, part, nb, tp);
    hipLaunchKernelGGL(tonemap_apply, dim3(nb), dim3(kTmThreads), 0, st, fb, n_pixels, tp);
    return hipGetLastError();
}

}  // namespace grrt_amd
