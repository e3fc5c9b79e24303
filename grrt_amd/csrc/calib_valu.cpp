// VALU issue-rate calibration: cycles per wave64 instruction of FP32 FMA,
// the hardware exp2 and reciprocal (the two transcendental operations of
// the surrogate's SiLU), and FP64 FMA, each measured on its own with eight
// independent dependency chains per lane so that latency is hidden.
// Usage: calib_valu [MHz]   (default 2400, the engine clock used to convert
// time to cycles)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

constexpr int kChains = 8;
constexpr int kIters = 4096;
constexpr int kThreads = 256;

template <int kOp>
__global__ __launch_bounds__(kThreads) void rate_kernel(float* out32, double* out64) {
    const unsigned tid = blockIdx.x * kThreads + threadIdx.x;
    if constexpr (kOp == 3) {
        double x[kChains];
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = 0.25 + 1e-3 * c + 1e-9 * tid;
        for (int i = 0; i < kIters; ++i) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) x[c] = __builtin_fma(x[c], 0.999, 1e-4);
        }
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < kChains; ++c) s += x[c];
        out64[tid] = s;
    } else {
        float x[kChains];
#pragma unroll
        for (int c = 0; c < kChains; ++c) x[c] = 0.25f + 1e-3f * c + 1e-7f * (tid & 1023);
        for (int i = 0; i < kIters; ++i) {
#pragma unroll
            for (int c = 0; c < kChains; ++c) {
                if constexpr (kOp == 0) x[c] = __builtin_fmaf(x[c], 0.999f, 1e-4f);
                if constexpr (kOp == 1) x[c] = __builtin_amdgcn_exp2f(-x[c]);     // stays in [0.5, 1]
                if constexpr (kOp == 2) x[c] = __builtin_amdgcn_rcpf(x[c]);          // alternates x, 1/x
            }
        }
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < kChains; ++c) s += x[c];
        out32[tid] = s;
    }
}

template <int kOp>
static int measure(const char* name, int blocks, float* o32, double* o64, double mhz, int n_simd) {
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(rate_kernel<kOp>, dim3(blocks), dim3(kThreads), 0, 0, o32, o64);   // warm
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(rate_kernel<kOp>, dim3(blocks), dim3(kThreads), 0, 0, o32, o64);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double wave_instr = static_cast<double>(blocks) * (kThreads / 64) * kIters * kChains;
    const double simd_cycles = ms * 1e-3 * mhz * 1e6 * n_simd;
    std::printf("%-10s %8.3f ms  %6.2f cycles per wave64 instruction per SIMD\n", name, ms, simd_cycles / wave_instr);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

int main(int argc, char** argv) {
    const double mhz = argc > 1 ? std::atof(argv[1]) : 2400.0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
    const int n_simd = prop.multiProcessorCount * 4;
    const int blocks = prop.multiProcessorCount * 16;   // 16 waves per SIMD: issue-bound, not latency-bound
    float* o32 = nullptr;
    double* o64 = nullptr;
    if (hipMalloc(&o32, sizeof(float) * blocks * kThreads) != hipSuccess) return 1;
    if (hipMalloc(&o64, sizeof(double) * blocks * kThreads) != hipSuccess) return 1;
    std::printf("%d CUs, %d SIMDs, clock taken as %.0f MHz\n", prop.multiProcessorCount, n_simd, mhz);
    int rc = measure<0>("fma_f32", blocks, o32, o64, mhz, n_simd);
    rc |= measure<1>("exp2_f32", blocks, o32, o64, mhz, n_simd);
    rc |= measure<2>("rcp_f32", blocks, o32, o64, mhz, n_simd);
    rc |= measure<3>("fma_f64", blocks, o32, o64, mhz, n_simd);
    (void)hipFree(o32);
    (void)hipFree(o64);
    return rc;
}
