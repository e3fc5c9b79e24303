// MFMA issue-rate calibration for the surrogate's network pass: cycles per
// v_mfma_f32_32x32x16_bf16 per wave (clock64 over the loop) and the chip's
// rate, for
//   - dependent accumulation chains with both operands in registers;
//   - the same chains interleaved with SiLU work (hardware exp2 and
//     reciprocal, the network's activation) on independent values;
//   - the A operand (weights) read from a 16 KB LDS window before each MFMA
//     with a rolling prefetch of three, as in the network pass, with one or
//     two 32-ray tiles sharing each weight read.
// Each form runs at one and at two wavefronts per SIMD (256-thread blocks,
// one or two per CU).  Every wave writes one cycle count and one checksum.
// Usage: calib_mfma
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

__device__ __forceinline__ f32x16 mfma(const bf16x8& a, const bf16x8& b, const f32x16& c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

constexpr int kIters = 2000;

// kTiles accumulation chains; kSilu SiLU evaluations per MFMA; kLds: the A
// operand comes from the LDS window (prefetch distance 3) instead of a register.
template <int kTiles, int kSilu, bool kLds>
__global__ __launch_bounds__(256) void mfma_kernel(float* out, long long* cyc, int iters) {
    __shared__ bf16x8 W[16 * 64];
    const int lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < 16 * 64; i += 256)
        for (int j = 0; j < 8; ++j) W[i][j] = static_cast<__bf16>(0.001f * ((i + j) & 31));
    __syncthreads();
    bf16x8 a, b[kTiles][16];
    for (int j = 0; j < 8; ++j) a[j] = static_cast<__bf16>(0.001f * (lane + j));
    for (int t = 0; t < kTiles; ++t)
        for (int s = 0; s < 16; ++s)
            for (int j = 0; j < 8; ++j) b[t][s][j] = static_cast<__bf16>(0.002f * (lane - j + s + t));
    f32x16 c[kTiles];
    for (int t = 0; t < kTiles; ++t)
        for (int v = 0; v < 16; ++v) c[t][v] = 0.01f * v;
    float h[16];
    for (int v = 0; v < 16; ++v) h[v] = 0.1f * (lane + v);
    const bf16x8* Wl = W + lane;     // entries lane + 64 s, s < 16: inside the window
    const long long t0 = clock64();
    for (int it = 0; it < iters; ++it) {
        bf16x8 w0 = a, w1 = a, w2 = a, w3 = a;
        if (kLds) {
            w0 = Wl[0];
            w1 = Wl[64];
            w2 = Wl[128];
        }
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            if (kLds && s + 3 < 16) w3 = Wl[(s + 3) * 64];
#pragma unroll
            for (int t = 0; t < kTiles; ++t) c[t] = mfma(w0, b[t][s], c[t]);
#pragma unroll
            for (int q = 0; q < kSilu * kTiles; ++q) h[(s * 3 + q) & 15] = silu(h[(s * 3 + q) & 15]);
            if (kLds) {
                w0 = w1;
                w1 = w2;
                w2 = w3;
            }
        }
        asm volatile("" ::: "memory");
    }
    const long long t1 = clock64();
    float sum = 0.0f;
    for (int t = 0; t < kTiles; ++t)
        for (int v = 0; v < 16; ++v) sum += c[t][v];
    for (int v = 0; v < 16; ++v) sum += h[v];
    const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (lane == 0) {
        out[wave] = sum;
        cyc[wave] = t1 - t0;
    }
}

template <int kTiles, int kSilu, bool kLds>
static bool run(int blocks, int per_simd, const char* name) {
    const int waves = blocks * 4;
    float* out = nullptr;
    long long* cyc = nullptr;
    if (hipMalloc(&out, waves * sizeof(float)) != hipSuccess || hipMalloc(&cyc, waves * sizeof(long long)) != hipSuccess)
        return false;
    hipLaunchKernelGGL((mfma_kernel<kTiles, kSilu, kLds>), dim3(blocks), dim3(256), 0, 0, out, cyc, 4);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return false;
    if (hipEventRecord(e0) != hipSuccess) return false;
    hipLaunchKernelGGL((mfma_kernel<kTiles, kSilu, kLds>), dim3(blocks), dim3(256), 0, 0, out, cyc, kIters);
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return false;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return false;
    std::vector<long long> h(waves);
    if (hipMemcpy(h.data(), cyc, waves * sizeof(long long), hipMemcpyDeviceToHost) != hipSuccess) return false;
    std::sort(h.begin(), h.end());
    const double per_wave = 16.0 * kTiles * kIters;
    std::printf("%-44s %d wave(s)/SIMD: %6.1f cycles per MFMA per wave, %7.1f TFLOP/s\n", name, per_simd,
                h[waves / 2] / per_wave, per_wave * waves * 32768.0 / (ms * 1e-3) / 1e12);
    (void)hipFree(out);
    (void)hipFree(cyc);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return true;
}

int main() {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus <= 0) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    std::printf("CUs %d; 32768 flops per MFMA\n", cus);
    bool ok = true;
    for (int per_simd = 1; per_simd <= 2; ++per_simd) {
        const int blocks = cus * per_simd;
        ok = ok && run<1, 0, false>(blocks, per_simd, "registers: 1 dependent chain");
        ok = ok && run<2, 0, false>(blocks, per_simd, "registers: 2 chains");
        ok = ok && run<1, 1, false>(blocks, per_simd, "registers: 1 chain + 1 SiLU per MFMA");
        ok = ok && run<1, 2, false>(blocks, per_simd, "registers: 1 chain + 2 SiLU per MFMA");
        ok = ok && run<1, 4, false>(blocks, per_simd, "registers: 1 chain + 4 SiLU per MFMA");
        ok = ok && run<1, 0, true>(blocks, per_simd, "LDS weights: 1 tile");
        ok = ok && run<2, 0, true>(blocks, per_simd, "LDS weights: 2 tiles per read");
        ok = ok && run<1, 1, true>(blocks, per_simd, "LDS weights: 1 tile + 1 SiLU per MFMA");
        ok = ok && run<2, 1, true>(blocks, per_simd, "LDS weights: 2 tiles + 1 SiLU per MFMA each");
    }
    return ok ? 0 : 1;
}
