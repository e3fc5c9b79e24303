// Copy-kernel calibration for HBM traffic counters: copies N bytes with a
// plain vectorised kernel so that rocprofv3's FETCH_SIZE / WRITE_SIZE for a
// known number of bytes can scale the same counters read on other kernels.
// Usage: calib_copy [MiB]   (default 256)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_u4(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 256;
    const size_t bytes = mib << 20;
    const size_t n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(copy_u4, dim3(4096), dim3(256), 0, 0, a, b, n);   // warm
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(copy_u4, dim3(4096), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    std::printf("calib_copy: %zu MiB read + %zu MiB written in %.3f ms (%.1f GB/s)\n", mib, mib, ms,
                2.0 * bytes / (ms * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
