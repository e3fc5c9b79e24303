// Opacity tables of the volumetric disk computed on the GPU
// (src/opacity.cpp:237-303): one lane per (rho, T) cell runs the cell's Saha
// solve, its absorption opacities at the table frequencies and its
// 50-frequency Rosseland mean, with the same point functions as the host
// builder (opacity_math.hpp).  The cells are independent and branchy FP64
// work (Saha passes, exp/pow per frequency), so the kernel is a plain
// one-lane-per-cell launch in 64-lane blocks spread over the CUs.
#include <hip/hip_runtime.h>

#include "volumetric_host.hpp"

namespace grrt_amd {

namespace {

__global__ __launch_bounds__(64) void opacity_cells(int n_rho, int n_T, double lr0, double lr1, double lt0,
                                                    double lt1, int n_nu, double ln0, double ln1, double* kabs,
                                                    double* es, double* mu, double* ross) {
    const int cell = static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x);
    const int n2 = n_rho * n_T;
    if (cell >= n2) return;
    const int j = cell / n_T, k = cell - j * n_T;
    const double rho = pow(10.0, lr0 + j * (lr1 - lr0) / (n_rho - 1));
    const double T = pow(10.0, lt0 + k * (lt1 - lt0) / (n_T - 1));
    double e, m, r;
    opacity_cell(rho, T, n_nu, ln0, ln1, kabs + cell, n2, e, m, r);
    es[cell] = e;
    mu[cell] = m;
    ross[cell] = r;
}

}  // namespace

hipError_t fill_opacity_tables_device(OpacityTables& o, hipStream_t st) {
    const size_t n2 = static_cast<size_t>(o.n_rho) * o.n_T;
    if (o.kappa_abs.size() != n2 * o.n_nu || o.kappa_es.size() != n2 || o.mu.size() != n2 ||
        o.kappa_ross.size() != n2)
        return hipErrorInvalidValue;
    double* d = nullptr;
    const size_t total = n2 * (o.n_nu + 3);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), total * sizeof(double));
    if (e != hipSuccess) return e;
    double* kabs = d;
    double* es = kabs + n2 * o.n_nu;
    double* mu = es + n2;
    double* ross = mu + n2;
    const unsigned blocks = static_cast<unsigned>((n2 + 63) / 64);
    opacity_cells<<<blocks, 64, 0, st>>>(o.n_rho, o.n_T, o.log_rho_min, o.log_rho_max, o.log_T_min, o.log_T_max,
                                         o.n_nu, o.log_nu_min, o.log_nu_max, kabs, es, mu, ross);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(o.kappa_abs.data(), kabs, n2 * o.n_nu * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(o.kappa_es.data(), es, n2 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(o.mu.data(), mu, n2 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(o.kappa_ross.data(), ross, n2 * sizeof(double), hipMemcpyDeviceToHost, st);
    const hipError_t es2 = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es2;
    const hipError_t ef = hipFree(d);
    return e != hipSuccess ? e : ef;
}

}  // namespace grrt_amd
