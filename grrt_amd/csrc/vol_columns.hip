// The volumetric disk's vertical columns on the GPU: one wavefront per
// column, the per-node steps of vol_column.hpp over its lanes (including
// every right-hand-side term that does not depend on the density) and the
// two running sums (optical depth from the top, the RK4 density march) on
// lane 0, the column's arrays in LDS.  Equal bit for bit to the host build over
// the same opacity tables.
#include <hip/hip_runtime.h>

#include "vol_column.hpp"

namespace grrt_amd {

namespace {

__global__ __launch_bounds__(64) void vol_columns(ColumnTables t, const ColumnIn* __restrict__ in, int n_z,
                                                  double tau_mid, double* __restrict__ z_max_out,
                                                  double* __restrict__ rho_out, double* __restrict__ T_out) {
#pragma clang fp contract(off)
    __shared__ double rho[kColMaxZ], T[kColMaxZ], tau[kColMaxZ], E[kColMaxZ], f[kColMaxZ], mu[kColMaxZ],
        dcs2[kColMaxZ], dfE[kColMaxZ];
    __shared__ RhsTerms q0[kColMaxZ], qh[kColMaxZ], q1[kColMaxZ];
    __shared__ double s_zmax;
    __shared__ int s_more;
    const int ri = blockIdx.x, lane = threadIdx.x, NZ = n_z;
    const ColumnIn c = in[ri];
    double* rp = rho_out + static_cast<long long>(ri) * NZ;
    double* tp = T_out + static_cast<long long>(ri) * NZ;
    if (c.H <= 0.0 || c.Te <= 0.0 || c.rm <= 0.0) {
        for (int z = lane; z < NZ; z += 64) {
            rp[z] = z == 0 ? 1.0 : 0.0;
            tp[z] = c.Te;
        }
        if (lane == 0) z_max_out[ri] = 3.0 * c.H;
        return;
    }
    const double rho_ref = col_rho_ref(t, c, tau_mid);
    const double Te4 = c.Te * c.Te * c.Te * c.Te;
    const double T_mid = col_T(Te4, tau_mid);
    if (lane == 0) s_zmax = 3.0 * c.H;
    __syncthreads();
    for (int iter = 0; iter < 8; ++iter) {
        const double zmax = s_zmax;
        const double dz = zmax / (NZ - 1);
        const double step = col_tau_step(t, rho_ref, T_mid, dz);
        if (lane == 0) {
            tau[NZ - 1] = 0.0;
            for (int z = NZ - 2; z >= 0; --z) tau[z] = tau[z + 1] + step;
        }
        __syncthreads();
        for (int z = lane; z < NZ; z += 64) {
            T[z] = col_T(Te4, tau[z]);
            E[z] = col_E(T[z]);
            mu[z] = col_mu(t, rho_ref, T[z]);
        }
        __syncthreads();
        for (int z = lane; z < NZ; z += 64) f[z] = col_f(t, E, T, z, NZ, dz, rho_ref);
        __syncthreads();
        for (int z = lane; z < NZ; z += 64) col_grads(T, mu, f, E, z, NZ, dz, dcs2[z], dfE[z]);
        __syncthreads();
        for (int z = lane; z < NZ - 1; z += 64)
            col_step_terms(T, mu, dcs2, dfE, NZ, dz, c.Omz2, rho_ref, z, q0[z], qh[z], q1[z]);
        __syncthreads();
        if (lane == 0) {
            col_rk4(q0, qh, q1, NZ, dz, rho);
            double zm = zmax;
            s_more = col_raise_top(rho, NZ, c.H, zm) ? 1 : 0;
            s_zmax = zm;
        }
        __syncthreads();
        if (!s_more) break;
        __syncthreads();   // every lane has read s_more before lane 0 rewrites it
    }
    for (int z = lane; z < NZ; z += 64) {
        rp[z] = rho[z];
        tp[z] = T[z];
    }
    if (lane == 0) z_max_out[ri] = s_zmax;
}

}  // namespace

hipError_t launch_vol_columns(const ColumnTables& t, const ColumnIn* in, int n_r, int n_z, double tau_mid,
                              double* z_max, double* rho, double* T, hipStream_t st) {
    if (n_z < 2 || n_z > kColMaxZ || n_r < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(vol_columns, dim3(static_cast<unsigned>(n_r)), dim3(64), 0, st, t, in, n_z, tau_mid, z_max,
                       rho, T);
    return hipGetLastError();
}

}  // namespace grrt_amd
