// post.hip -- the step after the solve: the full displacement from the free-dof solution,
// reconstruction.f:22-135 (U = scatter of the free values, prescribed values of the boundary
// conditions in their order, sol = U reshaped to (nb_cp, MCRD) column-major).
#include <algorithm>

#include "common.hpp"

namespace yeti {

__global__ void __launch_bounds__(256) recon_fill_kernel(int64_t n, double *U) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        U[i] = 0.0;
}

// U[ind_dof_free[d] - 1] = u_inv[d] (reconstruction.f:60-62); the free indices are distinct
__global__ void __launch_bounds__(256) recon_free_kernel(int32_t nfree, const double *u_inv, const int32_t *ind,
                                                         int64_t ndof, double *U) {
    for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < nfree; d += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = (int64_t)ind[d] - 1;
        if (g >= 0 && g < ndof) U[g] = u_inv[d];
    }
}

// one boundary condition: U[(target - 1) mcrd + dir - 1] = value (reconstruction.f:64-74)
__global__ void __launch_bounds__(256) recon_bc_kernel(int32_t n, const int32_t *targets, int32_t mcrd, int32_t dir,
                                                       double value, int64_t ndof, double *U) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = ((int64_t)targets[k] - 1) * mcrd + dir - 1;
        if (g >= 0 && g < ndof) U[g] = value;
    }
}

// sol(cp, i) = U(cp mcrd + i), sol column-major (reconstruction.f:131-135)
__global__ void __launch_bounds__(256) recon_sol_kernel(int32_t nb_cp, int32_t mcrd, const double *U, double *sol) {
    const int64_t n = (int64_t)nb_cp * mcrd;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / nb_cp, cp = t % nb_cp;
        sol[t] = U[cp * mcrd + i];
    }
}

static unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256)); }

}  // namespace yeti

using namespace yeti;

extern "C" int32_t yeti_reconstruction(const double *u_inv, const int32_t *ind_dof_free, int32_t nb_dof_free,
                                       const int32_t *bc_target, const int32_t *bc_target_nbelem_host,
                                       const double *bc_values_host, int32_t nb_bc, int32_t mcrd, int32_t nb_cp,
                                       double *sol, double *U, void *stream) {
    if (mcrd < 1 || nb_cp < 0 || nb_dof_free < 0 || nb_bc < 0 || !sol || !U)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "reconstruction: bad sizes or null output");
    if ((nb_dof_free > 0 && (!u_inv || !ind_dof_free)) || (nb_bc > 0 && (!bc_target_nbelem_host || !bc_values_host)))
        return set_error(YETI_ERR_INVALID_ARGUMENT, "reconstruction: null input");
    hipStream_t st = (hipStream_t)stream;
    const int64_t ndof = (int64_t)nb_cp * mcrd;
    if (ndof == 0) return 0;
    hipLaunchKernelGGL(recon_fill_kernel, dim3(grid_for(ndof)), dim3(256), 0, st, ndof, U);
    if (nb_dof_free > 0)
        hipLaunchKernelGGL(recon_free_kernel, dim3(grid_for(nb_dof_free)), dim3(256), 0, st, nb_dof_free, u_inv,
                           ind_dof_free, ndof, U);
    int64_t count = 0;
    for (int32_t b = 0; b < nb_bc; ++b) {
        const int32_t n = bc_target_nbelem_host[b];
        if (n < 0) return set_error(YETI_ERR_INVALID_ARGUMENT, "reconstruction: negative target count");
        if (bc_values_host[2 * b] <= 3.0 && n > 0) {
            if (!bc_target) return set_error(YETI_ERR_INVALID_ARGUMENT, "reconstruction: null bc_target");
            hipLaunchKernelGGL(recon_bc_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, bc_target + count, mcrd,
                               (int32_t)bc_values_host[2 * b], bc_values_host[2 * b + 1], ndof, U);
        }
        count += n;
    }
    hipLaunchKernelGGL(recon_sol_kernel, dim3(grid_for(ndof)), dim3(256), 0, st, nb_cp, mcrd, (const double *)U, sol);
    YETI_HIP(hipGetLastError());
    return 0;
}
