// post.hip -- the steps after the solve: the full displacement from the free-dof solution,
// reconstruction.f:22-135 (U = scatter of the free values, prescribed values of the boundary
// conditions in their order, sol = U reshaped to (nb_cp, MCRD) column-major), and the
// displacement at parametric points, evaluate_disp.f:20-189.
#include <algorithm>
#include <map>
#include <tuple>

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

constexpr int EVAL_MAXP = 8;

// BINARYSEARCH of igaparmetrization.f90:291-319 over uk[0..n-1] (1-based result, intervals
// [left, right); a value on no interval ends with the last midpoint, as there)
__device__ inline int eval_search(double val, const double *uk, int n) {
    int low = 1, high = n, middle = 1;
    bool found = false;
    while (low != high && !found) {
        middle = (low + high) / 2;
        if (val >= uk[middle - 1] && val < uk[middle]) found = true;
        else if (val < uk[middle - 1]) high = middle;
        else low = middle + 1;
    }
    return middle;
}

// B-spline values N_{span-p..span} at u (the triangular recurrence of dersbasisfuns.f90:30-131,
// values only); span1 is the 1-based span index
__device__ inline void eval_basis(int span1, int p, double u, const double *U, double *N) {
    const int i = span1 - 1;
    double left[EVAL_MAXP + 1], right[EVAL_MAXP + 1];
    N[0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[i + 1 - j];
        right[j] = U[i + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            const double tmp = N[r] / (right[r + 1] + left[j - r]);
            N[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        N[j] = saved;
    }
}

struct EvalPatch {
    int p[3], nkv[3], nint[3];
    const double *U[3];
    const int32_t *nijk, *ien, *table;   // table: interval triple -> element (0-based)
    const double *weight;
    int nnode;
};

// one thread per point (evaluate_disp.f:83-97 with evalnurbs, evaluatenurbsfcts.f:94-201)
__global__ void __launch_bounds__(256) evaldisp_kernel(EvalPatch ep, int32_t mcrd, const double *sol, int64_t npts,
                                                       const double *xi, double *disp) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npts; t += (int64_t)gridDim.x * blockDim.x) {
        int iv[3];
        for (int d = 0; d < 3; ++d) {
            const int idx = eval_search(xi[3 * t + d], ep.U[d] + ep.p[d], ep.nkv[d] - 2 * ep.p[d]);   // span - p
            iv[d] = min(max(idx - 1, 0), ep.nint[d] - 1);
        }
        const int e = ep.table[iv[0] + ep.nint[0] * (iv[1] + ep.nint[1] * iv[2])];
        double B[3][EVAL_MAXP + 1];
        for (int d = 0; d < 3; ++d) eval_basis(ep.nijk[3 * e + d], ep.p[d], xi[3 * t + d], ep.U[d], B[d]);
        double sum = 0.0;
        int a = 0;
        for (int k = 0; k <= ep.p[2]; ++k)
            for (int j = 0; j <= ep.p[1]; ++j)
                for (int i = 0; i <= ep.p[0]; ++i, ++a)
                    sum += B[0][ep.p[0] - i] * B[1][ep.p[1] - j] * B[2][ep.p[2] - k] * ep.weight[(int64_t)ep.nnode * e + a];
        const double inv = 1.0 / sum;
        double u[3] = {0.0, 0.0, 0.0};
        a = 0;
        for (int k = 0; k <= ep.p[2]; ++k)
            for (int j = 0; j <= ep.p[1]; ++j)
                for (int i = 0; i <= ep.p[0]; ++i, ++a) {
                    const double r = B[0][ep.p[0] - i] * B[1][ep.p[1] - j] * B[2][ep.p[2] - k] *
                                     ep.weight[(int64_t)ep.nnode * e + a] * inv;
                    const int64_t cp = (int64_t)ep.ien[(int64_t)ep.nnode * e + a] - 1;
                    for (int c = 0; c < mcrd && c < 3; ++c) u[c] += r * sol[cp * mcrd + c];
                }
        for (int c = 0; c < 3; ++c) disp[3 * t + c] = u[c];
    }
}

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

extern "C" int32_t yeti_evaldisp(const yeti_nurbs_patch *patch, int32_t nb_cp, int32_t mcrd, const double *sol,
                                 int64_t npts, const double *xi, double *disp) {
    if (!patch || !sol || (npts > 0 && (!xi || !disp)) || mcrd < 1 || nb_cp < 1 || npts < 0)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "evaldisp: bad arguments");
    const yeti_nurbs_patch &P = *patch;
    if (P.nelem < 1 || !P.nijk || !P.ien || !P.weight)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "evaldisp: empty patch");
    int nn = 1;
    EvalPatch ep{};
    for (int d = 0; d < 3; ++d) {
        if (P.p[d] < 0 || P.p[d] > EVAL_MAXP || !P.U[d] || P.nkv[d] < 2 * P.p[d] + 2)
            return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "evaldisp: degree %d / %d knots in direction %d", P.p[d],
                             P.nkv[d], d + 1);
        ep.p[d] = P.p[d];
        ep.nkv[d] = P.nkv[d];
        ep.nint[d] = P.nkv[d] - 2 * P.p[d] - 1;
        nn *= P.p[d] + 1;
    }
    if (nn != P.nnode) return set_error(YETI_ERR_INVALID_ARGUMENT, "evaldisp: nnode %d, degrees give %d", P.nnode, nn);
    // interval triple -> element: the unique element with those spans, else element 1
    // (igaparmetrization.f90:357-368)
    const int64_t ntab = (int64_t)ep.nint[0] * ep.nint[1] * ep.nint[2];
    std::vector<int32_t> table(ntab, -1), count(ntab, 0);
    for (int32_t e = 0; e < P.nelem; ++e) {
        int iv[3];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {
            iv[d] = P.nijk[3 * (int64_t)e + d] - P.p[d] - 1;
            ok &= iv[d] >= 0 && iv[d] < ep.nint[d];
        }
        if (!ok) continue;
        const int64_t k = iv[0] + (int64_t)ep.nint[0] * (iv[1] + (int64_t)ep.nint[1] * iv[2]);
        if (count[k]++ == 0) table[k] = e;
    }
    for (int64_t k = 0; k < ntab; ++k)
        if (count[k] != 1) table[k] = 0;
    for (int32_t e = 0; e < P.nelem; ++e)
        for (int a = 0; a < P.nnode; ++a) {
            const int32_t cp = P.ien[(int64_t)P.nnode * e + a];
            if (cp < 1 || cp > nb_cp) return set_error(YETI_ERR_INVALID_ARGUMENT, "evaldisp: IEN entry out of range");
        }
    if (npts == 0) return 0;
    DevBuf<double> dU[3], dw, dsol, dxi, ddisp;
    DevBuf<int32_t> dnijk, dien, dtab;
    hipError_t er = hipSuccess;
    for (int d = 0; d < 3 && er == hipSuccess; ++d) er = dU[d].upload(P.U[d], P.nkv[d]);
    if (er == hipSuccess) er = dnijk.upload(P.nijk, 3 * (size_t)P.nelem);
    if (er == hipSuccess) er = dien.upload(P.ien, (size_t)P.nnode * P.nelem);
    if (er == hipSuccess) er = dw.upload(P.weight, (size_t)P.nnode * P.nelem);
    if (er == hipSuccess) er = dtab.upload(table.data(), table.size());
    if (er == hipSuccess) er = dsol.upload(sol, (size_t)mcrd * nb_cp);
    if (er == hipSuccess) er = dxi.upload(xi, 3 * (size_t)npts);
    if (er == hipSuccess) er = ddisp.alloc(3 * (size_t)npts);
    if (er != hipSuccess) return set_error(YETI_ERR_DEVICE, "evaldisp: device memory: %s", hipGetErrorString(er));
    for (int d = 0; d < 3; ++d) ep.U[d] = dU[d].p;
    ep.nijk = dnijk.p;
    ep.ien = dien.p;
    ep.weight = dw.p;
    ep.table = dtab.p;
    ep.nnode = P.nnode;
    hipLaunchKernelGGL(evaldisp_kernel, dim3(grid_for(npts)), dim3(256), 0, 0, ep, mcrd, (const double *)dsol.p, npts,
                       (const double *)dxi.p, ddisp.p);
    YETI_HIP(hipGetLastError());
    YETI_HIP(hipMemcpy(disp, ddisp.p, 3 * (size_t)npts * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}
