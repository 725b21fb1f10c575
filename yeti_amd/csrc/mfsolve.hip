// mfsolve.hip -- matrix-free solves of K[free][:, free] u = F[free]:
//   * yeti_mf_pcg_f64: Jacobi-type PCG in FP64 with the FP64 matrix-free apply and a 3x3 nodal
//     block-Jacobi preconditioner (the plain FP64 GPU path);
//   * yeti_mf_solve_ir: FP64 iterative refinement whose corrections come from the same PCG run
//     in FP32 (FP32 apply, FP32 block-Jacobi, FP32 vectors) while the refinement residual
//     b - K x, the solution update and every CG dot product stay in FP64.
// Reference pattern: the CG loop of iga_mf_cg_3d (iga_wq_mf/fsrc/mf_iga_solver.f90:642-958,
// alpha = rsold / (p . Ap), with diagonal-type preconditioners). Stopping rule here:
// ||r||_2 <= rtol ||b||_2 over the free dofs. Reductions use a fixed grid and a fixed-order final
// sum, so every solve is bit-reproducible.
#include <cmath>
#include <vector>

#include "common.hpp"

extern "C" {
int32_t yeti_mf_apply_f64(const yeti_mf *m, const double *x, double *y, void *stream);
int32_t yeti_mf_apply_f32(const yeti_mf *m, const float *x, float *y, void *stream);
int32_t yeti_mf_diag_blocks_f64(const yeti_mf *m, double *blocks, void *stream);
int32_t yeti_mf_nb_cp(const yeti_mf *m);
}

namespace yeti {

constexpr int NRED = 1024;

__device__ inline double block_sum256(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    __syncthreads();
    return s;
}

__global__ void sum_partials_kernel(const double *part, int n, double *out) {
    __shared__ double sh[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    double t = block_sum256(v, sh);
    if (threadIdx.x == 0) *out = t;
}

// invert the 3x3 block of every CP after pinning fixed dofs (row/col -> identity), then zero the
// fixed entries: z = Minv r never touches fixed dofs
__global__ void block_inverse_kernel(int32_t ncp, const double *__restrict__ blk, const uint8_t *__restrict__ mask,
                                     double *__restrict__ inv64, float *__restrict__ inv32) {
    int cp = blockIdx.x * blockDim.x + threadIdx.x;
    if (cp >= ncp) return;
    double a[3][3];
    bool f[3];
    for (int k = 0; k < 3; ++k) f[k] = mask[3 * (int64_t)cp + k] != 0;
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
            double v = blk[9 * (int64_t)cp + 3 * k + l];
            if (!f[k] || !f[l]) v = (k == l) ? 1.0 : 0.0;
            a[k][l] = v;
        }
    double m00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    double m01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    double m02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    double det = a[0][0] * m00 + a[0][1] * m01 + a[0][2] * m02;
    double id = 1.0 / det;
    double iv[9];
    iv[0] = m00 * id;
    iv[1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
    iv[2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
    iv[3] = m01 * id;
    iv[4] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
    iv[5] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
    iv[6] = m02 * id;
    iv[7] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
    iv[8] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
            double v = (f[k] && f[l]) ? iv[3 * k + l] : 0.0;
            if (inv64) inv64[9 * (int64_t)cp + 3 * k + l] = v;
            if (inv32) inv32[9 * (int64_t)cp + 3 * k + l] = (float)v;
        }
}

// r = mask b (FP64 in, T out); partial b.b
template <typename T>
__global__ void __launch_bounds__(256) mask_copy_kernel(int64_t n, const double *b, const uint8_t *mask, T *r,
                                                        double *part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = mask[i] ? b[i] : 0.0;
        r[i] = (T)v;
        acc += v * v;
    }
    double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// z = Minv r per CP; partial r.z
template <typename T>
__global__ void __launch_bounds__(256) precond_kernel(int32_t ncp, const T *__restrict__ inv, const T *__restrict__ r,
                                                      T *__restrict__ z, double *part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t cp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; cp < ncp; cp += (int64_t)gridDim.x * blockDim.x) {
        const T *M = inv + 9 * cp;
        T r0 = r[3 * cp], r1 = r[3 * cp + 1], r2 = r[3 * cp + 2];
        T z0 = M[0] * r0 + M[1] * r1 + M[2] * r2;
        T z1 = M[3] * r0 + M[4] * r1 + M[5] * r2;
        T z2 = M[6] * r0 + M[7] * r1 + M[8] * r2;
        z[3 * cp] = z0;
        z[3 * cp + 1] = z1;
        z[3 * cp + 2] = z2;
        acc += (double)r0 * (double)z0 + (double)r1 * (double)z1 + (double)r2 * (double)z2;
    }
    double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// q = mask q; partial p.q
template <typename T>
__global__ void __launch_bounds__(256) mask_dot_kernel(int64_t n, const uint8_t *mask, const T *__restrict__ p,
                                                       T *__restrict__ q, double *part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T qi = mask[i] ? q[i] : T(0);
        q[i] = qi;
        acc += (double)p[i] * (double)qi;
    }
    double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// sc: [0] rz, [1] pq, [2] rr, [4] rz_new. x += a p, r -= a q; partial r.r
template <typename T>
__global__ void __launch_bounds__(256) xr_kernel(int64_t n, const double *sc, T *__restrict__ x, T *__restrict__ r,
                                                 const T *__restrict__ p, const T *__restrict__ q, double *part) {
    __shared__ double sh[4];
    const double alpha = sc[0] / sc[1];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] = (T)((double)x[i] + alpha * (double)p[i]);
        T ri = (T)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        acc += (double)ri * (double)ri;
    }
    double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename T>
__global__ void __launch_bounds__(256) p_kernel(int64_t n, const double *sc, T *__restrict__ p, const T *__restrict__ z,
                                                int first) {
    // first direction: p = z without reading p (fresh buffers may hold NaN bit patterns)
    const double beta = first ? 0.0 : sc[4] / sc[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = first ? z[i] : (T)((double)z[i] + beta * (double)p[i]);
}

__global__ void roll_kernel(double *sc) { sc[0] = sc[4]; }

template <typename T>
__global__ void zero_kernel(int64_t n, T *p) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = T(0);
}

// r = mask (b - y), FP64; partial r.r
__global__ void __launch_bounds__(256) residual_kernel(int64_t n, const double *b, const double *y, const uint8_t *mask,
                                                       double *r, double *part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = mask[i] ? b[i] - y[i] : 0.0;
        r[i] = v;
        acc += v * v;
    }
    double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void x_update_kernel(int64_t n, double *x, const float *d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] += (double)d[i];
}

template <typename T>
static int apply_op(const yeti_mf *m, const T *x, T *y, hipStream_t st) {
    if constexpr (sizeof(T) == 8) return yeti_mf_apply_f64(m, (const double *)x, (double *)y, st);
    else return yeti_mf_apply_f32(m, (const float *)x, (float *)y, st);
}

// PCG on the masked system from x = 0. r must hold mask*b on entry; returns iterations.
template <typename T>
static int pcg(const yeti_mf *m, const T *inv, const uint8_t *mask, T *x, T *r, double rtol_abs2, int maxit,
               int check_every, int *iters, double *rr_out, hipStream_t st, DevBuf<T> &z, DevBuf<T> &p,
               DevBuf<T> &q, DevBuf<double> &part, DevBuf<double> &sc) {
    const int32_t ncp = yeti_mf_nb_cp(m);
    const int64_t n = 3 * (int64_t)ncp;
    const unsigned nb = (unsigned)std::min<int64_t>(NRED, (n + 255) / 256);
    const unsigned nbc = (unsigned)std::min<int64_t>(NRED, (ncp + 255) / 256);
    hipLaunchKernelGGL(zero_kernel<T>, dim3(nb), dim3(256), 0, st, n, x);
    hipLaunchKernelGGL(precond_kernel<T>, dim3(nbc), dim3(256), 0, st, ncp, inv, r, z.p, part.p);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part.p, (int)nbc, sc.p + 0);
    hipLaunchKernelGGL(p_kernel<T>, dim3(nb), dim3(256), 0, st, n, sc.p, p.p, z.p, 1);
    YETI_HIP(hipGetLastError());
    int it = 0;
    double rr = 0.0;
    while (it < maxit) {
        int rc = apply_op<T>(m, p.p, q.p, st);
        if (rc) return rc;
        hipLaunchKernelGGL(mask_dot_kernel<T>, dim3(nb), dim3(256), 0, st, n, mask, p.p, q.p, part.p);
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part.p, (int)nb, sc.p + 1);
        hipLaunchKernelGGL(xr_kernel<T>, dim3(nb), dim3(256), 0, st, n, sc.p, x, r, p.p, q.p, part.p);
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part.p, (int)nb, sc.p + 2);
        hipLaunchKernelGGL(precond_kernel<T>, dim3(nbc), dim3(256), 0, st, ncp, inv, r, z.p, part.p);
        hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part.p, (int)nbc, sc.p + 4);
        hipLaunchKernelGGL(p_kernel<T>, dim3(nb), dim3(256), 0, st, n, sc.p, p.p, z.p, 0);
        hipLaunchKernelGGL(roll_kernel, dim3(1), dim3(1), 0, st, sc.p);
        YETI_HIP(hipGetLastError());
        ++it;
        if (it % check_every == 0 || it == maxit) {
            YETI_HIP(hipMemcpyAsync(&rr, sc.p + 2, sizeof(double), hipMemcpyDeviceToHost, st));
            YETI_HIP(hipStreamSynchronize(st));
            if (!(rr > rtol_abs2)) break;
        }
    }
    *iters = it;
    *rr_out = rr;
    return 0;
}

struct MfSolver {
    DevBuf<double> blk, inv64;
    DevBuf<float> inv32;
};

static int prepare_precond(const yeti_mf *m, const uint8_t *mask, bool f64, bool f32, MfSolver &s, hipStream_t st) {
    const int32_t ncp = yeti_mf_nb_cp(m);
    YETI_HIP(s.blk.alloc(9 * (size_t)ncp));
    int rc = yeti_mf_diag_blocks_f64(m, s.blk.p, st);
    if (rc) return rc;
    if (f64) YETI_HIP(s.inv64.alloc(9 * (size_t)ncp));
    if (f32) YETI_HIP(s.inv32.alloc(9 * (size_t)ncp));
    hipLaunchKernelGGL(block_inverse_kernel, dim3((ncp + 255) / 256), dim3(256), 0, st, ncp, s.blk.p, mask,
                       f64 ? s.inv64.p : nullptr, f32 ? s.inv32.p : nullptr);
    YETI_HIP(hipGetLastError());
    return 0;
}

}  // namespace yeti

using namespace yeti;

extern "C" {

int32_t yeti_mf_pcg_f64(const yeti_mf *m, const double *b, const uint8_t *free_mask, double *x, double rtol,
                        int32_t maxit, int32_t check_every, int32_t *iters, double *relres, void *stream) {
    if (!m || !b || !free_mask || !x || maxit < 0) return set_error(YETI_ERR_INVALID_ARGUMENT, "mf pcg: bad argument");
    if (check_every <= 0) check_every = 10;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = 3 * (int64_t)yeti_mf_nb_cp(m);
    MfSolver S;
    int rc = prepare_precond(m, free_mask, true, false, S, st);
    if (rc) return rc;
    DevBuf<double> r, z, p, q, part, sc;
    YETI_HIP(r.alloc(n));
    YETI_HIP(z.alloc(n));
    YETI_HIP(p.alloc(n));
    YETI_HIP(q.alloc(n));
    YETI_HIP(part.alloc(NRED));
    YETI_HIP(sc.alloc(8));
    const unsigned nb = (unsigned)std::min<int64_t>(NRED, (n + 255) / 256);
    hipLaunchKernelGGL(mask_copy_kernel<double>, dim3(nb), dim3(256), 0, st, n, b, free_mask, r.p, part.p);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part.p, (int)nb, sc.p + 3);
    double bb = 0.0;
    YETI_HIP(hipMemcpyAsync(&bb, sc.p + 3, sizeof(double), hipMemcpyDeviceToHost, st));
    YETI_HIP(hipStreamSynchronize(st));
    int it = 0;
    double rr = 0.0;
    if (bb > 0.0) {
        rc = pcg<double>(m, S.inv64.p, free_mask, x, r.p, rtol * rtol * bb, maxit, check_every, &it, &rr, st, z, p, q,
                         part, sc);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(zero_kernel<double>, dim3(nb), dim3(256), 0, st, n, x);
    }
    double rel = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    if (iters) *iters = it;
    if (relres) *relres = rel;
    if (!(rel <= rtol)) return set_error(YETI_ERR_NOT_CONVERGED, "mf pcg: %d iterations, relative residual %.3e", it, rel);
    return 0;
}

int32_t yeti_mf_solve_ir(const yeti_mf *m, const double *b, const uint8_t *free_mask, double *x, double rtol,
                         double inner_rtol, int32_t max_outer, int32_t max_inner, int32_t check_every,
                         int32_t *outer_iters, int32_t *inner_iters, double *relres, void *stream) {
    if (!m || !b || !free_mask || !x || max_outer < 0 || max_inner < 0)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "mf ir: bad argument");
    if (check_every <= 0) check_every = 10;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = 3 * (int64_t)yeti_mf_nb_cp(m);
    MfSolver S;
    int rc = prepare_precond(m, free_mask, false, true, S, st);
    if (rc) return rc;
    DevBuf<double> r64, y64, part, sc;
    DevBuf<float> r32, d32, z, p, q;
    YETI_HIP(r64.alloc(n));
    YETI_HIP(y64.alloc(n));
    YETI_HIP(part.alloc(NRED));
    YETI_HIP(sc.alloc(8));
    YETI_HIP(r32.alloc(n));
    YETI_HIP(d32.alloc(n));
    YETI_HIP(z.alloc(n));
    YETI_HIP(p.alloc(n));
    YETI_HIP(q.alloc(n));
    const unsigned nb = (unsigned)std::min<int64_t>(NRED, (n + 255) / 256);
    hipLaunchKernelGGL(zero_kernel<double>, dim3(nb), dim3(256), 0, st, n, x);
    hipLaunchKernelGGL(mask_copy_kernel<double>, dim3(nb), dim3(256), 0, st, n, b, free_mask, r64.p, part.p);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part.p, (int)nb, sc.p + 3);
    double bb = 0.0;
    YETI_HIP(hipMemcpyAsync(&bb, sc.p + 3, sizeof(double), hipMemcpyDeviceToHost, st));
    YETI_HIP(hipStreamSynchronize(st));
    int outer = 0, inner_total = 0;
    double rel = 0.0;
    if (bb > 0.0) {
        double rr = bb;
        rel = 1.0;
        while (outer < max_outer) {
            // FP32 correction: A32 d = r (r in FP32), inner tolerance relative to ||r||
            hipLaunchKernelGGL(mask_copy_kernel<float>, dim3(nb), dim3(256), 0, st, n, r64.p, free_mask, r32.p, part.p);
            int it = 0;
            double rr_in = 0.0;
            rc = pcg<float>(m, S.inv32.p, free_mask, d32.p, r32.p, inner_rtol * inner_rtol * rr, max_inner, check_every,
                            &it, &rr_in, st, z, p, q, part, sc);
            if (rc) return rc;
            inner_total += it;
            hipLaunchKernelGGL(x_update_kernel, dim3(nb), dim3(256), 0, st, n, x, d32.p);
            // FP64 refinement residual r = mask (b - K x)
            rc = yeti_mf_apply_f64(m, x, y64.p, st);
            if (rc) return rc;
            hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(256), 0, st, n, b, y64.p, free_mask, r64.p, part.p);
            hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, part.p, (int)nb, sc.p + 3);
            YETI_HIP(hipGetLastError());
            YETI_HIP(hipMemcpyAsync(&rr, sc.p + 3, sizeof(double), hipMemcpyDeviceToHost, st));
            YETI_HIP(hipStreamSynchronize(st));
            ++outer;
            rel = std::sqrt(rr / bb);
            if (!(rel > rtol)) break;
        }
    }
    if (outer_iters) *outer_iters = outer;
    if (inner_iters) *inner_iters = inner_total;
    if (relres) *relres = rel;
    if (!(rel <= rtol))
        return set_error(YETI_ERR_NOT_CONVERGED, "mf ir: %d outer / %d inner iterations, relative residual %.3e", outer,
                         inner_total, rel);
    return 0;
}

}  // extern "C"
