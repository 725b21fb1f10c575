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
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

#include "batch.hpp"

extern "C" {
int32_t yeti_mf_apply_f64(const yeti_mf *m, const double *x, double *y, void *stream);
int32_t yeti_mf_apply_f32(const yeti_mf *m, const float *x, float *y, void *stream);
int32_t yeti_mf_diag_blocks_f64(const yeti_mf *m, double *blocks, void *stream);
int32_t yeti_mf_nb_cp(const yeti_mf *m);
}

namespace yeti {

constexpr int NRED = 1024;

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

// partial[b] = max |v[i]| over the block's masked-in entries; out = max of the partials
template <typename T>
__global__ void __launch_bounds__(256) absmax_partial_kernel(int64_t n, const T *__restrict__ v, const uint8_t *mask,
                                                             double *__restrict__ partial) {
    __shared__ double sh[4];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (mask[i]) m = fmax(m, fabs((double)v[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ void max_partials_kernel(const double *part, int n, double *out) {
    if (threadIdx.x != 0) return;
    double m = 0.0;
    for (int i = 0; i < n; ++i) m = fmax(m, part[i]);
    *out = m;
}

template <typename T>
static int absmax_host(int64_t n, const T *v, const uint8_t *mask, double *part, double *slot, double *out,
                       hipStream_t st) {
    const unsigned nb = (unsigned)std::min<int64_t>(NRED, (n + 255) / 256);
    hipLaunchKernelGGL(absmax_partial_kernel<T>, dim3(nb), dim3(256), 0, st, n, v, mask, part);
    hipLaunchKernelGGL(max_partials_kernel, dim3(1), dim3(64), 0, st, (const double *)part, (int)nb, slot);
    YETI_HIP(hipGetLastError());
    YETI_HIP(hipMemcpyAsync(out, slot, sizeof(double), hipMemcpyDeviceToHost, st));
    YETI_HIP(hipStreamSynchronize(st));
    return 0;
}

template <typename T>
static int apply_op(const yeti_mf *m, const T *x, T *y, hipStream_t st) {
    if constexpr (sizeof(T) == 8) return yeti_mf_apply_f64(m, (const double *)x, (double *)y, st);
    else return yeti_mf_apply_f32(m, (const float *)x, (float *)y, st);
}

// PCG on the masked system from x = 0. r must hold mask*b on entry; returns iterations. Stops
// when r.r <= rtol_abs2, or, with inf_tol > 0, when max|r| <= inf_tol (then *rr_out is max|r|).
template <typename T>
static int pcg(const yeti_mf *m, const T *inv, const uint8_t *mask, T *x, T *r, double rtol_abs2, int maxit,
               int check_every, int *iters, double *rr_out, hipStream_t st, DevBuf<T> &z, DevBuf<T> &p,
               DevBuf<T> &q, DevBuf<double> &part, DevBuf<double> &sc, double inf_tol = 0.0) {
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
            if (inf_tol > 0.0) {
                int rc2 = absmax_host<T>(n, r, mask, part.p, sc.p + 6, &rr, st);
                if (rc2) return rc2;
                if (!(rr > inf_tol)) break;
            } else {
                YETI_HIP(hipMemcpyAsync(&rr, sc.p + 2, sizeof(double), hipMemcpyDeviceToHost, st));
                YETI_HIP(hipStreamSynchronize(st));
                if (!(rr > rtol_abs2)) break;
            }
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
                        int32_t maxit, int32_t check_every, int32_t norm, int32_t *iters, double *relres,
                        void *stream) {
    if (!m || !b || !free_mask || !x || maxit < 0 || (norm != 0 && norm != 2))
        return set_error(YETI_ERR_INVALID_ARGUMENT, "mf pcg: bad argument");
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
    double rr = 0.0, bmax = 0.0;
    if (bb > 0.0 && norm == 0) {   // the reference's max|b| (r = mask b here)
        rc = absmax_host<double>(n, r.p, free_mask, part.p, sc.p + 6, &bmax, st);
        if (rc) return rc;
    }
    if (bb > 0.0) {
        rc = pcg<double>(m, S.inv64.p, free_mask, x, r.p, rtol * rtol * bb, maxit, check_every, &it, &rr, st, z, p, q,
                         part, sc, norm == 0 ? rtol * bmax : 0.0);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(zero_kernel<double>, dim3(nb), dim3(256), 0, st, n, x);
    }
    double rel = bb > 0.0 ? (norm == 0 ? rr / bmax : std::sqrt(rr / bb)) : 0.0;
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

// ==========================================================================================
// Steppable, device-controlled solver for serving (yeti_ir_*).
//
// One session solves K[free][:, free] x = b[free] for one handle with preconditioned CG whose
// vectors and operator run in FP32 (or FP64) while the solution and the true residual are kept
// in FP64 and refreshed by residual replacement ("reliable updates"): the FP32 iteration
// accumulates its update y; when its recurred residual has fallen by `delta` since the last
// replacement (or meets the tolerance), then at the end of the current step x += y,
// r = mask (b - K x) with the FP64 apply, and the FP32 residual restarts from r while the search
// direction is kept. The Krylov space is not
// restarted, so an ill-conditioned system converges in about the iterations of FP64 CG, and the
// stopping test ||b - K x||_2 <= rtol ||b||_2 is always made on the FP64 residual.
// Control state lives in device memory: every kernel of a step is predicated on device flags,
// so a step (`chunk` iterations) is a fixed launch sequence that is captured once in a hipGraph
// and replayed until the session reports done; the host reads a small status block. Dot
// products are finished inside their producing launch (the last block to arrive adds the
// per-block partials in block order), so an iteration is 6 launches plus the preconditioner's. The
// iterates depend only on the session's own launch sequence: a request gives bit-identical
// results alone or in a mix. All device memory is taken at creation (optionally in an arena).
namespace yeti {

int mf_apply_pred(const yeti_mf *m, const double *x, double *y, hipStream_t st, const int32_t *pred);
int mf_apply_pred(const yeti_mf *m, const float *x, float *y, hipStream_t st, const int32_t *pred);
int fdm_create(const yeti_mf *m, const uint8_t *mask_host, int precision, hipStream_t st, yeti_fdm **out);
void fdm_destroy(yeti_fdm *f);
size_t fdm_bytes_bound(const yeti_mf *m, int precision);
int fdm_apply(const yeti_fdm *f, const double *r, const uint8_t *mask, double *z, double *part, int nblocks,
              const RzTail &tail, hipStream_t st, const int32_t *pred);
int fdm_apply(const yeti_fdm *f, const float *r, const uint8_t *mask, float *z, double *part, int nblocks,
              const RzTail &tail, hipStream_t st, const int32_t *pred);

// iterations after which a residual replacement is forced (so a solve at the FP64 rounding
// floor is recognised within a few such periods)
constexpr int32_t IR_FORCE_EVERY = 256;
// after a replacement that did not halve the true residual, the next one is forced this soon:
// the solve is at its FP64 floor, and stopping there should not cost IR_FORCE_EVERY iterations
// per stagnant replacement
constexpr int32_t IR_FORCE_STAGNANT = 32;

__global__ void pred_sum_partials_kernel(const double *part, int n, double *out, const int32_t *pred) {
    if (pred && *pred == 0) return;
    __shared__ double sh[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    double t = block_sum256(v, sh);
    if (threadIdx.x == 0) *out = t;
}

// 3x3 block-Jacobi z = M^-1 r; r.z reduced in the same launch and handed to rz_tail
template <typename T>
__global__ void __launch_bounds__(256) pred_precond_kernel(int32_t ncp, const T *__restrict__ inv, const T *__restrict__ r,
                                                           T *__restrict__ z, double *part, RzTail tail,
                                                           const int32_t *pred) {
    if (*pred == 0) return;
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
    double tot;
    if (last_block_total(block_sum256(acc, sh), part, tail.flags + F_COUNTER, sh, tot) && threadIdx.x == 0)
        rz_tail(tail, tot);
}

// q = mask q; sc[S_PQ] = p.q
template <typename T>
__global__ void __launch_bounds__(256) pred_mask_dot_kernel(int64_t n, const uint8_t *mask, const T *__restrict__ p,
                                                            T *__restrict__ q, double *part, double *sc,
                                                            int32_t *flags) {
    if (flags[F_ACTIVE] == 0) return;
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T qi = mask[i] ? q[i] : T(0);
        q[i] = qi;
        acc += (double)p[i] * (double)qi;
    }
    double tot;
    if (last_block_total(block_sum256(acc, sh), part, flags + F_COUNTER, sh, tot) && threadIdx.x == 0) sc[S_PQ] = tot;
}

// y += alpha p, r -= alpha q; sc[S_RR] = r.r, then the replacement test: a residual replacement
// is due when r.r has fallen by delta^2 since the last one, meets the tolerance, or after
// IR_FORCE_EVERY iterations (it is carried out at the end of the step).
template <typename T>
__global__ void __launch_bounds__(256) pred_xr_kernel(int64_t n, double *sc, T *__restrict__ y, T *__restrict__ r,
                                                      const T *__restrict__ p, const T *__restrict__ q, double *part,
                                                      int32_t *flags, double delta2) {
    if (flags[F_ACTIVE] == 0) return;
    __shared__ double sh[4];
    const double alpha = sc[S_RZ] / sc[S_PQ];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        y[i] = (T)((double)y[i] + alpha * (double)p[i]);
        T ri = (T)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        acc += (double)ri * (double)ri;
    }
    double rr;
    if (last_block_total(block_sum256(acc, sh), part, flags + F_COUNTER, sh, rr) && threadIdx.x == 0) {
        sc[S_RR] = rr;
        if (rr > sc[S_RRMAX]) sc[S_RRMAX] = rr;
        const int32_t since = flags[F_ITERS] + 1 - (int32_t)sc[S_LASTUPD];
        if (!(rr > delta2 * sc[S_RRMAX]) || !(rr > sc[S_TOL2])) {
            if (flags[F_UPDATE] == 0) flags[F_UPDATE] = 1;
        } else if (since >= (flags[F_STAGNANT] && flags[F_DEPARTED] ? IR_FORCE_STAGNANT : IR_FORCE_EVERY)) {
            if (flags[F_UPDATE] == 0) flags[F_UPDATE] = 2;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) pred_p_kernel(int64_t n, const double *sc, T *__restrict__ p, const T *__restrict__ z,
                                                     const int32_t *pred) {
    if (*pred == 0) return;
    const double beta = sc[S_BETA];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (T)((double)z[i] + beta * (double)p[i]);
}

// x += y, y = 0 (FP64 accumulation of the FP32/FP64 update), predicated on F_UPDATE
template <typename T>
__global__ void __launch_bounds__(256) ir_accumulate_kernel(int64_t n, double *x, T *y, const int32_t *flags) {
    if (flags[F_UPDATE] == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += (double)y[i];
        y[i] = T(0);
    }
}

// The end of a residual replacement whose true residual is rrt = |mask (b - K x)|^2 (one thread):
// stop at the tolerance; count replacements that do not improve the true residual; keep track of
// the best iterate (F_BEST) and of whether the iterated residual had departed from the true one.
// A solve that stagnates (max_stagnant replacements without a 4x gain) or diverges (true
// residual 1e4 x above its best) stops: at its FP64 rounding floor (F_FLOOR) when its best
// relative residual is at most IR_FLOOR_MAX, else as failed (F_FAILED), which is what an
// iteration that breaks down in its working precision gives. The best iterate is restored
// (F_RESTORE) when the solve stops above it.
constexpr double IR_FLOOR_MAX = 1e-6;

__device__ __forceinline__ void ir_replacement_end(double rrt, double *sc, int32_t *flags, double rtol2,
                                                   int32_t max_updates, int32_t max_stagnant) {
    const bool forced = flags[F_UPDATE] == 2;
    // the iterated residual has left the true one behind (by more than 4x in norm): the mark of
    // a rounding floor, as opposed to a solve that is merely converging slowly
    const bool departed = rrt > 16.0 * sc[S_RR];
    sc[S_RRT] = rrt;
    flags[F_UPDATE] = 0;
    flags[F_UPDATES] += 1;
    sc[S_LASTUPD] = (double)(flags[F_ITERS] + 1);
    sc[S_RR] = rrt;
    sc[S_RRMAX] = rrt;
    if (!(rrt == rrt)) {   // NaN: the iteration broke down; return the best iterate, failed
        flags[F_FAILED] = 1;
        flags[F_DONE] = 1;
        flags[F_ACTIVE] = 0;
        flags[F_RESTORE] = 1;
        sc[S_RRT] = sc[S_RRT_PREV];
        return;
    }
    if (rrt <= rtol2 * sc[S_BB]) {
        flags[F_DONE] = 1;
        flags[F_ACTIVE] = 0;
        return;
    }
    flags[F_DEPARTED] = departed ? 1 : 0;
    if (!(rrt < 0.25 * sc[S_RRT_PREV])) flags[F_STAGNANT] += 1;
    else flags[F_STAGNANT] = 0;
    // far above the best residual so far: the iterate is drifting away
    const bool diverging = rrt > 1e4 * sc[S_RRT_PREV];
    // S_RRT_PREV is the lowest true residual so far: save this iterate when it is the new best
    if (rrt < sc[S_RRT_PREV]) {
        sc[S_RRT_PREV] = rrt;
        flags[F_BEST] = 1;
    }
    // a forced replacement whose true residual has departed from the iterated one: the old
    // direction carries no information about the replaced residual, so the next direction starts
    // afresh (beta = 0 through an infinite previous r.z)
    if (forced && departed) sc[S_RZ] = __builtin_inf();
    bool stop = false;
    if (flags[F_STAGNANT] >= max_stagnant || diverging) {
        if (sc[S_RRT_PREV] <= IR_FLOOR_MAX * IR_FLOOR_MAX * sc[S_BB]) flags[F_FLOOR] = 1;
        else flags[F_FAILED] = 1;
        stop = true;
    } else if (flags[F_UPDATES] >= max_updates) {
        flags[F_FAILED] = 1;
        stop = true;
    }
    if (stop) {
        flags[F_DONE] = 1;
        flags[F_ACTIVE] = 0;
        if (sc[S_RRT_PREV] < rrt) {   // return the best iterate, and report its residual
            flags[F_RESTORE] = 1;
            sc[S_RRT] = sc[S_RRT_PREV];
        }
    }
}

// Residual replacement, predicated on F_UPDATE: r64 = mask (b - K x); r = (T) r64; then, in the
// last block, the end of the replacement: stop at the tolerance, and count replacements that do
// not improve the true residual (stagnation at the FP64 rounding floor of ill-conditioned systems).
template <typename T>
__global__ void __launch_bounds__(256) ir_true_residual_kernel(int64_t n, const double *b, const double *Kx,
                                                               const uint8_t *mask, T *r, double *part, double *sc,
                                                               int32_t *flags, double rtol2, int32_t max_updates,
                                                               int32_t max_stagnant) {
    if (flags[F_UPDATE] == 0) return;
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = mask[i] ? b[i] - Kx[i] : 0.0;
        r[i] = (T)v;
        acc += v * v;
    }
    double rrt;
    if (!last_block_total(block_sum256(acc, sh), part, flags + F_COUNTER, sh, rrt) || threadIdx.x != 0) return;
    ir_replacement_end(rrt, sc, flags, rtol2, max_updates, max_stagnant);
}

// After a replacement: keep the iterate as the best one (F_BEST), or put the best one back at
// the end of a solve that stopped above its best residual (F_RESTORE).
__global__ void __launch_bounds__(256) ir_best_kernel(int64_t n, double *x, double *xbest, const int32_t *flags) {
    const int32_t best = flags[F_BEST], restore = flags[F_RESTORE];
    if (!best && !restore) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (best) xbest[i] = x[i];
        else x[i] = xbest[i];
    }
}

__global__ void ir_best_clear_kernel(int32_t *flags) {
    flags[F_BEST] = 0;
    flags[F_RESTORE] = 0;
}

template <typename T>
__global__ void __launch_bounds__(256) ir_begin_copy_kernel(int64_t n, const double *b, const uint8_t *mask, double *x,
                                                            double *xbest, T *r, T *y, double *part) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = mask[i] ? b[i] : 0.0;
        r[i] = (T)v;
        x[i] = 0.0;
        xbest[i] = 0.0;   // the best iterate so far is the start, whose residual is b
        y[i] = T(0);
        acc += v * v;
    }
    double t = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void ir_begin_flags_kernel(double *sc, int32_t *flags, double rtol2) {
    const double bb = sc[S_BB];
    sc[S_RR] = bb;
    sc[S_RRT] = bb;
    sc[S_RRMAX] = bb;
    sc[S_RRT_PREV] = bb;
    sc[S_TOL2] = rtol2 * bb;
    sc[S_LASTUPD] = 0.0;
    sc[S_BETA] = 0.0;
    for (int i = 0; i < F_ALLOC; ++i) flags[i] = 0;   // includes the reduction counter
    if (bb > 0.0) flags[F_ACTIVE] = 1;
    else flags[F_DONE] = 1;
}

// p = z (first direction), rz, predicated on F_ACTIVE
template <typename T>
__global__ void __launch_bounds__(256) ir_first_p_kernel(int64_t n, const T *z, T *p, const int32_t *flags) {
    if (flags[F_ACTIVE] == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i];
}

}  // namespace yeti

struct yeti_ir {
    const yeti_mf *m = nullptr;
    const uint8_t *mask = nullptr;
    int64_t n = 0;
    int32_t ncp = 0;
    int inner64 = 0;          // 1: FP64 vectors / apply (plain FP64 PCG), 0: FP32
    double rtol = 0, delta = 0;
    int32_t chunk = 8, max_iters = 0, max_updates = 0;
    unsigned nb = 1, nbc = 1;
    yeti::DevBuf<double> blk, inv64, y64, part, sc, xbest;
    yeti::DevBuf<float> inv32;
    yeti::DevBuf<double> vr64, vy64, vz64, vp64, vq64;   // iteration vectors, FP64
    yeti::DevBuf<float> vr32, vy32, vz32, vp32, vq32;    // iteration vectors, FP32
    yeti::DevBuf<int32_t> flags;
    yeti_fdm *fdm = nullptr;  // fast diagonalisation (else 3x3 block-Jacobi)
    // FP64 operator: the matrix-free apply, or an assembled device CSR (yeti_ir_set_csr)
    const yeti_csr *csr = nullptr;
    const double *csr_vals = nullptr;
    const double *b = nullptr;
    double *x = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    ~yeti_ir() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (fdm) yeti::fdm_destroy(fdm);
    }
};

namespace yeti {

constexpr int32_t IR_MAX_STAGNANT = 4;

// The session's products: FP64 ones through the assembled CSR when one is set, FP32 ones always
// matrix-free.
static int ir_apply(yeti_ir *s, const double *x, double *y, hipStream_t st, const int32_t *pred) {
    if (s->csr) return csr_block_spmv(s->csr, s->csr_vals, x, nullptr, y, nullptr, 0, nullptr, st, pred);
    return mf_apply_pred(s->m, x, y, st, pred);
}
static int ir_apply(yeti_ir *s, const float *x, float *y, hipStream_t st, const int32_t *pred) {
    return mf_apply_pred(s->m, x, y, st, pred);
}

// z = M^-1 r; r.z goes to rz_tail in the same launch sequence (mode 0: first direction, 1: end
// of an iteration), predicated
template <typename T>
static int ir_precond(yeti_ir *s, const T *inv, const T *r, T *z, int mode, hipStream_t st, const int32_t *pred) {
    RzTail tail;
    tail.sc = s->sc.p;
    tail.flags = s->flags.p;
    tail.mode = mode;
    tail.max_iters = s->max_iters;
    if (s->fdm) return fdm_apply(s->fdm, r, s->mask, z, s->part.p, (int)s->nbc, tail, st, pred);
    hipLaunchKernelGGL(pred_precond_kernel<T>, dim3(s->nbc), dim3(256), 0, st, s->ncp, inv, r, z, s->part.p, tail, pred);
    YETI_HIP(hipGetLastError());
    return 0;
}

// The launch sequence of one step (captured into the session's graph): `chunk` PCG iterations,
// the residual replacement when one fell due during them, and the new direction. Per iteration:
// the apply (2 launches per patch), p.q, the x/r update with r.r and the replacement test, the
// preconditioner with r.z, beta and the iteration count, and p = z + beta p. The replacement
// (x += y, the FP64 apply, the true residual and the stopping tests) is placed once per step,
// before the last preconditioner, so it costs 4 launches per step instead of per iteration.
template <typename T>
static int ir_step_launches(yeti_ir *s, T *r, T *y, T *z, T *p, T *q, const T *inv, hipStream_t st) {
    const int32_t *act = s->flags.p + F_ACTIVE;
    const int32_t *upd = s->flags.p + F_UPDATE;
    const double delta2 = s->delta * s->delta;
    for (int k = 0; k < s->chunk; ++k) {
        int rc = ir_apply(s, p, q, st, act);
        if (rc) return rc;
        hipLaunchKernelGGL(pred_mask_dot_kernel<T>, dim3(s->nb), dim3(256), 0, st, s->n, s->mask, p, q, s->part.p,
                           s->sc.p, s->flags.p);
        hipLaunchKernelGGL(pred_xr_kernel<T>, dim3(s->nb), dim3(256), 0, st, s->n, s->sc.p, y, r, p, q, s->part.p,
                           s->flags.p, delta2);
        if (k == s->chunk - 1) {
            // residual replacement (predicated on F_UPDATE)
            hipLaunchKernelGGL(ir_accumulate_kernel<T>, dim3(s->nb), dim3(256), 0, st, s->n, s->x, y, s->flags.p);
            rc = ir_apply(s, (const double *)s->x, s->y64.p, st, upd);
            if (rc) return rc;
            hipLaunchKernelGGL(ir_true_residual_kernel<T>, dim3(s->nb), dim3(256), 0, st, s->n, s->b, s->y64.p, s->mask,
                               r, s->part.p, s->sc.p, s->flags.p, s->rtol * s->rtol, s->max_updates, IR_MAX_STAGNANT);
            hipLaunchKernelGGL(ir_best_kernel, dim3(s->nb), dim3(256), 0, st, s->n, s->x, s->xbest.p,
                               (const int32_t *)s->flags.p);
            hipLaunchKernelGGL(ir_best_clear_kernel, dim3(1), dim3(1), 0, st, s->flags.p);
        }
        // preconditioner and direction
        rc = ir_precond<T>(s, inv, r, z, 1, st, act);
        if (rc) return rc;
        hipLaunchKernelGGL(pred_p_kernel<T>, dim3(s->nb), dim3(256), 0, st, s->n, s->sc.p, p, z, act);
        YETI_HIP(hipGetLastError());
    }
    return 0;
}

static int ir_step_any(yeti_ir *s, hipStream_t st) {
    if (s->inner64)
        return ir_step_launches<double>(s, s->vr64.p, s->vy64.p, s->vz64.p, s->vp64.p, s->vq64.p, s->inv64.p, st);
    return ir_step_launches<float>(s, s->vr32.p, s->vy32.p, s->vz32.p, s->vp32.p, s->vq32.p, s->inv32.p, st);
}

template <typename T>
static int ir_begin_launches(yeti_ir *s, T *r, T *y, T *z, T *p, const T *inv, hipStream_t st) {
    const int32_t *act = s->flags.p + F_ACTIVE;
    hipLaunchKernelGGL(ir_begin_copy_kernel<T>, dim3(s->nb), dim3(256), 0, st, s->n, s->b, s->mask, s->x, s->xbest.p, r,
                       y, s->part.p);
    hipLaunchKernelGGL(pred_sum_partials_kernel, dim3(1), dim3(256), 0, st, s->part.p, (int)s->nb, s->sc.p + S_BB,
                       (const int32_t *)nullptr);
    hipLaunchKernelGGL(ir_begin_flags_kernel, dim3(1), dim3(1), 0, st, s->sc.p, s->flags.p, s->rtol * s->rtol);
    int rc = ir_precond<T>(s, inv, r, z, 0, st, act);
    if (rc) return rc;
    hipLaunchKernelGGL(ir_first_p_kernel<T>, dim3(s->nb), dim3(256), 0, st, s->n, z, p, s->flags.p);
    YETI_HIP(hipGetLastError());
    return 0;
}

}  // namespace yeti

namespace yeti {

// ---- batches: many sessions advanced by the same launches --------------------------------
// Member vector chunk [a vchunk, (a + 1) vchunk) of task (req, a) in the BT_VEC list.
#define BT_VEC_LOOP(bd, PRED)                                                          \
    const int32_t cnt_ = bd.count[BT_VEC];                                             \
    const BTask *tl_ = bd.list[BT_VEC];                                                \
    for (int32_t t_ = blockIdx.x; t_ < cnt_; t_ += gridDim.x) {                         \
        const BTask tk = tl_[t_];                                                      \
        const BReq &R = bd.req[tk.req];                                                \
        if (R.flags[PRED] == 0) continue;                                              \
        const int64_t i0 = (int64_t)tk.a * R.vchunk, i1 = std::min<int64_t>(R.n, i0 + R.vchunk);

// q = mask q; S_PQ = p.q
template <typename T>
__global__ void __launch_bounds__(256) bt_mask_dot_kernel(BatchDev bd) {
    __shared__ double sh[4];
    BT_VEC_LOOP(bd, F_ACTIVE)
        const T *p = (const T *)R.vp;
        T *q = (T *)R.vq;
        double acc = 0.0;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            T qi = R.mask[i] ? q[i] : T(0);
            q[i] = qi;
            acc += (double)p[i] * (double)qi;
        }
        double tot;
        if (member_total(block_sum256(acc, sh), R.red, R.flags + F_COUNTER, R.nvchunk, tk.a, sh, tot) &&
            threadIdx.x == 0)
            R.sc[S_PQ] = tot;
    }
}

// y += alpha p, r -= alpha q; S_RR = r.r and the replacement test (as pred_xr_kernel)
template <typename T>
__global__ void __launch_bounds__(256) bt_xr_kernel(BatchDev bd) {
    __shared__ double sh[4];
    BT_VEC_LOOP(bd, F_ACTIVE)
        const double alpha = R.sc[S_RZ] / R.sc[S_PQ];
        T *y = (T *)R.vy;
        T *r = (T *)R.vr;
        const T *p = (const T *)R.vp;
        const T *q = (const T *)R.vq;
        double acc = 0.0;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            y[i] = (T)((double)y[i] + alpha * (double)p[i]);
            T ri = (T)((double)r[i] - alpha * (double)q[i]);
            r[i] = ri;
            acc += (double)ri * (double)ri;
        }
        double rr;
        if (member_total(block_sum256(acc, sh), R.red, R.flags + F_COUNTER, R.nvchunk, tk.a, sh, rr) &&
            threadIdx.x == 0) {
            double *sc = R.sc;
            int32_t *flags = R.flags;
            sc[S_RR] = rr;
            if (rr > sc[S_RRMAX]) sc[S_RRMAX] = rr;
            const int32_t since = flags[F_ITERS] + 1 - (int32_t)sc[S_LASTUPD];
            if (!(rr > R.delta2 * sc[S_RRMAX]) || !(rr > sc[S_TOL2])) {
                if (flags[F_UPDATE] == 0) flags[F_UPDATE] = 1;
            } else if (since >= (flags[F_STAGNANT] && flags[F_DEPARTED] ? IR_FORCE_STAGNANT : IR_FORCE_EVERY)) {
                if (flags[F_UPDATE] == 0) flags[F_UPDATE] = 2;
            }
        }
    }
}

// p = z + beta p
template <typename T>
__global__ void __launch_bounds__(256) bt_p_kernel(BatchDev bd) {
    BT_VEC_LOOP(bd, F_ACTIVE)
        const double beta = R.sc[S_BETA];
        T *p = (T *)R.vp;
        const T *z = (const T *)R.vz;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) p[i] = (T)((double)z[i] + beta * (double)p[i]);
    }
}

// x += y, y = 0 (members with a replacement due)
template <typename T>
__global__ void __launch_bounds__(256) bt_accumulate_kernel(BatchDev bd) {
    BT_VEC_LOOP(bd, F_UPDATE)
        T *y = (T *)R.vy;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            R.x[i] += (double)y[i];
            y[i] = T(0);
        }
    }
}

// r = mask (b - K x) from the FP64 apply, and the session's replacement logic (as
// ir_true_residual_kernel)
template <typename T>
__global__ void __launch_bounds__(256) bt_true_residual_kernel(BatchDev bd) {
    __shared__ double sh[4];
    BT_VEC_LOOP(bd, F_UPDATE)
        T *r = (T *)R.vr;
        double acc = 0.0;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            double v = R.mask[i] ? R.b[i] - R.y64[i] : 0.0;
            r[i] = (T)v;
            acc += v * v;
        }
        double rrt;
        if (!member_total(block_sum256(acc, sh), R.red, R.flags + F_COUNTER, R.nvchunk, tk.a, sh, rrt) ||
            threadIdx.x != 0)
            continue;
        ir_replacement_end(rrt, R.sc, R.flags, R.rtol2, R.max_updates, IR_MAX_STAGNANT);
    }
}

// keep the best iterate / put it back (as ir_best_kernel); the flags are cleared afterwards by
// bt_best_clear_kernel
__global__ void __launch_bounds__(256) bt_best_kernel(BatchDev bd) {
    const int32_t cnt = bd.count[BT_VEC];
    const BTask *tl = bd.list[BT_VEC];
    for (int32_t t = blockIdx.x; t < cnt; t += gridDim.x) {
        const BTask tk = tl[t];
        const BReq &R = bd.req[tk.req];
        const int32_t best = R.flags[F_BEST], restore = R.flags[F_RESTORE];
        if (!best && !restore) continue;
        const int64_t i0 = (int64_t)tk.a * R.vchunk, i1 = std::min<int64_t>(R.n, i0 + R.vchunk);
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            if (best) R.xbest[i] = R.x[i];
            else R.x[i] = R.xbest[i];
        }
    }
}

__global__ void bt_best_clear_kernel(BatchDev bd) {
    for (int32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < bd.count[BT_NLISTS]; m += gridDim.x * blockDim.x) {
        bd.req[m].flags[F_BEST] = 0;
        bd.req[m].flags[F_RESTORE] = 0;
    }
}

// statuses of the members, in member order, in the layout of yeti_ir_status
__global__ void bt_status_kernel(BatchDev bd, yeti_ir_status *out) {
    const int32_t nm = bd.count[BT_NLISTS];
    for (int32_t m = blockIdx.x; m < nm; m += gridDim.x) {
        const BReq &R = bd.req[m];
        int32_t *fl = &out[m].active;
        double *sc = &out[m].rz;
        if (threadIdx.x < F_NFLAGS) fl[threadIdx.x] = R.flags[threadIdx.x];
        if (threadIdx.x < S_NSC) sc[threadIdx.x] = R.sc[threadIdx.x];
    }
}

constexpr unsigned BT_VGRID = 2048;

template <typename T>
static int batch_step_launches(const BatchDev &bd, int prec64, int chunk, hipStream_t st) {
    for (int k = 0; k < chunk; ++k) {
        int rc = batch_mf_apply(bd, prec64, false, st);
        if (rc) return rc;
        hipLaunchKernelGGL(bt_mask_dot_kernel<T>, dim3(BT_VGRID), dim3(256), 0, st, bd);
        hipLaunchKernelGGL(bt_xr_kernel<T>, dim3(BT_VGRID), dim3(256), 0, st, bd);
        if (k == chunk - 1) {
            hipLaunchKernelGGL(bt_accumulate_kernel<T>, dim3(BT_VGRID), dim3(256), 0, st, bd);
            rc = batch_mf_apply(bd, prec64, true, st);
            if (rc) return rc;
            hipLaunchKernelGGL(bt_true_residual_kernel<T>, dim3(BT_VGRID), dim3(256), 0, st, bd);
            hipLaunchKernelGGL(bt_best_kernel, dim3(BT_VGRID), dim3(256), 0, st, bd);
            hipLaunchKernelGGL(bt_best_clear_kernel, dim3(1), dim3(256), 0, st, bd);
        }
        rc = batch_fdm_apply(bd, prec64, 1, st);
        if (rc) return rc;
        hipLaunchKernelGGL(bt_p_kernel<T>, dim3(BT_VGRID), dim3(256), 0, st, bd);
        YETI_HIP(hipGetLastError());
    }
    return 0;
}

}  // namespace yeti

struct yeti_batch {
    int prec64 = 0;
    int32_t chunk = 16;
    std::vector<yeti_ir *> members;
    std::vector<yeti::BReq> hreq;
    yeti::DevBuf<yeti::BReq> dreq;
    yeti::DevBuf<yeti::BTask> dlist[yeti::BT_NLISTS];
    yeti::DevBuf<int32_t> dcount;
    yeti::DevBuf<yeti_ir_status> dstat;
    int64_t cap = 0;             // tasks per list
    bool dirty = true;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    ~yeti_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    yeti::BatchDev dev() const {
        yeti::BatchDev bd;
        bd.req = dreq.p;
        for (int l = 0; l < yeti::BT_NLISTS; ++l) bd.list[l] = dlist[l].p;
        bd.count = dcount.p;
        return bd;
    }
};

namespace yeti {

// session fields of a member (the operator's come from batch_fill_operator)
static int batch_fill_session(const yeti_ir *s, int prec64, BReq *R) {
    if (!s->b || !s->x) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: session not begun");
    if (s->csr) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: sessions with an assembled operator are not batched");
    R->mask = s->mask;
    R->b = s->b;
    R->x = s->x;
    R->xbest = s->xbest.p;
    R->y64 = s->y64.p;
    if (prec64) {
        R->vr = s->vr64.p; R->vy = s->vy64.p; R->vz = s->vz64.p; R->vp = s->vp64.p; R->vq = s->vq64.p;
    } else {
        R->vr = s->vr32.p; R->vy = s->vy32.p; R->vz = s->vz32.p; R->vp = s->vp32.p; R->vq = s->vq32.p;
    }
    R->sc = s->sc.p;
    R->flags = s->flags.p;
    R->red = s->part.p;
    const int64_t n = s->n, ncp = s->ncp;
    R->vchunk = (int32_t)std::max<int64_t>(BATCH_CHUNK_MIN, (n + BATCH_RED_MAX - 1) / BATCH_RED_MAX);
    R->nvchunk = (int32_t)((n + R->vchunk - 1) / R->vchunk);
    R->cchunk = (int32_t)std::max<int64_t>(BATCH_CHUNK_MIN / 3, (ncp + BATCH_RED_MAX - 1) / BATCH_RED_MAX);
    R->ncchunk = (int32_t)((ncp + R->cchunk - 1) / R->cchunk);
    R->rtol2 = s->rtol * s->rtol;
    R->delta2 = s->delta * s->delta;
    R->max_iters = s->max_iters;
    R->max_updates = s->max_updates;
    return 0;
}

// task lists of the current members, uploaded to the device (on `st`, which is synchronised)
static int batch_upload(yeti_batch *b, hipStream_t st) {
    std::vector<BTask> lists[BT_NLISTS];
    for (int32_t mi = 0; mi < (int32_t)b->hreq.size(); ++mi) {
        const BReq &R = b->hreq[mi];
        const int pi = b->prec64 ? 0 : 1;
        const int64_t nbr = (int64_t)R.nbrick[pi][0] * R.nbrick[pi][1] * R.nbrick[pi][2];
        for (int64_t k = 0; k < nbr; ++k) lists[BT_BRICK1 + R.p - 1].push_back({mi, (int32_t)k, 0, 0});
        if (!b->prec64) {   // the FP64 replacement apply sweeps the FP64 bricks
            const int64_t nb64 = (int64_t)R.nbrick[0][0] * R.nbrick[0][1] * R.nbrick[0][2];
            for (int64_t k = 0; k < nb64; ++k) lists[BT_RBRICK1 + R.p - 1].push_back({mi, (int32_t)k, 0, 0});
        }
        for (int64_t k = 0; k < (R.ncp + BATCH_GATHER_BLOCK - 1) / BATCH_GATHER_BLOCK; ++k)
            lists[BT_GATHER].push_back({mi, (int32_t)k, 0, 0});
        for (int32_t k = 0; k < R.nvchunk; ++k) lists[BT_VEC].push_back({mi, k, 0, 0});
        for (int32_t k = 0; k < R.ncchunk; ++k) lists[BT_CP].push_back({mi, k, 0, 0});
        for (int prod = 0; prod < 6; ++prod) {
            const FdmGemm &g = R.g[prod];
            for (int c = 0; c < 3; ++c) {
                const int tiles = ((g.M[c] + 63) / 64) * ((g.N[c] + 63) / 64);
                for (int z = 0; z < g.nbatch[c]; ++z)
                    for (int tt = 0; tt < tiles; ++tt) lists[BT_GEMM0 + prod].push_back({mi, c, z, tt});
            }
        }
    }
    std::vector<int32_t> counts(BT_NLISTS + 1);
    for (int l = 0; l < BT_NLISTS; ++l) {
        if ((int64_t)lists[l].size() > b->cap)
            return set_error(YETI_ERR_BUDGET, "batch: %zu tasks in list %d, capacity %lld", lists[l].size(), l,
                             (long long)b->cap);
        counts[l] = (int32_t)lists[l].size();
        if (!lists[l].empty())
            YETI_HIP(hipMemcpyAsync(b->dlist[l].p, lists[l].data(), lists[l].size() * sizeof(BTask),
                                    hipMemcpyHostToDevice, st));
    }
    counts[BT_NLISTS] = (int32_t)b->hreq.size();
    if (!b->hreq.empty())
        YETI_HIP(hipMemcpyAsync(b->dreq.p, b->hreq.data(), b->hreq.size() * sizeof(BReq), hipMemcpyHostToDevice, st));
    YETI_HIP(hipMemcpyAsync(b->dcount.p, counts.data(), counts.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    YETI_HIP(hipStreamSynchronize(st));   // the host vectors go out of scope
    b->dirty = false;
    return 0;
}

}  // namespace yeti

extern "C" {

int32_t yeti_ir_create(const yeti_mf *m, const uint8_t *free_mask, int32_t precision, int32_t precond, double rtol,
                       double delta, int32_t chunk, int32_t max_iters, int32_t max_updates, void *arena,
                       size_t arena_bytes, void *stream, yeti_ir **out, size_t *used) {
    if (!m || !out) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: null argument");
    *out = nullptr;
    if (precision != 32 && precision != 64)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: precision must be 32 or 64");
    if (!(rtol > 0.0) || !(delta > 0.0) || delta >= 1.0 || chunk < 1 || max_iters < 1 || max_updates < 1)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: bad tolerance or iteration limits");
    if (arena && !free_mask) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: null free mask");
    hipStream_t st = (hipStream_t)stream;
    Arena a;
    a.base = (char *)arena;
    a.size = arena ? arena_bytes : 0;
    AllocScope scope(&a, st);
    auto *s = new yeti_ir();
    auto fail = [&](int code) {
        delete s;
        return code;
    };
    s->m = m;
    s->mask = free_mask;
    s->ncp = yeti_mf_nb_cp(m);
    s->n = 3 * (int64_t)s->ncp;
    s->inner64 = precision == 64;
    s->rtol = rtol;
    s->delta = delta;
    s->chunk = chunk;
    s->max_iters = max_iters;
    s->max_updates = max_updates;
    s->nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(NRED, (s->n + 255) / 256));
    s->nbc = (unsigned)std::max<int64_t>(1, std::min<int64_t>(NRED, (s->ncp + 255) / 256));
    const size_t n = (size_t)s->n;
    hipError_t e = s->blk.alloc(9 * (size_t)s->ncp);
    if (e == hipSuccess) e = s->y64.alloc(n);
    if (e == hipSuccess) e = s->xbest.alloc(n);
    if (e == hipSuccess) e = s->part.alloc(NRED);
    if (e == hipSuccess) e = s->sc.alloc(S_NSC);
    if (e == hipSuccess) e = s->flags.alloc(F_ALLOC);
    if (s->inner64) {
        if (e == hipSuccess) e = s->inv64.alloc(9 * (size_t)s->ncp);
        for (auto *v : {&s->vr64, &s->vy64, &s->vz64, &s->vp64, &s->vq64})
            if (e == hipSuccess) e = v->alloc(n);
    } else {
        if (e == hipSuccess) e = s->inv32.alloc(9 * (size_t)s->ncp);
        for (auto *v : {&s->vr32, &s->vy32, &s->vz32, &s->vp32, &s->vq32})
            if (e == hipSuccess) e = v->alloc(n);
    }
    if (e != hipSuccess) return fail(set_error(YETI_ERR_BUDGET, "ir: device memory: %s", hipGetErrorString(e)));
    if (precond != 0 && precond != 1) return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "ir: unknown preconditioner"));
    if (!arena) {
        if (used) *used = a.used + (precond == 1 ? fdm_bytes_bound(m, precision) : 0);
        delete s;   // counting only
        return 0;
    }
    if (precond == 1) {
        std::vector<uint8_t> mh(n);
        if (hipMemcpyAsync(mh.data(), free_mask, n, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return fail(set_error(YETI_ERR_DEVICE, "ir: mask copy failed"));
        int rcf = fdm_create(m, mh.data(), precision, st, &s->fdm);
        if (rcf) return fail(rcf);
    }
    if (used) *used = a.used;
    if (precond == 1) {   // fast diagonalisation: the block-Jacobi arrays stay unused
        *out = s;
        return 0;
    }
    // block-Jacobi preconditioner of the free dofs
    int rc = yeti_mf_diag_blocks_f64(m, s->blk.p, st);
    if (rc) return fail(rc);
    hipLaunchKernelGGL(block_inverse_kernel, dim3((s->ncp + 255) / 256), dim3(256), 0, st, s->ncp, s->blk.p, free_mask,
                       s->inner64 ? s->inv64.p : nullptr, s->inner64 ? nullptr : s->inv32.p);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "ir: preconditioner set-up failed"));
    *out = s;
    return 0;
}

int32_t yeti_ir_set_csr(yeti_ir *s, const yeti_csr *c, const double *vals) {
    if (!s || (c && !vals)) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: null argument");
    if (s->gexec) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: set the operator before the first graph step");
    if (c && csr_nrows(c) != s->n) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: CSR has %lld rows, session %lld",
                                                    (long long)csr_nrows(c), (long long)s->n);
    s->csr = c;
    s->csr_vals = c ? vals : nullptr;
    return 0;
}

int32_t yeti_ir_begin(yeti_ir *s, const double *b, double *x, void *stream) {
    if (!s || !b || !x) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: null argument");
    hipStream_t st = (hipStream_t)stream;
    if ((s->b != b || s->x != x) && s->gexec) {
        // the captured step refers to b and x: capture again for new buffers
        (void)hipGraphExecDestroy(s->gexec);
        (void)hipGraphDestroy(s->graph);
        s->gexec = nullptr;
        s->graph = nullptr;
    }
    s->b = b;
    s->x = x;
    if (s->inner64) return ir_begin_launches<double>(s, s->vr64.p, s->vy64.p, s->vz64.p, s->vp64.p, s->inv64.p, st);
    return ir_begin_launches<float>(s, s->vr32.p, s->vy32.p, s->vz32.p, s->vp32.p, s->inv32.p, st);
}

int32_t yeti_ir_step(yeti_ir *s, int32_t use_graph, void *stream) {
    if (!s || !s->b) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: step before begin");
    hipStream_t st = (hipStream_t)stream;
    if (!use_graph) return ir_step_any(s, st);
    if (!st) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: graph steps need a non-null stream");
    if (!s->gexec) {
        YETI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = ir_step_any(s, st);
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return set_error(YETI_ERR_DEVICE, "ir: capture failed: %s", hipGetErrorString(ec));
        hipGraphExec_t ex = nullptr;
        hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            return set_error(YETI_ERR_DEVICE, "ir: graph instantiation failed: %s", hipGetErrorString(ei));
        }
        s->graph = g;
        s->gexec = ex;
    }
    YETI_HIP(hipGraphLaunch(s->gexec, st));
    return 0;
}

int32_t yeti_ir_status_async(const yeti_ir *s, yeti_ir_status *host_status, void *stream) {
    if (!s || !host_status) return set_error(YETI_ERR_INVALID_ARGUMENT, "ir: null argument");
    hipStream_t st = (hipStream_t)stream;
    // flags then scalars, in the layout of yeti_ir_status (the caller waits on the stream)
    static_assert(offsetof(yeti_ir_status, rz) == sizeof(int32_t) * F_NFLAGS, "status layout");
    static_assert(sizeof(yeti_ir_status) == sizeof(int32_t) * F_NFLAGS + sizeof(double) * S_NSC, "status layout");
    YETI_HIP(hipMemcpyAsync(&host_status->active, s->flags.p, sizeof(int32_t) * F_NFLAGS, hipMemcpyDeviceToHost, st));
    YETI_HIP(hipMemcpyAsync(&host_status->rz, s->sc.p, sizeof(double) * S_NSC, hipMemcpyDeviceToHost, st));
    return 0;
}

void yeti_ir_destroy(yeti_ir *s) { delete s; }

int32_t yeti_batch_create(int32_t precision, int32_t chunk, int64_t task_capacity, yeti_batch **out) {
    if (!out) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: null handle pointer");
    *out = nullptr;
    if ((precision != 32 && precision != 64) || chunk < 1 || task_capacity < 1 || task_capacity > INT32_MAX)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: bad precision, chunk or capacity");
    auto *b = new yeti_batch();
    b->prec64 = precision == 64;
    b->chunk = chunk;
    b->cap = task_capacity;
    hipError_t e = b->dreq.alloc(BATCH_MAX_MEMBERS);
    for (int l = 0; l < BT_NLISTS && e == hipSuccess; ++l) e = b->dlist[l].alloc((size_t)task_capacity);
    if (e == hipSuccess) e = b->dcount.alloc(BT_NLISTS + 1);
    if (e == hipSuccess) e = b->dstat.alloc(BATCH_MAX_MEMBERS);
    if (e == hipSuccess) {
        std::vector<int32_t> zero(BT_NLISTS + 1, 0);
        e = hipMemcpy(b->dcount.p, zero.data(), zero.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        delete b;
        return set_error(YETI_ERR_DEVICE, "batch: device tables: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

int32_t yeti_batch_add(yeti_batch *b, yeti_ir *s) {
    if (!b || !s) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: null argument");
    if ((int)b->members.size() >= BATCH_MAX_MEMBERS) return set_error(YETI_ERR_BUDGET, "batch: full");
    if (s->inner64 != b->prec64) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: session precision differs");
    for (auto *m : b->members)
        if (m == s) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: session already a member");
    BReq R{};
    int rc = batch_fill_operator(s->m, s->fdm, b->prec64, &R);
    if (rc) return rc;
    rc = batch_fill_session(s, b->prec64, &R);
    if (rc) return rc;
    b->members.push_back(s);
    b->hreq.push_back(R);
    b->dirty = true;
    return 0;
}

int32_t yeti_batch_remove(yeti_batch *b, yeti_ir *s) {
    if (!b || !s) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: null argument");
    for (size_t k = 0; k < b->members.size(); ++k)
        if (b->members[k] == s) {
            b->members.erase(b->members.begin() + k);
            b->hreq.erase(b->hreq.begin() + k);
            b->dirty = true;
            return 0;
        }
    return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: not a member");
}

int32_t yeti_batch_size(const yeti_batch *b) { return b ? (int32_t)b->members.size() : -1; }

int32_t yeti_batch_step(yeti_batch *b, int32_t use_graph, void *stream) {
    if (!b) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: null handle");
    hipStream_t st = (hipStream_t)stream;
    if (b->dirty) {
        int rc = batch_upload(b, st);
        if (rc) return rc;
    }
    if (b->members.empty()) return 0;
    const BatchDev bd = b->dev();
    auto launches = [&]() {
        return b->prec64 ? batch_step_launches<double>(bd, 1, b->chunk, st)
                         : batch_step_launches<float>(bd, 0, b->chunk, st);
    };
    if (!use_graph) return launches();
    if (!st) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: graph steps need a non-null stream");
    if (!b->gexec) {
        // the step reads every member and count from the device tables, so one capture serves
        // every later membership
        YETI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = launches();
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return set_error(YETI_ERR_DEVICE, "batch: capture failed: %s", hipGetErrorString(ec));
        hipGraphExec_t ex = nullptr;
        hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            return set_error(YETI_ERR_DEVICE, "batch: graph instantiation failed: %s", hipGetErrorString(ei));
        }
        b->graph = g;
        b->gexec = ex;
    }
    YETI_HIP(hipGraphLaunch(b->gexec, st));
    return 0;
}

int32_t yeti_batch_status_async(yeti_batch *b, yeti_ir_status *host_status, void *stream) {
    if (!b || !host_status) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: null argument");
    hipStream_t st = (hipStream_t)stream;
    if (b->dirty) {
        int rc = batch_upload(b, st);
        if (rc) return rc;
    }
    if (b->members.empty()) return 0;
    hipLaunchKernelGGL(bt_status_kernel, dim3(64), dim3(64), 0, st, b->dev(), b->dstat.p);
    YETI_HIP(hipGetLastError());
    YETI_HIP(hipMemcpyAsync(host_status, b->dstat.p, b->members.size() * sizeof(yeti_ir_status), hipMemcpyDeviceToHost,
                            st));
    return 0;
}

void yeti_batch_destroy(yeti_batch *b) { delete b; }

}  // extern "C"
