// solve.hip -- device-resident Jacobi-preconditioned conjugate gradients on the assembled CSR
// (FP64), with fixed-order reductions so a solve is bit-reproducible.
//
// Reference behaviour replaced: the callers solve (K + K^T)[free][:, free] u = F[free] with
// scipy (spsolve, benchs/volume_force/test_volume_force_3d.py:67-79), and YETI's matrix-free CG
// iga_mf_cg_3d (iga_wq_mf/fsrc/mf_iga_solver.f90:642-958) iterates alpha = rsold / (p . Ap).
// Here the free-dof restriction is a 0/1 mask on the full vector: masked entries of x, r, z and
// p stay exactly zero, which is the reduced system. Stopping rule: ||r||_2 <= rtol ||b||_2.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "common.hpp"

namespace yeti {

constexpr int RED_BLOCKS = 1024;     // partial sums per reduction (fixed -> fixed summation order)
constexpr int RED_THREADS = 256;

__device__ inline double block_sum(double v, double *sh) {
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

// y = mask .* (A x) ; partial[b] = sum over the block's rows of x . y
__global__ void __launch_bounds__(256) spmv_dot_kernel(int64_t nrows, const int64_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ colind,
                                                       const double *__restrict__ vals,
                                                       const double *__restrict__ x,
                                                       const uint8_t *__restrict__ mask,
                                                       double *__restrict__ y, double *__restrict__ partial) {
    __shared__ double sh[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int64_t row = (int64_t)blockIdx.x * 4 + w; row < nrows; row += (int64_t)gridDim.x * 4) {
        double s = 0.0;
        if (mask[row]) {
            for (int64_t k = rowptr[row] + lane; k < rowptr[row + 1]; k += 64) s += vals[k] * x[colind[k]];
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        }
        if (lane == 0) {
            y[row] = s;
            acc += s * x[row];
        }
    }
    double t = block_sum(lane == 0 ? acc : 0.0, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void finalize_kernel(const double *partial, int n, double *out) {
    __shared__ double sh[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += partial[i];
    double t = block_sum(v, sh);
    if (threadIdx.x == 0) *out = t;
}

// scalars: [0] rz, [1] pq, [2] rr, [3] bb, [4] rz_new
__global__ void __launch_bounds__(256) update_xr_kernel(int64_t n, const double *__restrict__ sc,
                                                        double *__restrict__ x, double *__restrict__ r,
                                                        double *__restrict__ z, const double *__restrict__ p,
                                                        const double *__restrict__ q,
                                                        const double *__restrict__ dinv,
                                                        double *__restrict__ part_rz, double *__restrict__ part_rr) {
    __shared__ double sh[4];
    const double alpha = sc[0] / sc[1];
    double a_rz = 0.0, a_rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        double ri = r[i] - alpha * q[i];
        r[i] = ri;
        double zi = dinv[i] * ri;
        z[i] = zi;
        a_rz += ri * zi;
        a_rr += ri * ri;
    }
    double t1 = block_sum(a_rz, sh);
    double t2 = block_sum(a_rr, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = t1;
        part_rr[blockIdx.x] = t2;
    }
}

__global__ void __launch_bounds__(256) update_p_kernel(int64_t n, double *__restrict__ sc, double *__restrict__ p,
                                                       const double *__restrict__ z) {
    const double beta = sc[4] / sc[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

__global__ void roll_scalars_kernel(double *sc) { sc[0] = sc[4]; }

// r = mask b, z = dinv r, p = z; partials of r.z and b.b
__global__ void __launch_bounds__(256) init_kernel(int64_t n, const double *__restrict__ b, const uint8_t *mask,
                                                   const double *__restrict__ dinv, double *__restrict__ x,
                                                   double *__restrict__ r, double *__restrict__ z,
                                                   double *__restrict__ p, double *__restrict__ part_rz,
                                                   double *__restrict__ part_bb) {
    __shared__ double sh[4];
    double a_rz = 0.0, a_bb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double bi = mask[i] ? b[i] : 0.0;
        x[i] = 0.0;
        r[i] = bi;
        double zi = dinv[i] * bi;
        z[i] = zi;
        p[i] = zi;
        a_rz += bi * zi;
        a_bb += bi * bi;
    }
    double t1 = block_sum(a_rz, sh);
    double t2 = block_sum(a_bb, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = t1;
        part_bb[blockIdx.x] = t2;
    }
}

// partial[b] = max over the block's entries of |v[i]| (masked entries count as 0)
__global__ void __launch_bounds__(256) absmax_kernel(int64_t n, const double *__restrict__ v, const uint8_t *mask,
                                                     double *__restrict__ partial) {
    __shared__ double sh[4];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!mask || mask[i]) m = fmax(m, fabs(v[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ void finalize_max_kernel(const double *partial, int n, double *out) {
    if (threadIdx.x != 0) return;
    double m = 0.0;
    for (int i = 0; i < n; ++i) m = fmax(m, partial[i]);
    *out = m;
}

// dinv[i] = mask[i] / A[i][i]
__global__ void jacobi_kernel(int64_t nrows, const int64_t *rowptr, const int32_t *colind, const double *vals,
                              const uint8_t *mask, double *dinv) {
    int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    int64_t lo = rowptr[row], hi = rowptr[row + 1] - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (colind[mid] < row) lo = mid + 1; else hi = mid;
    }
    double d = (lo < rowptr[row + 1] && colind[lo] == row) ? vals[lo] : 0.0;
    dinv[row] = (mask[row] && d != 0.0) ? 1.0 / d : 0.0;
}

// Work vectors of one PCG solve: r, z, p, q, dinv (nrows each), two partial-sum arrays
// (RED_BLOCKS each, or PCG_PARTS for the nodal-block product) and 8 scalars.
struct PcgWork {
    double *r, *z, *p, *q, *dinv, *part1, *part2, *sc;
};
constexpr int PCG_PARTS = 4096;

// The operator of a PCG solve: a scalar CSR (rowptr, colind, vals), or a device CSR handle whose
// nodal 3x3 blocks the product walks (blocked != null; rowptr / colind are then its scalar
// pattern, used for the diagonal).
struct PcgOp {
    int64_t nrows;
    const int64_t *rowptr;
    const int32_t *colind;
    const double *vals;
    const yeti_csr *blocked;
};

// Stop test. norm 2: ||r||_2 <= rtol ||b||_2; norm 0: max|r| <= rtol max|b| over the free dofs,
// the rule of iga_mf_cg_3d (maxval(abs(r)) / maxval(abs(b)), mf_iga_solver.f90:779,935).
static int pcg_run(const PcgOp &op, const double *b, const uint8_t *free_mask, double *x, double rtol, int32_t maxit,
                   int32_t check_every, int norm, int32_t *iters_out, double *relres_out, hipStream_t st,
                   const PcgWork &w) {
    const int64_t nrows = op.nrows;
    const int64_t *rowptr = op.rowptr;
    const int32_t *colind = op.colind;
    const double *vals = op.vals;
    if (check_every <= 0) check_every = 10;
    const unsigned nb = (unsigned)std::min<int64_t>(RED_BLOCKS, (nrows + 255) / 256);
    hipLaunchKernelGGL(jacobi_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, st, nrows, rowptr,
                       colind, vals, free_mask, w.dinv);
    hipLaunchKernelGGL(init_kernel, dim3(nb), dim3(256), 0, st, nrows, b, free_mask, w.dinv, x, w.r, w.z, w.p,
                       w.part1, w.part2);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, st, w.part1, (int)nb, w.sc + 0);
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, st, w.part2, (int)nb, w.sc + 3);
    YETI_HIP(hipGetLastError());
    double hs[8] = {0};
    YETI_HIP(hipMemcpyAsync(hs, w.sc, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
    YETI_HIP(hipStreamSynchronize(st));
    const double bb = hs[3];
    if (bb == 0.0) {
        if (iters_out) *iters_out = 0;
        if (relres_out) *relres_out = 0.0;
        return 0;
    }
    // norm 0: the reference's max|b| over the free dofs (r = mask b at the start)
    double bmax = 0.0;
    if (norm == 0) {
        hipLaunchKernelGGL(absmax_kernel, dim3(nb), dim3(256), 0, st, nrows, (const double *)w.r, free_mask, w.part2);
        hipLaunchKernelGGL(finalize_max_kernel, dim3(1), dim3(64), 0, st, (const double *)w.part2, (int)nb, w.sc + 6);
        YETI_HIP(hipGetLastError());
        YETI_HIP(hipMemcpyAsync(&bmax, w.sc + 6, sizeof(double), hipMemcpyDeviceToHost, st));
        YETI_HIP(hipStreamSynchronize(st));
    }
    const unsigned nspmv = (unsigned)std::min<int64_t>(RED_BLOCKS, (nrows + 3) / 4);
    int it = 0;
    double rel = 1.0;
    while (it < maxit) {
        int nparts = (int)nspmv;
        if (op.blocked) {
            int rc = csr_block_spmv(op.blocked, vals, w.p, free_mask, w.q, w.part1, 2, &nparts, st);
            if (rc) return rc;
        } else {
            hipLaunchKernelGGL(spmv_dot_kernel, dim3(nspmv), dim3(256), 0, st, nrows, rowptr, colind, vals, w.p,
                               free_mask, w.q, w.part1);
        }
        hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, st, w.part1, nparts, w.sc + 1);
        hipLaunchKernelGGL(update_xr_kernel, dim3(nb), dim3(256), 0, st, nrows, w.sc, x, w.r, w.z, w.p, w.q, w.dinv,
                           w.part1, w.part2);
        hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, st, w.part1, (int)nb, w.sc + 4);
        hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(256), 0, st, w.part2, (int)nb, w.sc + 2);
        hipLaunchKernelGGL(update_p_kernel, dim3(nb), dim3(256), 0, st, nrows, w.sc, w.p, w.z);
        hipLaunchKernelGGL(roll_scalars_kernel, dim3(1), dim3(1), 0, st, w.sc);
        YETI_HIP(hipGetLastError());
        ++it;
        if (it % check_every == 0 || it == maxit) {
            double rr = 0.0;
            if (norm == 0) {
                hipLaunchKernelGGL(absmax_kernel, dim3(nb), dim3(256), 0, st, nrows, (const double *)w.r, free_mask,
                                   w.part2);
                hipLaunchKernelGGL(finalize_max_kernel, dim3(1), dim3(64), 0, st, (const double *)w.part2, (int)nb,
                                   w.sc + 6);
                YETI_HIP(hipGetLastError());
                YETI_HIP(hipMemcpyAsync(&rr, w.sc + 6, sizeof(double), hipMemcpyDeviceToHost, st));
                YETI_HIP(hipStreamSynchronize(st));
                rel = rr / bmax;
            } else {
                YETI_HIP(hipMemcpyAsync(&rr, w.sc + 2, sizeof(double), hipMemcpyDeviceToHost, st));
                YETI_HIP(hipStreamSynchronize(st));
                rel = std::sqrt(rr / bb);
            }
            if (!(rel > rtol)) break;     // also stops on NaN
        }
    }
    if (iters_out) *iters_out = it;
    if (relres_out) *relres_out = rel;
    if (std::isnan(rel)) return set_error(YETI_ERR_NOT_CONVERGED, "pcg: residual is NaN");
    if (rel > rtol) return set_error(YETI_ERR_NOT_CONVERGED, "pcg: %d iterations, relative residual %.3e > %.3e", it, rel, rtol);
    return 0;
}

// ---- the allocation-free variant's scratch ------------------------------------------------
// One file-scope device array holds every work buffer of yeti_pcg_csr_noalloc: 8 scalars, two
// partial-sum arrays and five vectors of up to PCG_NOALLOC_MAX_ROWS rows (about 42 MB, held on
// each device the library runs on until the process ends). Host code never passes `scratch`
// itself (in host code the name is a stand-in, not a device address): its address on each
// device is looked up once with hipGetSymbolAddress, that device current, and kept here.
constexpr int64_t PCG_NOALLOC_MAX_ROWS = int64_t(1) << 20;
constexpr int64_t PCG_SCRATCH_DOUBLES = 8 + 2 * RED_BLOCKS + 5 * PCG_NOALLOC_MAX_ROWS;
static __device__ __align__(256) double scratch[PCG_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
static std::mutex scratch_mutex;
static double *scratch_ptr[MAX_DEVICES] = {};

static int scratch_on_current_device(double **out) {
    int dev = 0;
    YETI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(YETI_ERR_DEVICE, "pcg: device id %d out of range", dev);
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(YETI_ERR_DEVICE, "pcg: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        scratch_ptr[dev] = (double *)ptr;
    }
    *out = scratch_ptr[dev];
    return 0;
}

}  // namespace yeti

using namespace yeti;

extern "C" {

int32_t yeti_pcg_csr(int64_t nrows, const int64_t *rowptr, const int32_t *colind, const double *vals,
                     const double *b, const uint8_t *free_mask, double *x, double rtol, int32_t maxit,
                     int32_t check_every, int32_t *iters_out, double *relres_out, void *stream) {
    if (!rowptr || !colind || !vals || !b || !free_mask || !x || nrows <= 0 || maxit < 0)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "pcg: bad argument");
    DevBuf<double> r, z, p, q, dinv, part1, part2, sc;
    YETI_HIP(r.alloc(nrows));
    YETI_HIP(z.alloc(nrows));
    YETI_HIP(p.alloc(nrows));
    YETI_HIP(q.alloc(nrows));
    YETI_HIP(dinv.alloc(nrows));
    YETI_HIP(part1.alloc(RED_BLOCKS));
    YETI_HIP(part2.alloc(RED_BLOCKS));
    YETI_HIP(sc.alloc(8));
    const PcgWork w{r.p, z.p, p.p, q.p, dinv.p, part1.p, part2.p, sc.p};
    const PcgOp op{nrows, rowptr, colind, vals, nullptr};
    return pcg_run(op, b, free_mask, x, rtol, maxit, check_every, 2, iters_out, relres_out, (hipStream_t)stream, w);
}

int32_t yeti_csr_pcg(const yeti_csr *c, const double *vals, const double *b, const uint8_t *free_mask, double *x,
                     double rtol, int32_t maxit, int32_t check_every, int32_t norm, int32_t *iters_out,
                     double *relres_out, void *stream) {
    if (!c || !vals || !b || !free_mask || !x || maxit < 0 || (norm != 0 && norm != 2))
        return set_error(YETI_ERR_INVALID_ARGUMENT, "pcg: bad argument");
    const int64_t nrows = csr_nrows(c);
    if (nrows <= 0) return set_error(YETI_ERR_INVALID_ARGUMENT, "pcg: empty matrix");
    DevBuf<double> r, z, p, q, dinv, part1, part2, sc;
    YETI_HIP(r.alloc(nrows));
    YETI_HIP(z.alloc(nrows));
    YETI_HIP(p.alloc(nrows));
    YETI_HIP(q.alloc(nrows));
    YETI_HIP(dinv.alloc(nrows));
    YETI_HIP(part1.alloc(PCG_PARTS));
    YETI_HIP(part2.alloc(PCG_PARTS));
    YETI_HIP(sc.alloc(8));
    const PcgWork w{r.p, z.p, p.p, q.p, dinv.p, part1.p, part2.p, sc.p};
    const PcgOp op{nrows, csr_rowptr_dev(c), csr_colind_dev(c), vals, c};
    return pcg_run(op, b, free_mask, x, rtol, maxit, check_every, norm, iters_out, relres_out, (hipStream_t)stream, w);
}

int32_t yeti_pcg_csr_noalloc(int64_t nrows, const int64_t *rowptr, const int32_t *colind, const double *vals,
                             const double *b, const uint8_t *free_mask, double *x, double rtol, int32_t maxit,
                             int32_t check_every, int32_t *iters_out, double *relres_out, void *stream) {
    if (!rowptr || !colind || !vals || !b || !free_mask || !x || nrows <= 0 || maxit < 0)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "pcg: bad argument");
    if (nrows > PCG_NOALLOC_MAX_ROWS)
        return set_error(YETI_ERR_SIZE_OVERFLOW, "pcg (no allocation): %lld rows exceed the scratch maximum of %lld",
                         (long long)nrows, (long long)PCG_NOALLOC_MAX_ROWS);
    double *s = nullptr;
    if (int rc = scratch_on_current_device(&s)) return rc;
    // every entry the solve reads is written by it first (scalars by the reductions, vectors by
    // the Jacobi and start kernels), so nothing left by an earlier call needs clearing
    const PcgWork w{s + 8 + 2 * RED_BLOCKS,
                    s + 8 + 2 * RED_BLOCKS + nrows,
                    s + 8 + 2 * RED_BLOCKS + 2 * nrows,
                    s + 8 + 2 * RED_BLOCKS + 3 * nrows,
                    s + 8 + 2 * RED_BLOCKS + 4 * nrows,
                    s + 8,
                    s + 8 + RED_BLOCKS,
                    s};
    const PcgOp op{nrows, rowptr, colind, vals, nullptr};
    return pcg_run(op, b, free_mask, x, rtol, maxit, check_every, 2, iters_out, relres_out, (hipStream_t)stream, w);
}

}  // extern "C"
