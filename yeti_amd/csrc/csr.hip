// csr.hip -- the CSR form of K_side + K_side^T built on the device, and its FP64 SpMV.
//
// Reference behaviour replaced: the callers of sys_linmat_lindef_static turn its COO output into
// scipy's canonical CSC/CSR and add the transpose, `(K_side + K_side^T)`
// (overlay/iga_model.py:391-395; benchs/volume_force/test_volume_force_3d.py:69-79).
//
// Pattern (once, from the resident connectivity): CP -> (element, local node) incidence,
// sorted by element; per CP row the sorted set of CPs sharing an element (3x3 block pattern);
// per block the list of element contributions in ascending element order. Values (every
// assembly): one thread per block sums its contributions read from the COO values in that fixed
// order -- deterministic, no floating-point atomics. Scalar CSR: dof row 3 I + k holds, for each
// block column J of CP row I (ascending), the columns 3 J, 3 J + 1, 3 J + 2: scipy's canonical
// layout of the structural pattern.
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace yeti {
// declared in assembly.hip
struct AsmView {
    int32_t nb_cp;
    int32_t npatch;
    const int32_t *const *ien;      // per patch, device
    const int32_t *const *elems;    // per patch active elements (patch-local), device
    const int64_t *const *offs;     // per patch COO offsets, device
    const int32_t *nnode;           // per patch, host
    const int32_t *n_active;        // per patch, host
};
int asm_view(const yeti_asm *h, AsmView *v);
}  // namespace yeti

struct yeti_csr {
    int32_t nb_cp = 0;
    int64_t nblocks = 0;                 // 3x3 blocks
    yeti::DevBuf<int64_t> bptr;          // [nb_cp + 1]
    yeti::DevBuf<int32_t> bcol;          // [nblocks]
    yeti::DevBuf<int64_t> cptr;          // [nblocks + 1] contribution offsets
    yeti::DevBuf<int64_t> contrib;       // (COO position << 2) | kind
    yeti::DevBuf<int64_t> rowptr;        // scalar CSR [3 nb_cp + 1]
    yeti::DevBuf<int32_t> colind;        // scalar CSR [9 nblocks]
};

namespace yeti {

// ---- incidence --------------------------------------------------------------------------
__global__ void inc_count_kernel(const int32_t *ien, const int32_t *elems, int n_active, int nn,
                                 int32_t *count) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_active * nn) return;
    int ea = (int)(t / nn), a = (int)(t % nn);
    int e = elems[ea];
    atomicAdd(&count[ien[(int64_t)nn * e + a] - 1], 1);
}

// entry = (patch << 58) | (active index << 10) | local node ; sorted per CP afterwards
__global__ void inc_fill_kernel(const int32_t *ien, const int32_t *elems, int n_active, int nn, int patch,
                                const int64_t *ptr, int32_t *cursor, int64_t *ent) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_active * nn) return;
    int ea = (int)(t / nn), a = (int)(t % nn);
    int e = elems[ea];
    int cp = ien[(int64_t)nn * e + a] - 1;
    int slot = atomicAdd(&cursor[cp], 1);
    ent[ptr[cp] + slot] = ((int64_t)patch << 58) | ((int64_t)ea << 10) | a;
}

__global__ void inc_sort_kernel(int32_t nb_cp, const int64_t *ptr, int64_t *ent) {
    int cp = blockIdx.x * blockDim.x + threadIdx.x;
    if (cp >= nb_cp) return;
    int64_t b = ptr[cp], n = ptr[cp + 1] - b;
    for (int64_t i = 1; i < n; ++i) {          // insertion sort: lists are short (<= 729)
        int64_t v = ent[b + i];
        int64_t j = i - 1;
        while (j >= 0 && ent[b + j] > v) {
            ent[b + j + 1] = ent[b + j];
            --j;
        }
        ent[b + j + 1] = v;
    }
}

// exclusive scan of int32/int64 counts into int64 offsets (single-pass over blocks of 1024,
// then block sums, then add): three small kernels.
template <typename T>
__global__ void scan_block_kernel(const T *in, int64_t n, int64_t *out, int64_t *bsum) {
    __shared__ int64_t s[1024];
    int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    int64_t v = i < n ? (int64_t)in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t add = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == 1023) bsum[blockIdx.x] = s[1023];
}

__global__ void scan_add_kernel(int64_t *out, int64_t n, const int64_t *boff) {
    int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    if (i < n) out[i] += boff[blockIdx.x];
}

// exclusive scan of n values into out[0..n], out[n] = total. Host-driven recursion.
template <typename T>
static int exclusive_scan(const T *in, int64_t n, int64_t *out, hipStream_t st) {
    if (n == 0) {
        YETI_HIP(hipMemsetAsync(out, 0, sizeof(int64_t), st));
        return 0;
    }
    int64_t nb = (n + 1023) / 1024;
    DevBuf<int64_t> bsum, boff;
    YETI_HIP(bsum.alloc(nb));
    YETI_HIP(boff.alloc(nb + 1));
    hipLaunchKernelGGL(scan_block_kernel<T>, dim3((unsigned)nb), dim3(1024), 0, st, in, n, out, bsum.p);
    YETI_HIP(hipGetLastError());
    if (nb > 1) {   // one block needs no carry; recursing on it would never end
        int rc = exclusive_scan<int64_t>(bsum.p, nb, boff.p, st);
        if (rc) return rc;
        hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)nb), dim3(1024), 0, st, out, n, boff.p);
        YETI_HIP(hipGetLastError());
    }
    // total = last offset + last count
    int64_t last_off = 0;
    T last_in{};
    YETI_HIP(hipMemcpyAsync(&last_off, out + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    YETI_HIP(hipMemcpyAsync(&last_in, in + n - 1, sizeof(T), hipMemcpyDeviceToHost, st));
    YETI_HIP(hipStreamSynchronize(st));
    int64_t total = last_off + (int64_t)last_in;
    YETI_HIP(hipMemcpyAsync(out + n, &total, sizeof(int64_t), hipMemcpyHostToDevice, st));
    YETI_HIP(hipStreamSynchronize(st));
    return 0;
}

// ---- block pattern ----------------------------------------------------------------------
struct PatchIen {
    const int32_t *ien[8];
    const int32_t *elems[8];
    const int64_t *offs[8];
    int32_t nn[8];
};

constexpr int HASH = 4096;     // per-wave open-addressing set (>= 2 x 729 unique neighbours)

// One wave per CP row: collect the CPs of every element touching the row into an LDS hash set,
// then (pass 2) sort them into the row's block columns.
template <bool FILL>
__global__ void __launch_bounds__(256) row_pattern_kernel(int32_t nb_cp, const int64_t *inc_ptr,
                                                          const int64_t *inc, PatchIen pi,
                                                          int32_t *row_len, const int64_t *bptr,
                                                          int32_t *bcol, int32_t *overflow) {
    __shared__ int32_t table[4][HASH];
    __shared__ int32_t cnt[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + w;
    int32_t *T = table[w];
    for (int i = lane; i < HASH; i += 64) T[i] = -1;
    if (lane == 0) cnt[w] = 0;
    __syncthreads();
    if (row < nb_cp) {
        for (int64_t k = inc_ptr[row]; k < inc_ptr[row + 1]; ++k) {
            int64_t en = inc[k];
            int patch = (int)(en >> 58), ea = (int)((en >> 10) & ((1LL << 48) - 1));
            int nn = pi.nn[patch];
            int e = pi.elems[patch][ea];
            const int32_t *ie = pi.ien[patch] + (int64_t)nn * e;
            for (int b = lane; b < nn; b += 64) {
                int32_t J = ie[b] - 1;
                uint32_t h = ((uint32_t)J * 2654435761u) & (HASH - 1);
                int probe = 0;
                for (; probe < HASH; ++probe) {        // bounded: the table holds HASH keys
                    int32_t old = atomicCAS(&T[h], -1, J);
                    if (old == -1) {
                        atomicAdd(&cnt[w], 1);
                        break;
                    }
                    if (old == J) break;
                    h = (h + 1) & (HASH - 1);
                }
                if (probe == HASH) atomicOr(overflow, 1);
            }
        }
    }
    __syncthreads();
    const bool live = row < nb_cp;
    const int n = cnt[w];
    if (!FILL) {
        if (live && lane == 0) {
            row_len[row] = n;
            if (n > 1024) atomicOr(overflow, 1);   // more neighbours than the sort buffer holds
        }
        return;
    }
    // compact then rank-sort (n <= 729): each lane ranks its entries against all others.
    // Every wave reaches every barrier (rows past the end only skip the writes).
    __shared__ int32_t comp[4][1024];
    __shared__ int32_t pos[4];
    if (lane == 0) pos[w] = 0;
    __syncthreads();
    for (int i = lane; i < HASH; i += 64) {
        int32_t v = T[i];
        if (v >= 0) {
            int slot = atomicAdd(&pos[w], 1);
            if (slot < 1024) comp[w][slot] = v;
        }
    }
    __syncthreads();
    if (!live || n > 1024) return;
    const int64_t out = bptr[row];
    for (int i = lane; i < n; i += 64) {
        int32_t v = comp[w][i];
        int r = 0;
        for (int j = 0; j < n; ++j) r += comp[w][j] < v;
        bcol[out + r] = v;
    }
}

// Per row: count (pass 1) or write (pass 2) the contributions of each block, in element order.
// kind 0: block (a, b) stored as is (a < b); 1: stored as (b, a) (a > b, transpose);
// 2: diagonal block (a == b), stored as its 6 upper entries with the diagonal halved.
template <bool FILL>
__global__ void __launch_bounds__(64) contrib_kernel(int32_t nb_cp, const int64_t *inc_ptr, const int64_t *inc,
                                                     PatchIen pi, const int64_t *bptr, const int32_t *bcol,
                                                     int32_t *bcount, const int64_t *cptr, int64_t *contrib) {
    __shared__ int32_t cur[1024];
    const int row = blockIdx.x, lane = threadIdx.x;
    if (row >= nb_cp) return;
    const int64_t b0 = bptr[row];
    const int nb = (int)(bptr[row + 1] - b0);
    if (nb > 1024) return;      // rejected by the pattern pass (overflow flag)
    for (int i = lane; i < nb; i += 64) cur[i] = 0;
    __syncthreads();
    for (int64_t k = inc_ptr[row]; k < inc_ptr[row + 1]; ++k) {
        int64_t en = inc[k];
        int patch = (int)(en >> 58), ea = (int)((en >> 10) & ((1LL << 48) - 1)), a = (int)(en & 1023);
        int nn = pi.nn[patch];
        int e = pi.elems[patch][ea];
        int64_t base = pi.offs[patch][ea];
        const int32_t *ie = pi.ien[patch] + (int64_t)nn * e;
        for (int b = lane; b < nn; b += 64) {
            int32_t J = ie[b] - 1;
            int lo = 0, hi = nb - 1;       // binary search of J among the row's sorted columns
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (bcol[b0 + mid] < J) lo = mid + 1; else hi = mid;
            }
            // one contribution per (element, block): no two lanes hit the same block here
            int slot = cur[lo]++;
            if (FILL) {
                int64_t code;
                if (a < b) code = ((base + 9LL * b * (b - 1) / 2 + 6LL * b + 9LL * a) << 2) | 0;
                else if (a > b) code = ((base + 9LL * a * (a - 1) / 2 + 6LL * a + 9LL * b) << 2) | 1;
                else code = ((base + 9LL * a * (a - 1) / 2 + 6LL * a + 9LL * a) << 2) | 2;
                contrib[cptr[b0 + lo] + slot] = code;
            }
        }
        __syncthreads();
    }
    if (!FILL)
        for (int i = lane; i < nb; i += 64) bcount[b0 + i] = cur[i];
}

__global__ void scalar_pattern_kernel(int32_t nb_cp, const int64_t *bptr, const int32_t *bcol,
                                      int64_t *rowptr, int32_t *colind) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I > nb_cp) return;
    if (I == nb_cp) {
        rowptr[3 * (int64_t)nb_cp] = 9 * bptr[nb_cp];
        return;
    }
    int64_t b0 = bptr[I];
    int nb = (int)(bptr[I + 1] - b0);
    for (int k = 0; k < 3; ++k) {
        int64_t r0 = 9 * b0 + 3LL * k * nb;
        rowptr[3 * (int64_t)I + k] = r0;
        for (int j = 0; j < nb; ++j) {
            int32_t J = bcol[b0 + j];
            colind[r0 + 3 * j + 0] = 3 * J;
            colind[r0 + 3 * j + 1] = 3 * J + 1;
            colind[r0 + 3 * j + 2] = 3 * J + 2;
        }
    }
}

// ---- values ------------------------------------------------------------------------------
__global__ void csr_values_kernel(int32_t nb_cp, const int64_t *bptr, const int64_t *cptr,
                                  const int64_t *contrib, const double *__restrict__ Kdata,
                                  double *__restrict__ vals) {
    // one thread per block; blocks of a row are consecutive threads
    int64_t blk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nblocks = bptr[nb_cp];
    if (blk >= nblocks) return;
    // row of this block: binary search in bptr
    int lo = 0, hi = nb_cp - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (bptr[mid] <= blk) lo = mid; else hi = mid - 1;
    }
    const int I = lo;
    const int64_t b0 = bptr[I];
    const int nb = (int)(bptr[I + 1] - b0);
    const int j = (int)(blk - b0);
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // v[k][l] = v[3k + l]
    for (int64_t c = cptr[blk]; c < cptr[blk + 1]; ++c) {
        int64_t code = contrib[c];
        int kind = (int)(code & 3);
        const double *s = Kdata + (code >> 2);
        if (kind == 0) {            // stored (k, l) at s[k + 3 l]
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) v[3 * k + l] += s[k + 3 * l];
        } else if (kind == 1) {     // transpose of the stored block
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) v[3 * k + l] += s[l + 3 * k];
        } else {                    // diagonal block: side + side^T
            const double d00 = s[0], d01 = s[1], d11 = s[2], d02 = s[3], d12 = s[4], d22 = s[5];
            v[0] += 2.0 * d00; v[4] += 2.0 * d11; v[8] += 2.0 * d22;
            v[1] += d01; v[3] += d01;
            v[2] += d02; v[6] += d02;
            v[5] += d12; v[7] += d12;
        }
    }
    const int64_t r0 = 9 * b0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) vals[r0 + 3LL * k * nb + 3 * j + l] = v[3 * k + l];
}

// ---- SpMV by nodal 3x3 blocks: one wave per CP row ---------------------------------------
// The three dof rows 3I, 3I+1, 3I+2 of a CP share their block columns; row 3I+k of block j is
// vals[9 b0 + 3 k nb + 3 j + l], l = 0..2 (b0 = bptr[I], nb blocks in the row). A lane takes
// blocks j = lane, lane + 64, ...: it reads the block column once (4 bytes per 9 entries instead
// of 4 per entry), x at 3J..3J+2, and 3 x 3 consecutive values, consecutive lanes on consecutive
// 24-byte runs. Rows of a masked-out dof give 0. MODE bits: 1 = |vals| |x| (the FP64 rounding
// bound of a product), 2 = block partials of x . y for the PCG.
template <int MODE>
__global__ void __launch_bounds__(256) block_spmv_kernel(int32_t nb_cp, const int64_t *__restrict__ bptr,
                                                         const int32_t *__restrict__ bcol,
                                                         const double *__restrict__ vals,
                                                         const double *__restrict__ x,
                                                         const uint8_t *__restrict__ mask,
                                                         double *__restrict__ y, double *__restrict__ partial,
                                                         const int32_t *__restrict__ pred) {
    if (pred && *pred == 0) return;   // predicated launch (device-side solver control)
    constexpr bool ABS = (MODE & 1) != 0, DOT = (MODE & 2) != 0;
    __shared__ double sh[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int64_t I = (int64_t)blockIdx.x * 4 + w; I < nb_cp; I += (int64_t)gridDim.x * 4) {
        const int64_t b0 = bptr[I];
        const int nb = (int)(bptr[I + 1] - b0);
        const double *v0 = vals + 9 * b0;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int j = lane; j < nb; j += 64) {
            const int64_t J = bcol[b0 + j];
            double x0 = x[3 * J], x1 = x[3 * J + 1], x2 = x[3 * J + 2];
            const double *a0 = v0 + 3 * j, *a1 = a0 + 3 * (int64_t)nb, *a2 = a1 + 3 * (int64_t)nb;
            double c[9] = {a0[0], a0[1], a0[2], a1[0], a1[1], a1[2], a2[0], a2[1], a2[2]};
            if (ABS) {
                x0 = fabs(x0); x1 = fabs(x1); x2 = fabs(x2);
#pragma unroll
                for (int t = 0; t < 9; ++t) c[t] = fabs(c[t]);
            }
            s0 += c[0] * x0 + c[1] * x1 + c[2] * x2;
            s1 += c[3] * x0 + c[4] * x1 + c[5] * x2;
            s2 += c[6] * x0 + c[7] * x1 + c[8] * x2;
        }
        for (int o = 32; o > 0; o >>= 1) {
            s0 += __shfl_xor(s0, o, 64);
            s1 += __shfl_xor(s1, o, 64);
            s2 += __shfl_xor(s2, o, 64);
        }
        if (lane == 0) {
            double r[3] = {s0, s1, s2};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (mask && !mask[3 * I + k]) r[k] = 0.0;
                y[3 * I + k] = r[k];
                if (DOT) acc += r[k] * x[3 * I + k];
            }
        }
    }
    if (DOT) {
        double v = lane == 0 ? acc : 0.0;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) sh[w] = v;
        __syncthreads();
        if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    }
}

constexpr int SPMV_MAX_BLOCKS = 4096;

// y = mask .* (A x) (mode 0), the same with |A| |x| (mode 1), or mode 2: y = mask .* (A x) and
// partial[b] = the grid block's share of x . y (grid of *nparts blocks, at most
// SPMV_MAX_BLOCKS). mask may be null.
int csr_block_spmv(const yeti_csr *c, const double *vals, const double *x, const uint8_t *mask, double *y,
                   double *partial, int mode, int *nparts, hipStream_t st, const int32_t *pred) {
    if (c->nb_cp <= 0) return 0;
    const unsigned g = (unsigned)std::min<int64_t>(SPMV_MAX_BLOCKS, ((int64_t)c->nb_cp + 3) / 4);
    if (nparts) *nparts = (int)g;
    switch (mode) {
    case 0: hipLaunchKernelGGL(block_spmv_kernel<0>, dim3(g), dim3(256), 0, st, c->nb_cp, c->bptr.p, c->bcol.p, vals, x,
                               mask, y, partial, pred); break;
    case 1: hipLaunchKernelGGL(block_spmv_kernel<1>, dim3(g), dim3(256), 0, st, c->nb_cp, c->bptr.p, c->bcol.p, vals, x,
                               mask, y, partial, pred); break;
    default: hipLaunchKernelGGL(block_spmv_kernel<2>, dim3(g), dim3(256), 0, st, c->nb_cp, c->bptr.p, c->bcol.p, vals,
                                x, mask, y, partial, pred); break;
    }
    YETI_HIP(hipGetLastError());
    return 0;
}

int64_t csr_nrows(const yeti_csr *c) { return 3 * (int64_t)c->nb_cp; }
const int64_t *csr_rowptr_dev(const yeti_csr *c) { return c->rowptr.p; }
const int32_t *csr_colind_dev(const yeti_csr *c) { return c->colind.p; }

}  // namespace yeti

using namespace yeti;

namespace yeti {
// ---- canonical form: scipy's K_side + K_side^T drops the entries that sum to exactly 0 ----
__global__ void __launch_bounds__(256) nz_count_kernel(int64_t nrows, const int64_t *rowptr, const double *vals,
                                                       int32_t *cnt) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int32_t n = 0;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) n += vals[k] != 0.0;
    cnt[r] = n;
}

__global__ void __launch_bounds__(256) nz_compact_kernel(int64_t nrows, const int64_t *rowptr, const int32_t *colind,
                                                         const double *vals, const int64_t *newptr,
                                                         int32_t *colind_out, double *vals_out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int64_t o = newptr[r];
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        const double v = vals[k];
        if (v != 0.0) {
            colind_out[o] = colind[k];
            vals_out[o] = v;
            ++o;
        }
    }
}

}  // namespace yeti

extern "C" {

int32_t yeti_csr_drop_zeros(const yeti_csr *c, const double *vals_dev, int64_t *rowptr_out, int32_t *colind_out,
                            double *vals_out, int64_t *nnz_out, void *stream) {
    if (!c || !vals_dev || !rowptr_out || !colind_out || !vals_out || !nnz_out)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nrows = 3 * (int64_t)c->nb_cp;
    DevBuf<int32_t> cnt;
    YETI_HIP(cnt.alloc(std::max<int64_t>(nrows, 1)));
    const unsigned g = (unsigned)((nrows + 255) / 256);
    if (nrows > 0) {
        hipLaunchKernelGGL(nz_count_kernel, dim3(g), dim3(256), 0, st, nrows, c->rowptr.p, vals_dev, cnt.p);
        YETI_HIP(hipGetLastError());
    }
    int rc = exclusive_scan<int32_t>(cnt.p, nrows, rowptr_out, st);   // ends with a stream sync
    if (rc) return rc;
    if (nrows > 0) {
        hipLaunchKernelGGL(nz_compact_kernel, dim3(g), dim3(256), 0, st, nrows, c->rowptr.p, c->colind.p, vals_dev,
                           (const int64_t *)rowptr_out, colind_out, vals_out);
        YETI_HIP(hipGetLastError());
    }
    YETI_HIP(hipMemcpyAsync(nnz_out, rowptr_out + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    YETI_HIP(hipStreamSynchronize(st));
    return 0;
}

int32_t yeti_csr_create(const yeti_asm *h, void *stream, yeti_csr **out) {
    if (!h || !out) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    hipStream_t st = (hipStream_t)stream;
    AsmView v;
    int rc = asm_view(h, &v);
    if (rc) return rc;
    if (v.npatch > 8) return set_error(YETI_ERR_INVALID_ARGUMENT, "CSR build supports up to 8 patches");
    auto *c = new yeti_csr();
    auto fail = [&](int code) {
        delete c;
        return code;
    };
    c->nb_cp = v.nb_cp;
    PatchIen pi{};
    for (int p = 0; p < v.npatch; ++p) {
        pi.ien[p] = v.ien[p];
        pi.elems[p] = v.elems[p];
        pi.offs[p] = v.offs[p];
        pi.nn[p] = v.nnode[p];
        if (v.nnode[p] > 729) return fail(set_error(YETI_ERR_UNSUPPORTED_DEGREE, "CSR: too many nodes per element"));
    }
    DevBuf<int32_t> cnt, cursor, rowlen, bcount, oflow;
    DevBuf<int64_t> inc_ptr, inc;
    if (cnt.alloc(v.nb_cp) != hipSuccess || cursor.alloc(v.nb_cp) != hipSuccess ||
        inc_ptr.alloc(v.nb_cp + 1) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: allocation failed"));
    if (hipMemsetAsync(cnt.p, 0, 4 * (size_t)v.nb_cp, st) != hipSuccess ||
        hipMemsetAsync(cursor.p, 0, 4 * (size_t)v.nb_cp, st) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: memset failed"));
    for (int p = 0; p < v.npatch; ++p) {
        int64_t n = (int64_t)v.n_active[p] * v.nnode[p];
        if (n == 0) continue;
        hipLaunchKernelGGL(inc_count_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, v.ien[p],
                           v.elems[p], v.n_active[p], v.nnode[p], cnt.p);
    }
    if ((rc = exclusive_scan<int32_t>(cnt.p, v.nb_cp, inc_ptr.p, st))) return fail(rc);
    int64_t ninc = 0;
    if (hipMemcpy(&ninc, inc_ptr.p + v.nb_cp, 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: copy failed"));
    if (inc.alloc(std::max<int64_t>(ninc, 1)) != hipSuccess) return fail(set_error(YETI_ERR_DEVICE, "CSR: alloc inc"));
    for (int p = 0; p < v.npatch; ++p) {
        int64_t n = (int64_t)v.n_active[p] * v.nnode[p];
        if (n == 0) continue;
        hipLaunchKernelGGL(inc_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, v.ien[p],
                           v.elems[p], v.n_active[p], v.nnode[p], p, inc_ptr.p, cursor.p, inc.p);
    }
    hipLaunchKernelGGL(inc_sort_kernel, dim3((v.nb_cp + 255) / 256), dim3(256), 0, st, v.nb_cp, inc_ptr.p, inc.p);
    // block pattern
    if (rowlen.alloc(v.nb_cp) != hipSuccess || c->bptr.alloc(v.nb_cp + 1) != hipSuccess ||
        oflow.alloc(1) != hipSuccess || hipMemsetAsync(oflow.p, 0, 4, st) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: alloc rows"));
    hipLaunchKernelGGL(row_pattern_kernel<false>, dim3((v.nb_cp + 3) / 4), dim3(256), 0, st, v.nb_cp, inc_ptr.p,
                       inc.p, pi, rowlen.p, nullptr, nullptr, oflow.p);
    int32_t of = 0;
    if (hipMemcpyAsync(&of, oflow.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: copy failed"));
    if (of) return fail(set_error(YETI_ERR_SIZE_OVERFLOW, "CSR: a control point has more than 1024 neighbours"));
    if ((rc = exclusive_scan<int32_t>(rowlen.p, v.nb_cp, c->bptr.p, st))) return fail(rc);
    if (hipMemcpy(&c->nblocks, c->bptr.p + v.nb_cp, 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: copy failed"));
    if (c->bcol.alloc(std::max<int64_t>(c->nblocks, 1)) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: alloc bcol"));
    hipLaunchKernelGGL(row_pattern_kernel<true>, dim3((v.nb_cp + 3) / 4), dim3(256), 0, st, v.nb_cp, inc_ptr.p,
                       inc.p, pi, nullptr, c->bptr.p, c->bcol.p, oflow.p);
    // contributions
    if (bcount.alloc(std::max<int64_t>(c->nblocks, 1)) != hipSuccess ||
        c->cptr.alloc(c->nblocks + 1) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: alloc contributions"));
    hipLaunchKernelGGL(contrib_kernel<false>, dim3(v.nb_cp), dim3(64), 0, st, v.nb_cp, inc_ptr.p, inc.p, pi,
                       c->bptr.p, c->bcol.p, bcount.p, nullptr, nullptr);
    if ((rc = exclusive_scan<int32_t>(bcount.p, c->nblocks, c->cptr.p, st))) return fail(rc);
    int64_t ncontrib = 0;
    if (hipMemcpy(&ncontrib, c->cptr.p + c->nblocks, 8, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: copy failed"));
    if (c->contrib.alloc(std::max<int64_t>(ncontrib, 1)) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: alloc contrib list (%lld entries)", (long long)ncontrib));
    hipLaunchKernelGGL(contrib_kernel<true>, dim3(v.nb_cp), dim3(64), 0, st, v.nb_cp, inc_ptr.p, inc.p, pi,
                       c->bptr.p, c->bcol.p, nullptr, c->cptr.p, c->contrib.p);
    // scalar pattern
    if (c->rowptr.alloc(3 * (int64_t)v.nb_cp + 1) != hipSuccess ||
        c->colind.alloc(std::max<int64_t>(9 * c->nblocks, 1)) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: alloc scalar pattern"));
    hipLaunchKernelGGL(scalar_pattern_kernel, dim3((v.nb_cp + 256) / 256), dim3(256), 0, st, v.nb_cp, c->bptr.p,
                       c->bcol.p, c->rowptr.p, c->colind.p);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "CSR: pattern kernels failed"));
    *out = c;
    return 0;
}

int64_t yeti_csr_nnz(const yeti_csr *c) { return c ? 9 * c->nblocks : -1; }
int32_t yeti_csr_nrows(const yeti_csr *c) { return c ? 3 * c->nb_cp : -1; }
const int64_t *yeti_csr_rowptr(const yeti_csr *c) { return c ? c->rowptr.p : nullptr; }
const int32_t *yeti_csr_colind(const yeti_csr *c) { return c ? c->colind.p : nullptr; }

int32_t yeti_csr_values(const yeti_csr *c, const double *Kdata_dev, double *vals_dev, void *stream) {
    if (!c || !Kdata_dev || !vals_dev) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    if (c->nblocks == 0) return 0;
    hipLaunchKernelGGL(csr_values_kernel, dim3((unsigned)((c->nblocks + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, c->nb_cp, c->bptr.p, c->cptr.p, c->contrib.p, Kdata_dev, vals_dev);
    YETI_HIP(hipGetLastError());
    return 0;
}

int32_t yeti_csr_spmv(const yeti_csr *c, const double *vals_dev, const double *x_dev, double *y_dev, void *stream) {
    if (!c || !vals_dev || !x_dev || !y_dev) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return csr_block_spmv(c, vals_dev, x_dev, nullptr, y_dev, nullptr, 0, nullptr, (hipStream_t)stream);
}

int32_t yeti_csr_spmv_abs(const yeti_csr *c, const double *vals_dev, const double *x_dev, double *y_dev, void *stream) {
    if (!c || !vals_dev || !x_dev || !y_dev) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return csr_block_spmv(c, vals_dev, x_dev, nullptr, y_dev, nullptr, 1, nullptr, (hipStream_t)stream);
}

int32_t yeti_csr_pattern_to_host(const yeti_csr *c, int64_t *rowptr_host, int32_t *colind_host) {
    if (!c || !rowptr_host || !colind_host) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    YETI_HIP(hipDeviceSynchronize());
    YETI_HIP(hipMemcpy(rowptr_host, c->rowptr.p, (3 * (size_t)c->nb_cp + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (c->nblocks > 0)
        YETI_HIP(hipMemcpy(colind_host, c->colind.p, 9 * (size_t)c->nblocks * sizeof(int32_t), hipMemcpyDeviceToHost));
    return 0;
}

void yeti_csr_destroy(yeti_csr *c) { delete c; }

}  // extern "C"
