// session.hpp -- device state shared by the solve sessions (mfsolve.hip) and the preconditioner
// kernels that end a session iteration (mf.hip): the scalar / flag slots and the one-launch grid
// reductions that update them.
#pragma once

#include "common.hpp"

namespace yeti {

// sc[] slots (doubles), in the order of yeti_ir_status
enum : int {
    S_RZ = 0, S_PQ, S_RR, S_RZN, S_BB, S_RRT, S_RRMAX, S_TOL2, S_RRT_PREV, S_LASTUPD, S_BETA, S_NSC = 16
};
// flags[] slots (int32), in the order of yeti_ir_status; F_COUNTER (not reported) is the
// arrival counter of the grid reductions; F_BEST / F_RESTORE (not reported) ask the step to save
// the iterate as the best one so far / to put the best one back; F_DEPARTED (not reported): at the
// last replacement the true residual stood far above the iterated one. F_UPDATE is 1 for a replacement
// the residual's fall (or the tolerance) called for, 2 for a forced one.
enum : int {
    F_ACTIVE = 0, F_UPDATE, F_STAGNANT, F_DONE, F_FAILED, F_ITERS, F_FLOOR, F_UPDATES, F_NFLAGS = 8,
    F_COUNTER = 8, F_BEST = 9, F_RESTORE = 10, F_DEPARTED = 11, F_ALLOC = 16
};

// Sum of v over a 256-thread block, returned in thread 0.
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

// Grid reduction inside the producing launch. Every block stores its partial (blk, valid in
// thread 0) in part[block]; the block that arrives last on the device-scope counter adds all
// partials in block order, exactly as a separate one-block pass over part[] would, and returns
// true with the total in thread 0. It leaves the counter at zero for the next reduction. The
// kernel's blocks must all call this (none may return before it).
__device__ inline bool last_block_total(double blk, double *part, int32_t *counter, double *sh, double &total) {
    __shared__ int32_t is_last;
    const int nblk = (int)(gridDim.x * gridDim.y);
    const int bid = (int)(blockIdx.x + gridDim.x * blockIdx.y);
    if (threadIdx.x == 0) {
        part[bid] = blk;
        __threadfence();
        is_last = __hip_atomic_fetch_add(counter, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == nblk - 1;
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double v = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x)
        v += __hip_atomic_load(part + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    total = block_sum256(v, sh);
    if (threadIdx.x == 0) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// What the last block of the preconditioner launch does with r.z (thread 0 only).
struct RzTail {
    double *sc = nullptr;
    int32_t *flags = nullptr;
    int32_t mode = 0;        // 0: first direction (rz = r.z); 1: end of an iteration
    int32_t max_iters = 0;
};

__device__ inline void rz_tail(const RzTail &t, double rz) {
    t.sc[S_RZN] = rz;
    if (t.mode == 0) {
        t.sc[S_RZ] = rz;
        return;
    }
    // beta for p = z + beta p, roll rz, count the iteration, cap
    t.sc[S_BETA] = rz / t.sc[S_RZ];
    t.sc[S_RZ] = rz;
    t.flags[F_ITERS] += 1;
    if (t.flags[F_ITERS] >= t.max_iters) {
        t.flags[F_DONE] = 1;
        t.flags[F_FAILED] = 1;
        t.flags[F_ACTIVE] = 0;
    }
}

}  // namespace yeti
