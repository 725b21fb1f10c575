// batch.hpp -- types shared by the matrix-free operator (mf.hip) and the solve sessions
// (mfsolve.hip): the patch and preconditioner descriptors the kernels take, and the device
// tables of a batch of sessions (many solves advanced by the same launches).
#pragma once

#include "session.hpp"

namespace yeti {

struct MfPatch {
    int32_t N[3];            // elements per direction
    int32_t ncp[3];          // control points per direction
    const int32_t *cpmap;    // patch grid CP -> global CP (0-based)
    const double *Btab;      // [d][e_d][NG][NP1] values (basis function span - p + r)
    const double *dBtab;     // same, derivatives
    int32_t boff[3];         // offset of direction d in the tables (in elements)
    int32_t gnb[2][3];       // brick grids of the sweep apply, [0] FP64, [1] FP32; each orders that
                             // precision's stored geometry
    double lam, mu;
};

// Fast diagonalisation of one patch: free index boxes per displacement component and the
// offsets of the 1D eigenvector matrices and inverse eigenvalue sums.
struct FdmArgs {
    int32_t n[3];            // patch CP grid
    int32_t lo[3][3];        // [component][direction] first free index
    int32_t m[3][3];         // [component][direction] free count
    int64_t tot[3];          // m0 m1 m2 per component
    int64_t maxtot;
    int64_t uoff[3][3];      // offset of U_cd (m x m, row-major) in the matrix buffer
    int64_t doff[3];         // offset of 1/D_c in the scaling buffer
    const int32_t *cpmap;
};

// One mode product of the preconditioner, out(.., a, ..) = sum_b A[a][b] in(.., b, ..) along one
// axis for the three components, written as a batched strided GEMM C = L R: along axis 0,
// L = A (m0 x m0) and R = the tensor seen as m0 x (m1 m2); along axis 1, per k, L = the tensor
// slice (m0 x m1) and R = A^T; along axis 2, L = the tensor seen as (m0 m1) x m2 and R = A^T.
// Element (row, col) of operand X of component c, batch z, is at X[off[c] + z bs[c] + row rs[c]
// + col cs[c]].
struct FdmGemm {
    int32_t M[3], N[3], K[3], nbatch[3];
    int64_t loff[3], lrs[3], lcs[3], lbs[3];
    int64_t roff[3], rrs[3], rcs[3], rbs[3];
    int64_t coff[3], crs[3], ccs[3], cbs[3];
    int32_t gather;   // R is read from the DOF vector r through the CP map (first forward product)
    int32_t scale;    // C is multiplied by 1 / D_c (last forward product)
};

// ---- batches ------------------------------------------------------------------------------
// A batch advances many single-patch solve sessions of one precision with the same launches:
// every launch walks a device list of tasks, each naming a member and a piece of its work (a
// brick of elements, a block of control points, a chunk of a vector, a tile of a preconditioner
// product). A member's reductions are split into chunks that depend only on its own size and
// are summed in chunk order by the last chunk to finish, so its iterates do not depend on which
// other sessions share the batch.
constexpr int BATCH_MAX_MEMBERS = 256;
constexpr int BATCH_RED_MAX = 1024;          // reduction chunks per member (the session's partials)
constexpr int BATCH_CHUNK_MIN = 2048;        // vector entries per reduction chunk, at least
constexpr int BATCH_GATHER_BLOCK = 256;      // control points per gather task

struct BReq {
    // operator
    MfPatch mp;
    int32_t p, ng;
    int32_t nbrick[2][3];    // apply brick grids (the sweep's), [0] FP64, [1] FP32
    int32_t br[2][3], bn[2][3];  // apply brick edges per direction in elements / nodes, same index
    int32_t ncp;
    int64_t n;               // 3 ncp
    const void *geo[2];      // quadrature geometry, [0] FP64, [1] FP32
    const void *w[2];        // control-point weights
    void *part[2];           // brick partial sums
    // preconditioner, in the session's precision: the six mode products' operands
    FdmArgs fa;
    FdmGemm g[6];
    const void *L[6];
    const void *R[6];
    void *C[6];
    const void *invD;
    const void *wout;        // result of the last product (read by the scatter)
    // session state
    const uint8_t *mask;
    const double *b;
    double *x, *xbest, *y64;
    void *vr, *vy, *vz, *vp, *vq;  // iteration vectors r, y, z, p, q (session precision)
    double *sc;
    int32_t *flags;
    double *red;             // reduction partials (BATCH_RED_MAX)
    int32_t vchunk, nvchunk; // entries per vector chunk (of n) and their number
    int32_t cchunk, ncchunk; // control points per scatter chunk (of ncp) and their number
    double rtol2, delta2;
    int32_t max_iters, max_updates;
};

struct BTask {
    int32_t req;             // member index
    int32_t a, b, c;         // task fields (brick / block / chunk id, or component, batch, tile)
};

enum : int {
    BT_BRICK1 = 0, BT_BRICK2, BT_BRICK3, BT_BRICK4,   // apply bricks per degree 1..4 (batch precision)
    BT_RBRICK1, BT_RBRICK2, BT_RBRICK3, BT_RBRICK4,   // FP64 bricks of an FP32 batch (replacement applies)
    BT_GATHER, BT_VEC, BT_CP, BT_GEMM0, BT_NLISTS = BT_GEMM0 + 6
};

struct BatchDev {
    const BReq *req;
    const BTask *list[BT_NLISTS];
    const int32_t *count;    // [BT_NLISTS] (device, read by the kernels)
};

// A member's partial sums over a task list are combined by the member's last task to finish:
// task i of the member stores its partial in part[i], the member's arrival counter decides the
// last one, which adds part[0 .. ntask) in order (the same sum whatever other members share the
// launch). Leaves the counter at zero. Every thread of the workgroup calls it.
__device__ inline bool member_total(double blk, double *part, int32_t *counter, int ntask, int task, double *sh,
                                    double &total) {
    __shared__ int32_t is_last;
    if (threadIdx.x == 0) {
        part[task] = blk;
        __threadfence();
        is_last = __hip_atomic_fetch_add(counter, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == ntask - 1;
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double v = 0.0;
    for (int i = threadIdx.x; i < ntask; i += blockDim.x)
        v += __hip_atomic_load(part + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    total = block_sum256(v, sh);
    if (threadIdx.x == 0) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

// mf.hip: the batched operator (apply through the bricks and gather) and preconditioner.
// rpl: the FP64 apply of the residual replacement (x -> y64, members with an update due);
// otherwise the iteration apply in the batch precision (p -> q, active members).
int batch_mf_apply(const BatchDev &bd, int prec64, bool rpl, hipStream_t st);
int batch_fdm_apply(const BatchDev &bd, int prec64, int mode, hipStream_t st);
// fills the operator and preconditioner fields of a member; 0 or an error code
int batch_fill_operator(const yeti_mf *m, const yeti_fdm *f, int prec64, BReq *out);

}  // namespace yeti
