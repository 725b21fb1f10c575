// elem_kernel.hpp -- the FP64 element stiffness kernels (included by assembly.hip, which defines
// GaussRule and bspline_ders before including it).
//
// 1. elem_geo_kernel, one thread per (element, Gauss point): the Jacobian of the NURBS map
//    (shap.f90:94-178) reduced to 12 numbers per point, A[m][c] = sqrt(w |det J|) / W
//    dxi_m/dx_c and dW/dxi_m / W, written to HBM (12 doubles per Gauss point).
// 2. elem_stiffness_kernel, persistent and warp-specialised: one workgroup of 12 waves per CU
//    walks its share of the elements; each element is cut into steps (pass, Gauss-point chunk).
//    In step s
//      * 4 producer waves build step s+1: for every node a and Gauss point q of the chunk,
//        G_c[a][q] = sqrt(w|det J|) dR_a/dx_c = w_a sum_m (dN_a/dxi_m - N_a dW_m/W) A[m][c]
//        into one half of a double-buffered LDS array;
//      * 8 consumer waves accumulate S^{kl}_{ab} = sum_q G_k[a][q] G_l[b][q] over the chunk on
//        v_mfma_f64_16x16x4f64. A consumer task is (16-node tile pair (I, J), k): the three
//        16x16 tiles S^{k0}, S^{k1}, S^{k2} of rows I, columns J; a wave keeps 3-4 tasks and
//        loads each distinct operand fragment once per K step (about 0.8 fragment per MFMA);
//    one __syncthreads() ends each step. After an element's last chunk of a pass the consumers
//    write the raw S tiles of the pass's columns into an LDS stage; all 12 waves then emit
//    K_ab[k][l] = lam S^{kl}_ab + mu S^{lk}_ab + mu d_kl tr S_ab (the isotropic form of
//    stiffmatrix_byCP's BoI^T D BoJ, stiffmatrix.F90:175-216) with coalesced stores, the row
//    and column indices computed from the COO position.
#pragma once

#include <utility>

typedef double f64x2 __attribute__((ext_vector_type(2)));

// Compile-time loop: f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N-1>{}),
// so schedule tables indexed by the loop variable fold to constants.
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr int ELEM_NGEO = 12;   // per Gauss point: A[m][c] (9), dW/dxi_m / W (3)

template <int P, int NG>
struct ElemCfg {
    static constexpr int NP1 = P + 1;
    static constexpr int NN = NP1 * NP1 * NP1;          // nodes (CPs) per element
    static constexpr int NT = (NN + 15) / 16;           // 16-node tiles
    static constexpr int NNP = NT * 16;                 // padded node count
    static constexpr int NQ = NG * NG * NG;             // Gauss points per element
    static constexpr int NQ8 = (NQ + 7) / 8 * 8;        // padded: two MFMA K steps per LDS read
    static constexpr int QCW = NNP >= 128 ? 8 : 16;     // Gauss points per chunk (LDS budget)
    static constexpr int QC = NQ8 < QCW ? NQ8 : QCW;
    static constexpr int NCH = (NQ8 + QC - 1) / QC;
    static constexpr int QCP = QC + 2;                  // padded G row (doubles)
    static constexpr int NPAIR = NT * (NT + 1) / 2;     // node-tile pairs I <= J
    static constexpr int CW = 8;                        // consumer waves
    static constexpr int PW = 4;                        // producer waves
    static constexpr int NTH = (CW + PW) * 64;
    static constexpr int MAXT = NNP >= 128 ? 3 : 4;     // tasks per consumer wave per pass (registers)
    static constexpr int MAXPASS = 8;
    static constexpr int MAXR = 8;                      // staging rounds per pass
    static constexpr int GPW = QC / PW;                 // Gauss points per producer wave per chunk
    static constexpr int64_t NDE = 9LL * NN * (NN - 1) / 2 + 6LL * NN;   // COO entries per element
    // LDS stage of raw S values (9 per node pair a <= b): what is left of ~150 KiB after G
    static constexpr int64_t SSTN = 9LL * NN * (NN + 1) / 2;
    static constexpr int64_t GBYTES = 2LL * 3 * NNP * QCP * 8;
    static constexpr int64_t STGMAX = (150 * 1024 - GBYTES - 4 * 1024) / 8;
    static constexpr int STG = (int)(SSTN < STGMAX ? SSTN : STGMAX);
    static_assert(QC % PW == 0, "chunk must split over producer waves");
};

__host__ __device__ constexpr int64_t coo_start(int b) { return (9LL * b * b + 3LL * b) / 2; }   // COO
__host__ __device__ constexpr int64_t sst_start(int b) { return 9LL * b * (b + 1) / 2; }          // stage

// Static schedule of consumer tasks (tile pair, k) over waves and passes. A pass takes whole tile
// columns J (all pairs I <= J), so the S values of nodes b in [16 J0, 16 J1) are complete after
// it; its tasks are dealt to the waves in pair-major order in contiguous runs (sizes differing by
// at most one), so a wave's tasks share operand fragments. Each pass's b-range is emitted in
// rounds whose raw S values fit the LDS stage.
template <int NT, int WAVES, int MAXT, int MAXPASS, int MAXR>
struct TaskSchedule {
    int npass;
    int nt[MAXPASS][WAVES];                 // tasks of a wave
    int task[MAXPASS][WAVES][MAXT];         // pair * 3 + k
    int na[MAXPASS][WAVES], nb[MAXPASS][WAVES];
    int afr[MAXPASS][WAVES][MAXT];          // distinct A fragments: I * 3 + k
    int bfr[MAXPASS][WAVES][3 * MAXT];      // distinct B fragments: J * 3 + l
    int aidx[MAXPASS][WAVES][MAXT];         // task -> A fragment slot
    int bidx[MAXPASS][WAVES][MAXT][3];      // (task, l) -> B fragment slot
    int nround[MAXPASS];
    int rb[MAXPASS][MAXR + 1];              // round boundaries in b
    bool ok;
};

__host__ __device__ constexpr int pair_J(int pi) {
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= pi) ++j;
    return j;
}
__host__ __device__ constexpr int pair_I(int pi) { return pi - pair_J(pi) * (pair_J(pi) + 1) / 2; }

template <int NT, int WAVES, int MAXT, int MAXPASS, int MAXR>
constexpr TaskSchedule<NT, WAVES, MAXT, MAXPASS, MAXR> make_schedule(int nn, int stg) {
    TaskSchedule<NT, WAVES, MAXT, MAXPASS, MAXR> s{};
    s.ok = true;
    int J = 0, pass = 0;
    while (J < NT && pass < MAXPASS) {
        int J1 = J, ntask = 0;
        while (J1 < NT && ntask + 3 * (J1 + 1) <= MAXT * WAVES) {
            ntask += 3 * (J1 + 1);
            ++J1;
        }
        if (J1 == J) {
            s.ok = false;
            return s;
        }
        const int t0 = 3 * (J * (J + 1) / 2);
        int t = t0;
        for (int w = 0; w < WAVES; ++w) {
            const int cnt = ntask / WAVES + (w < ntask % WAVES ? 1 : 0);
            s.nt[pass][w] = cnt;
            s.na[pass][w] = s.nb[pass][w] = 0;
            for (int u = 0; u < cnt; ++u, ++t) {
                s.task[pass][w][u] = t;
                const int pi = t / 3, k = t % 3;
                const int akey = pair_I(pi) * 3 + k;
                int ai = -1;
                for (int x = 0; x < s.na[pass][w]; ++x)
                    if (s.afr[pass][w][x] == akey) ai = x;
                if (ai < 0) {
                    ai = s.na[pass][w]++;
                    s.afr[pass][w][ai] = akey;
                }
                s.aidx[pass][w][u] = ai;
                for (int l = 0; l < 3; ++l) {
                    const int bkey = pair_J(pi) * 3 + l;
                    int bi = -1;
                    for (int x = 0; x < s.nb[pass][w]; ++x)
                        if (s.bfr[pass][w][x] == bkey) bi = x;
                    if (bi < 0) {
                        bi = s.nb[pass][w]++;
                        s.bfr[pass][w][bi] = bkey;
                    }
                    s.bidx[pass][w][u][l] = bi;
                }
            }
        }
        // rounds over b in [16 J, min(16 J1, nn))
        int b0 = 16 * J, bend = 16 * J1 < nn ? 16 * J1 : nn, r = 0;
        s.rb[pass][0] = b0;
        while (b0 < bend) {
            int b1 = b0 + 1;
            while (b1 < bend && sst_start(b1 + 1) - sst_start(b0) <= stg) ++b1;
            if (sst_start(b1) - sst_start(b0) > stg || r >= MAXR) {
                s.ok = false;
                return s;
            }
            s.rb[pass][++r] = b1;
            b0 = b1;
        }
        s.nround[pass] = r;
        J = J1;
        ++pass;
    }
    s.npass = pass;
    if (J < NT) s.ok = false;
    return s;
}

template <int P, int NG>
struct Sched {
    using C = ElemCfg<P, NG>;
    static constexpr auto S = make_schedule<C::NT, C::CW, C::MAXT, C::MAXPASS, C::MAXR>(C::NN, C::STG);
    static constexpr int NPASS = S.npass;
    static_assert(S.ok, "task schedule does not fit the wave / staging capacities");
};

struct PatchArgs {
    const double *coords;     // (3, nb_cp)
    const int32_t *ien;       // (nnode, nelem_patch), 1-based
    const double *weight;     // (nnode, nelem_patch)
    const int32_t *nijk;      // (3, nelem_patch), 1-based spans
    const double *U0, *U1, *U2;  // knot vectors
    const int32_t *elems;     // active elements of this launch (patch-local 0-based)
    const int64_t *offs;      // COO offset per active element
    int32_t n_active;
    double lam, mu;
    GaussRule g;
};

// ---------------------------------------------------------------------------------------
// Geometry pass: per active element gi and Gauss point q, geo[(gi * 12 + v) * NQ + q] with
// v = 3 m + c: A[m][c] = sqrt(w |det J|) / W dxi_m/dx_c; v = 9 + m: dW/dxi_m / W. J is the
// derivative with respect to the knot parameters (shap.f90:133-153); w |det J| includes the
// parent-to-parameter factors (shap.f90:156-178) and the Gauss weights (UELMAT.f90:118-119).
template <int P, int NG>
__global__ void __launch_bounds__(256) elem_geo_kernel(PatchArgs pa, double *__restrict__ geo) {
    using C = ElemCfg<P, NG>;
    constexpr int NP1 = C::NP1, NN = C::NN, NQ = C::NQ;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)pa.n_active * NQ) return;
    const int64_t gi = t / NQ;
    const int q = (int)(t % NQ);
    const int e = pa.elems[gi];
    const int gq[3] = {q % NG, (q / NG) % NG, q / (NG * NG)};
    double B[3][NP1], dB[3][NP1], half[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double *U = d == 0 ? pa.U0 : d == 1 ? pa.U1 : pa.U2;
        const int span1 = pa.nijk[3 * (int64_t)e + d];
        const double u0 = U[span1 - 1], u1 = U[span1];
        const double u = ((u1 - u0) * pa.g.x[gq[d]] + (u1 + u0)) * 0.5;
        double N[NP1], dN[NP1];
        bspline_ders<P>(span1 - 1, u, U, N, dN);
#pragma unroll
        for (int i = 0; i < NP1; ++i) {   // local index i <-> basis function P - i
            B[d][i] = N[P - i];
            dB[d][i] = dN[P - i];
        }
        half[d] = 0.5 * (u1 - u0);
    }
    const int32_t *ien = pa.ien + (int64_t)NN * e;
    const double *wt = pa.weight + (int64_t)NN * e;
    double s[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) s[m] = 0.0;
    for (int a = 0; a < NN; ++a) {
        const int i = a % NP1, j = (a / NP1) % NP1, k = a / (NP1 * NP1);
        const double w = wt[a];
        const double *xp = pa.coords + 3 * (int64_t)(ien[a] - 1);
        const double x0 = xp[0], x1 = xp[1], x2 = xp[2];
        const double b12 = B[1][j] * B[2][k] * w;
        const double Na = B[0][i] * b12;
        const double d0 = dB[0][i] * b12;
        const double d1 = B[0][i] * dB[1][j] * B[2][k] * w;
        const double d2 = B[0][i] * B[1][j] * dB[2][k] * w;
        s[0] += Na;
        s[1] += d0; s[2] += d1; s[3] += d2;
        s[4] += x0 * Na; s[5] += x1 * Na; s[6] += x2 * Na;
        s[7] += x0 * d0; s[8] += x0 * d1; s[9] += x0 * d2;
        s[10] += x1 * d0; s[11] += x1 * d1; s[12] += x1 * d2;
        s[13] += x2 * d0; s[14] += x2 * d1; s[15] += x2 * d2;
    }
    const double Wi = 1.0 / s[0];
    const double dw[3] = {s[1] * Wi, s[2] * Wi, s[3] * Wi};
    double J[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[r][c] = (s[7 + 3 * r + c] - s[4 + r] * dw[c]) * Wi;
    const double m00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double m01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double m02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * m00 + J[0][1] * m01 + J[0][2] * m02;
    const double om = pa.g.w[gq[0]] * pa.g.w[gq[1]] * pa.g.w[gq[2]] * fabs(det * half[0] * half[1] * half[2]);
    const double f = sqrt(om) * Wi / det;
    double A[9];   // A[3 m + c] = f * (adj J)[m][c]
    A[0] = m00 * f;
    A[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * f;
    A[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * f;
    A[3] = m01 * f;
    A[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * f;
    A[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * f;
    A[6] = m02 * f;
    A[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * f;
    A[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * f;
    double *out = geo + gi * (ELEM_NGEO * NQ) + q;
#pragma unroll
    for (int v = 0; v < 9; ++v) out[v * NQ] = A[v];
#pragma unroll
    for (int m = 0; m < 3; ++m) out[(9 + m) * NQ] = dw[m];
}

__device__ __forceinline__ double bcast(double v, int src) {
    // wave-uniform source lane -> scalar broadcast
    int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------------------
template <int P, int NG>
struct ElemSmem {
    using C = ElemCfg<P, NG>;
    // G_c[a][q] = sqrt(w |det J|) dR_a/dx_c at the chunk's Gauss points, double buffered
    __attribute__((aligned(16))) double G[2][3][C::NNP][C::QCP];
    double B[C::PW][3][NG][C::NP1];     // per producer wave: B(reversed local index) at the abscissae
    double dB[C::PW][3][NG][C::NP1];
    double stage[C::STG];               // raw S of one round: stage[sst(b) + 9 a + k + 3 l] = S^{kl}_{ab}
    int32_t dof[C::NN];                 // 3 (IEN - 1) of the element being emitted
};

// Producer: per-element set-up (1D bases in the wave's own LDS table, node weights in registers).
template <int P, int NG>
struct ProducerElem {
    using C = ElemCfg<P, NG>;
    double wb[(C::NNP + 63) / 64];   // weight of the G-row nodes a = lane + 64 m
};

template <int P, int NG>
__device__ __forceinline__ void producer_load_element(ElemSmem<P, NG> &sm, const PatchArgs &pa, int e, int pw,
                                                      ProducerElem<P, NG> &pe) {
    using C = ElemCfg<P, NG>;
    constexpr int NP1 = C::NP1, NN = C::NN;
    const int lane = threadIdx.x & 63;
    const double *wt = pa.weight + (int64_t)NN * e;
#pragma unroll
    for (int m = 0; m < (C::NNP + 63) / 64; ++m) {
        const int a = lane + 64 * m;
        pe.wb[m] = a < NN ? wt[a] : 0.0;
    }
    if (lane < 3 * NG) {
        const int d = lane / NG, g = lane % NG;
        const double *U = d == 0 ? pa.U0 : d == 1 ? pa.U1 : pa.U2;
        const int span1 = pa.nijk[3 * (int64_t)e + d];
        const double u0 = U[span1 - 1], u1 = U[span1];
        const double u = ((u1 - u0) * pa.g.x[g] + (u1 + u0)) * 0.5;
        double N[NP1], dN[NP1];
        bspline_ders<P>(span1 - 1, u, U, N, dN);
#pragma unroll
        for (int i = 0; i < NP1; ++i) {   // local index i <-> basis function P - i
            sm.B[pw][d][g][i] = N[P - i];
            sm.dB[pw][d][g][i] = dN[P - i];
        }
    }
    // the table is read back by this wave only: order the LDS writes before the reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Producer: the geometry of this wave's Gauss points of a chunk, one value per lane (lane
// 12 t + v holds value v of Gauss point q0 + qw + t), loaded one step before it is used.
template <int P, int NG>
__device__ __forceinline__ double producer_prefetch(const double *__restrict__ geo_e, int pw, int q0) {
    using C = ElemCfg<P, NG>;
    const int lane = threadIdx.x & 63;
    const int t = lane / ELEM_NGEO, v = lane - ELEM_NGEO * t;
    const int q = q0 + pw * C::GPW + t;
    return (t < C::GPW && q < C::NQ) ? geo_e[v * C::NQ + q] : 0.0;
}

// Producer: Gauss points [q0 + qw, q0 + qw + GPW) of the chunk into buffer `buf`; lane = node.
template <int P, int NG>
__device__ __forceinline__ void producer_chunk(ElemSmem<P, NG> &sm, double gv, int pw, int buf, int q0,
                                               const ProducerElem<P, NG> &pe) {
    using C = ElemCfg<P, NG>;
    constexpr int NP1 = C::NP1, NN = C::NN, NNP = C::NNP, NQ = C::NQ;
    static_assert(ELEM_NGEO * C::GPW <= 64, "geometry of a wave's Gauss points must fit a wave");
    const int lane = threadIdx.x & 63;
    const int qw = pw * C::GPW;
#pragma unroll
    for (int t = 0; t < C::GPW; ++t) {
        const int qt = q0 + qw + t;   // wave-uniform
        double A[9], dw[3];
#pragma unroll
        for (int v = 0; v < 9; ++v) A[v] = bcast(gv, ELEM_NGEO * t + v);
#pragma unroll
        for (int m = 0; m < 3; ++m) dw[m] = bcast(gv, ELEM_NGEO * t + 9 + m);
        const int h0 = qt % NG, h1 = (qt / NG) % NG, h2 = (qt / (NG * NG)) % NG;
#pragma unroll
        for (int m = 0; m < (NNP + 63) / 64; ++m) {
            const int a = lane + 64 * m;
            double g0v = 0.0, g1v = 0.0, g2v = 0.0;
            if (a < NN && qt < NQ) {
                const int i = a % NP1, j = (a / NP1) % NP1, k = a / (NP1 * NP1);
                const double b0 = sm.B[pw][0][h0][i], b1 = sm.B[pw][1][h1][j], b2 = sm.B[pw][2][h2][k];
                const double wa = pe.wb[m];
                const double b12 = b1 * b2;
                const double N = b0 * b12;
                const double t0 = (sm.dB[pw][0][h0][i] * b12 - N * dw[0]) * wa;
                const double t1 = (b0 * sm.dB[pw][1][h1][j] * b2 - N * dw[1]) * wa;
                const double t2 = (b0 * b1 * sm.dB[pw][2][h2][k] - N * dw[2]) * wa;
                g0v = t0 * A[0] + t1 * A[3] + t2 * A[6];
                g1v = t0 * A[1] + t1 * A[4] + t2 * A[7];
                g2v = t0 * A[2] + t1 * A[5] + t2 * A[8];
            }
            if (a < NNP) {
                sm.G[buf][0][a][qw + t] = g0v;
                sm.G[buf][1][a][qw + t] = g1v;
                sm.G[buf][2][a][qw + t] = g2v;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Consumer: MFMA over one chunk for consumer wave W of pass PASS. K step 2t uses Gauss points
// 8t + 2g, step 2t+1 uses 8t + 2g + 1 (g = lane K group): the order of the K sum is free as long
// as A and B use the same permutation. Each distinct fragment is read once per pair of K steps.
template <int P, int NG>
struct ConsumerAcc {
    using C = ElemCfg<P, NG>;
    f64x4 S[C::MAXT][3];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int u = 0; u < C::MAXT; ++u)
#pragma unroll
            for (int l = 0; l < 3; ++l) S[u][l] = f64x4{0.0, 0.0, 0.0, 0.0};
    }
};

template <int P, int NG, int PASS, int W>
__device__ __forceinline__ void mfma_chunk(const ElemSmem<P, NG> &sm, int buf, ConsumerAcc<P, NG> &acc) {
    using C = ElemCfg<P, NG>;
    constexpr int nt = Sched<P, NG>::S.nt[PASS][W];
    constexpr int na = Sched<P, NG>::S.na[PASS][W], nb = Sched<P, NG>::S.nb[PASS][W];
    if constexpr (nt > 0) {
        const int lane = threadIdx.x & 63;
        const int nl = lane & 15, kg = lane >> 4;
#pragma unroll 1
        for (int t = 0; t < C::QC / 8; ++t) {
            const int qo = 8 * t + 2 * kg;
            f64x2 Af[na], Bf[nb];
            static_for<na>([&](auto ic) {
                constexpr int x = decltype(ic)::value;
                constexpr int key = Sched<P, NG>::S.afr[PASS][W][x];
                Af[x] = *reinterpret_cast<const f64x2 *>(&sm.G[buf][key % 3][16 * (key / 3) + nl][qo]);
            });
            static_for<nb>([&](auto ic) {
                constexpr int x = decltype(ic)::value;
                constexpr int key = Sched<P, NG>::S.bfr[PASS][W][x];
                Bf[x] = *reinterpret_cast<const f64x2 *>(&sm.G[buf][key % 3][16 * (key / 3) + nl][qo]);
            });
            static_for<nt>([&](auto ic) {
                constexpr int u = decltype(ic)::value;
                constexpr int ai = Sched<P, NG>::S.aidx[PASS][W][u];
                static_for<3>([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    constexpr int bi = Sched<P, NG>::S.bidx[PASS][W][u][l];
                    acc.S[u][l] = __builtin_amdgcn_mfma_f64_16x16x4f64(Af[ai].x, Bf[bi].x, acc.S[u][l], 0, 0, 0);
                });
            });
            static_for<nt>([&](auto ic) {
                constexpr int u = decltype(ic)::value;
                constexpr int ai = Sched<P, NG>::S.aidx[PASS][W][u];
                static_for<3>([&](auto lc) {
                    constexpr int l = decltype(lc)::value;
                    constexpr int bi = Sched<P, NG>::S.bidx[PASS][W][u][l];
                    acc.S[u][l] = __builtin_amdgcn_mfma_f64_16x16x4f64(Af[ai].y, Bf[bi].y, acc.S[u][l], 0, 0, 0);
                });
            });
        }
    }
}

// Epilogue, first half: the raw S tiles of wave W whose column node b lies in round R of pass
// PASS, into the stage. Accumulator lane layout (v_mfma_f64_16x16x4f64): register r of lane l
// holds row (l >> 4) + 4 r and column l & 15 of the tile; rows are nodes a of tile I, columns
// nodes b of tile J.
template <int P, int NG, int PASS, int W, int R>
__device__ __forceinline__ void stage_round(ElemSmem<P, NG> &sm, const ConsumerAcc<P, NG> &acc) {
    constexpr auto &S = Sched<P, NG>::S;
    constexpr int nt = S.nt[PASS][W];
    constexpr int b0 = S.rb[PASS][R], b1 = S.rb[PASS][R + 1];
    constexpr int s0 = (int)sst_start(b0);
    (void)S;
    const int lane = threadIdx.x & 63;
    const int nl = lane & 15, kg = lane >> 4;
    static_for<nt>([&](auto ic) {
        constexpr int u = decltype(ic)::value;
        constexpr int code = Sched<P, NG>::S.task[PASS][W][u];
        constexpr int pi = code / 3, k = code % 3;
        constexpr int I = pair_I(pi), J = pair_J(pi);
        if constexpr (16 * J + 15 >= b0 && 16 * J < b1) {
            const int b = 16 * J + nl;
            if (b >= b0 && b < b1) {
                double *st = sm.stage + ((int)sst_start(b) - s0) + k;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 16 * I + kg + 4 * r;
                    if (I == J && a > b) continue;
#pragma unroll
                    for (int l = 0; l < 3; ++l) st[9 * a + 3 * l] = acc.S[u][l][r];
                }
            }
        }
    });
}

// Epilogue, second half (all 768 threads): the COO entries of round R of pass PASS. Node b's
// entries are the contiguous range [start(b), start(b+1)), start(b) = 9 b (b-1)/2 + 6 b: first
// the 9-entry blocks a < b with kk fastest, then the 6-entry diagonal block (0,0) (0,1) (1,1)
// (0,2) (1,2) (2,2) with the diagonal halved (build_system_elemStorage.f90:260-293). Wave w
// takes rows b = b0 + w, b0 + w + 12, ...: its lanes first contract the row's blocks in place
// (lane = block a), then copy the row with consecutive lanes on consecutive entries; entry
// j < 9 b of the row sits at stage offset 9 a + kk + 3 ll = j. row = 3 (IEN_a - 1) + kk,
// col = 3 (IEN_b - 1) + ll.
template <int P, int NG, bool WRITE_IDX, int PASS, int R>
__device__ __forceinline__ void copy_round(ElemSmem<P, NG> &sm, double lam, double mu, int64_t base,
                                           double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                           int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    constexpr auto &S = Sched<P, NG>::S;
    constexpr int b0 = S.rb[PASS][R], b1 = S.rb[PASS][R + 1];
    constexpr int s0 = (int)sst_start(b0);
    constexpr int NW = C::NTH / 64;
    (void)S;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    for (int b = b0 + wave; b < b1; b += NW) {
        double *st_b = sm.stage + ((int)sst_start(b) - s0);
        // K_ab[kk][ll] = lam S^{kk ll} + mu S^{ll kk} + mu d_kk,ll tr S, in place
        for (int a = lane; a <= b; a += 64) {
            double *s9 = st_b + 9 * a;
            double Sv[9], Kv[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) Sv[i] = s9[i];
            const double tr = mu * (Sv[0] + Sv[4] + Sv[8]);
            const double half = a == b ? 0.5 : 1.0;
#pragma unroll
            for (int ll = 0; ll < 3; ++ll)
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    if (kk == ll) Kv[kk + 3 * ll] = ((lam + mu) * Sv[kk + 3 * ll] + tr) * half;
                    else Kv[kk + 3 * ll] = lam * Sv[kk + 3 * ll] + mu * Sv[ll + 3 * kk];
                }
#pragma unroll
            for (int i = 0; i < 9; ++i) s9[i] = Kv[i];
        }
        // the copy reads values other lanes of this wave wrote
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int len = 9 * b + 6;
        const int64_t gb = base + coo_start(b);
        const int32_t dofb = WRITE_IDX ? sm.dof[b] : 0;
        // four entries per lane per trip, their LDS reads issued together
        for (int j0 = 0; j0 < len; j0 += 256) {
            double v[4];
            int pos[4];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int j = min(j0 + 64 * h + lane, len - 1);
                const int d = j - 9 * b;
                pos[h] = d < 0 ? j : 9 * b + ((0x210100 >> (4 * d)) & 15) + 3 * ((0x222110 >> (4 * d)) & 15);
                v[h] = st_b[pos[h]];
            }
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int j = j0 + 64 * h + lane;
                if (j < len) {
                    Kdata[gb + j] = v[h];
                    if (WRITE_IDX) {
                        const int a = pos[h] / 9;
                        const int m = pos[h] - 9 * a;   // kk + 3 ll within the block
                        const int ll = m / 3, kk = m - 3 * ll;
                        Krow[gb + j] = sm.dof[a] + kk;
                        Kcol[gb + j] = dofb + ll;
                    }
                }
            }
        }
    }
}

// Rounds of pass PASS for consumer wave W: stage, barrier, copy, (barrier before the next round).
template <int P, int NG, bool WRITE_IDX, int PASS, int W, int R = 0>
__device__ __forceinline__ void consumer_epilogue(ElemSmem<P, NG> &sm, double lam, double mu, int64_t base,
                                                  const ConsumerAcc<P, NG> &acc, double *__restrict__ Kdata,
                                                  int32_t *__restrict__ Krow, int32_t *__restrict__ Kcol) {
    constexpr int NR = Sched<P, NG>::S.nround[PASS];
    stage_round<P, NG, PASS, W, R>(sm, acc);
    __syncthreads();
    copy_round<P, NG, WRITE_IDX, PASS, R>(sm, lam, mu, base, Kdata, Krow, Kcol);
    if constexpr (R + 1 < NR) {
        __syncthreads();
        consumer_epilogue<P, NG, WRITE_IDX, PASS, W, R + 1>(sm, lam, mu, base, acc, Kdata, Krow, Kcol);
    }
}

// The producers' side of the same rounds (same barriers, they help with the copies).
template <int P, int NG, bool WRITE_IDX, int PASS, int R = 0>
__device__ __forceinline__ void producer_epilogue_pass(ElemSmem<P, NG> &sm, double lam, double mu, int64_t base,
                                                       double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                                       int32_t *__restrict__ Kcol) {
    constexpr int NR = Sched<P, NG>::S.nround[PASS];
    __syncthreads();
    copy_round<P, NG, WRITE_IDX, PASS, R>(sm, lam, mu, base, Kdata, Krow, Kcol);
    if constexpr (R + 1 < NR) {
        __syncthreads();
        producer_epilogue_pass<P, NG, WRITE_IDX, PASS, R + 1>(sm, lam, mu, base, Kdata, Krow, Kcol);
    }
}

template <int P, int NG, bool WRITE_IDX, int PASS = 0>
__device__ __forceinline__ void producer_epilogue(ElemSmem<P, NG> &sm, int pass, double lam, double mu,
                                                  int64_t base, double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                                  int32_t *__restrict__ Kcol) {
    if (pass == PASS) {
        producer_epilogue_pass<P, NG, WRITE_IDX, PASS>(sm, lam, mu, base, Kdata, Krow, Kcol);
    } else if constexpr (PASS + 1 < Sched<P, NG>::NPASS) {
        producer_epilogue<P, NG, WRITE_IDX, PASS + 1>(sm, pass, lam, mu, base, Kdata, Krow, Kcol);
    }
}

// One consumer step for wave W: the pass is dispatched to compile-time code.
template <int P, int NG, bool WRITE_IDX, int W, int PASS = 0>
__device__ __forceinline__ void consumer_step(ElemSmem<P, NG> &sm, const PatchArgs &pa, int buf, int pass, int ch,
                                              int e, int64_t base, ConsumerAcc<P, NG> &acc,
                                              double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                              int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    if (pass == PASS) {
        if (ch == 0) acc.zero();
        mfma_chunk<P, NG, PASS, W>(sm, buf, acc);
        if (ch == C::NCH - 1) {
            if (WRITE_IDX && W < 2) {
                const int t = threadIdx.x;
                if (t < C::NN) sm.dof[t] = 3 * (pa.ien[(int64_t)C::NN * e + t] - 1);
            }
            consumer_epilogue<P, NG, WRITE_IDX, PASS, W>(sm, pa.lam, pa.mu, base, acc, Kdata, Krow, Kcol);
        }
    } else if constexpr (PASS + 1 < Sched<P, NG>::NPASS) {
        consumer_step<P, NG, WRITE_IDX, W, PASS + 1>(sm, pa, buf, pass, ch, e, base, acc, Kdata, Krow, Kcol);
    }
}

// The whole step loop of consumer wave W (one compile-time specialisation per wave, so each wave
// keeps only its own accumulators live). One barrier per step, plus the epilogue's rounds.
template <int P, int NG, bool WRITE_IDX, int W>
__device__ __forceinline__ void consumer_loop(ElemSmem<P, NG> &sm, const PatchArgs &pa, int nsteps,
                                              double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                              int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    constexpr int SPE = Sched<P, NG>::NPASS * C::NCH;
    ConsumerAcc<P, NG> acc;
    for (int s = 0; s < nsteps; ++s) {
        const int el = s / SPE, rem = s % SPE;
        const int pass = rem / C::NCH, ch = rem % C::NCH;
        const int64_t gi = blockIdx.x + (int64_t)el * gridDim.x;
        const int e = pa.elems[gi];
        const int64_t base = pa.offs[gi];
        consumer_step<P, NG, WRITE_IDX, W>(sm, pa, s & 1, pass, ch, e, base, acc, Kdata, Krow, Kcol);
        __syncthreads();
    }
}

template <int P, int NG, bool WRITE_IDX>
__global__ void __launch_bounds__(768, 1)
elem_stiffness_kernel(PatchArgs pa, const double *__restrict__ geo, double *__restrict__ Kdata,
                      int32_t *__restrict__ Krow, int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    constexpr int NPASS = Sched<P, NG>::NPASS;
    constexpr int SPE = NPASS * C::NCH;              // steps per element
    static_assert(NPASS <= C::MAXPASS, "schedule does not fit");
    __shared__ ElemSmem<P, NG> sm;

    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool producer = wave >= C::CW;
    const int pw = wave - C::CW;
    // elements of this workgroup: blockIdx.x, blockIdx.x + gridDim.x, ... (same count for all waves)
    const int nloc = (pa.n_active - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
    const int nsteps = nloc * SPE;

    // Producers and consumers run separate loops with the same number of barriers (one before
    // the first step, one after each step, and the epilogue rounds), so each role keeps only its
    // own registers live.
    if (producer) {
        ProducerElem<P, NG> pe;
        int pe_elem = -1;
        auto step_geo = [&](int st) {   // element batch index and first Gauss point of step st
            const int64_t gi = blockIdx.x + (int64_t)(st / SPE) * gridDim.x;
            return producer_prefetch<P, NG>(geo + gi * (ELEM_NGEO * C::NQ), pw, ((st % SPE) % C::NCH) * C::QC);
        };
        double gv = 0.0;   // geometry of the chunk built next
        if (nsteps > 0) {
            gv = step_geo(0);
            const double gnext = nsteps > 1 ? step_geo(1) : 0.0;
            producer_load_element<P, NG>(sm, pa, pa.elems[blockIdx.x], pw, pe);
            pe_elem = 0;
            producer_chunk<P, NG>(sm, gv, pw, 0, 0, pe);
            gv = gnext;
        }
        __syncthreads();
        for (int s = 0; s < nsteps; ++s) {
            const int s1 = s + 1;
            if (s1 < nsteps) {
                const double gnext = s1 + 1 < nsteps ? step_geo(s1 + 1) : 0.0;   // one step ahead
                const int el = s1 / SPE, ch = (s1 % SPE) % C::NCH;
                if (el != pe_elem) {
                    producer_load_element<P, NG>(sm, pa, pa.elems[blockIdx.x + (int64_t)el * gridDim.x], pw, pe);
                    pe_elem = el;
                }
                producer_chunk<P, NG>(sm, gv, pw, s1 & 1, ch * C::QC, pe);
                gv = gnext;
            }
            const int rem = s % SPE;
            if (rem % C::NCH == C::NCH - 1) {
                const int64_t gi = blockIdx.x + (int64_t)(s / SPE) * gridDim.x;
                producer_epilogue<P, NG, WRITE_IDX>(sm, rem / C::NCH, pa.lam, pa.mu, pa.offs[gi], Kdata, Krow, Kcol);
            }
            __syncthreads();
        }
    } else {
        __syncthreads();
        switch (wave) {
        case 0: consumer_loop<P, NG, WRITE_IDX, 0>(sm, pa, nsteps, Kdata, Krow, Kcol); break;
        case 1: consumer_loop<P, NG, WRITE_IDX, 1>(sm, pa, nsteps, Kdata, Krow, Kcol); break;
        case 2: consumer_loop<P, NG, WRITE_IDX, 2>(sm, pa, nsteps, Kdata, Krow, Kcol); break;
        case 3: consumer_loop<P, NG, WRITE_IDX, 3>(sm, pa, nsteps, Kdata, Krow, Kcol); break;
        case 4: consumer_loop<P, NG, WRITE_IDX, 4>(sm, pa, nsteps, Kdata, Krow, Kcol); break;
        case 5: consumer_loop<P, NG, WRITE_IDX, 5>(sm, pa, nsteps, Kdata, Krow, Kcol); break;
        case 6: consumer_loop<P, NG, WRITE_IDX, 6>(sm, pa, nsteps, Kdata, Krow, Kcol); break;
        default: consumer_loop<P, NG, WRITE_IDX, 7>(sm, pa, nsteps, Kdata, Krow, Kcol); break;
        }
    }
}
