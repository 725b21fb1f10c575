// mf.hip -- sum-factorised matrix-free apply y = K x of the NURBS linear-elasticity stiffness, in
// FP32 or FP64, plus its 3x3 nodal block diagonal.
//
// Reference pattern: YETI's matrix-free operator mf_iga_get_ku_3d
// (iga_wq_mf/fsrc/mf_iga_solver.f90:292-555) applies y = sum_ij B_i^T diag(c_ij W) B_j x with
// tensor-product (Kronecker) 1D basis matrices (tensor_algebra.f90:369-486) for a scalar
// B-spline operator. This generalises it to 3-component elasticity on rational (NURBS) patches:
//   grad_xi u_l = (1/W) [ sum_b w_b x_b[l] dB_b - (dW/W) sum_b w_b x_b[l] B_b ]   (quotient rule)
//   H = grad_xi u J^{-1},  sigma = lam tr(H) I + mu (H + H^T)
//   y_a[k] = w_a sum_q [ sum_m T_km dB_a/dxi_m - (sum_m T_km dW_m/W) B_a ],
//   T_km = w|det J| sum_l sigma_kl J^{-1}_ml / W
// which equals K x for K_ab[k][l] = sum_q w|J| (lam dR_a/dx_k dR_b/dx_l + mu dR_a/dx_l dR_b/dx_k
// + mu d_kl grad R_a . grad R_b), the matrix the assembly kernel writes.
//
// Per quadrature point the stored geometry is 13 numbers: A = J^{-1}/W (9), dW/W (3), w|det J|
// (1), computed once in FP64 and kept in the apply's precision. The tensor contractions run one
// thread per quadrature point (then per node) with LDS staging. Elements of a patch are grouped
// in bricks of 2(p+1) elements per direction; a workgroup walks a brick colour by colour
// ((p+1)^3 colours, elements of one colour share no node), adding element results into an LDS
// brick accumulator in a fixed order; a gather kernel then sums the <= 8 bricks covering each
// control point in a fixed order. No floating-point atomics anywhere: the apply is
// bit-reproducible.
#include <algorithm>
#include <cmath>
#include <vector>

#include "batch.hpp"
#include "mf_sweep.hpp"

namespace yeti {

// -- host copy of the B-spline routine (same algorithm as the device one in assembly.hip) --
static void host_bspline(int i, int p, double u, const double *U, double *N, double *dN) {
    double left[16], right[16], cur[16], low[16];
    cur[0] = 1.0;
    low[0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[i + 1 - j];
        right[j] = U[i + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            double tmp = cur[r] / (right[r + 1] + left[j - r]);
            cur[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        cur[j] = saved;
        if (j == p - 1)
            for (int r = 0; r < p; ++r) low[r] = cur[r];
    }
    for (int r = 0; r <= p; ++r) {
        N[r] = cur[r];
        int k = i - p + r;
        double d = 0.0;
        if (r >= 1) {
            double den = U[k + p] - U[k];
            if (den != 0.0) d += p * low[r - 1] / den;
        }
        if (r <= p - 1) {
            double den = U[k + p + 1] - U[k + 1];
            if (den != 0.0) d -= p * low[r] / den;
        }
        dN[r] = d;
    }
}

void gauss_rule_host(int n, double *x, double *w);   // assembly.hip

template <int P, int NG>
struct MfCfg {
    static constexpr int NP1 = P + 1;
    static constexpr int NQ = NG * NG * NG;
    static constexpr int NN = NP1 * NP1 * NP1;
    static constexpr int TPE0 = NQ > NN ? NQ : NN;
    // threads per element: padded to a power of two up to a wave, so that an element's threads
    // lie in one wave and its contraction stages need only wave-level ordering
    static constexpr int TPE = TPE0 <= 8 ? 8 : TPE0 <= 16 ? 16 : TPE0 <= 32 ? 32 : TPE0 <= 64 ? 64 : TPE0;
    static constexpr bool WAVE_LOCAL = 64 % TPE == 0;
    static constexpr int EPB = 256 / TPE;               // elements per batch
    static constexpr int BR = 2 * NP1;                  // brick edge (elements)
    static constexpr int BN = BR + P;                   // brick edge (nodes)
    static constexpr int MAXD = NG > NP1 ? NG : NP1;
    static constexpr int SLOT = 12 * MAXD * MAXD * MAXD; // LDS values per element slot per buffer
    static_assert(EPB >= 1, "element too large for one workgroup");
};


template <typename T, int P, int NG, int NC, int BN>
struct MfSmem {
    using C = MfCfg<P, NG>;
    T buf0[C::EPB][C::SLOT];
    T buf1[C::EPB][C::SLOT];
    T acc[BN * BN * BN * NC];
    T B[C::EPB][3][NG][C::NP1];
    T dB[C::EPB][3][NG][C::NP1];
    T w[C::EPB][C::NN];
};

// geometry of one patch, in the sweep apply's order (sweep_geo_off)
template <int P, int NG>
__global__ void mf_geometry_kernel(MfPatch mp, const double *__restrict__ coords, const double *__restrict__ wcp,
                                   const double *__restrict__ half, GaussRule g, double *__restrict__ geo64,
                                   float *__restrict__ geo32) {
    using C = MfCfg<P, NG>;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nel = (int64_t)mp.N[0] * mp.N[1] * mp.N[2];
    if (t >= nel * C::NQ) return;
    const int64_t e = t / C::NQ;
    const int q = (int)(t % C::NQ);
    const int e1 = (int)(e % mp.N[0]), e2 = (int)((e / mp.N[0]) % mp.N[1]), e3 = (int)(e / ((int64_t)mp.N[0] * mp.N[1]));
    const int g0 = q % NG, g1 = (q / NG) % NG, g2 = q / (NG * NG);
    const int ed[3] = {e1, e2, e3};
    const int gg[3] = {g0, g1, g2};
    double Bv[3][C::NP1], dBv[3][C::NP1];
    for (int d = 0; d < 3; ++d)
        for (int r = 0; r < C::NP1; ++r) {
            int64_t o = ((int64_t)(mp.boff[d] + ed[d]) * NG + gg[d]) * C::NP1 + r;
            Bv[d][r] = mp.Btab[o];
            dBv[d][r] = mp.dBtab[o];
        }
    double s[16] = {0};
    for (int k = 0; k < C::NP1; ++k)
        for (int j = 0; j < C::NP1; ++j)
            for (int i = 0; i < C::NP1; ++i) {
                int64_t cpl = (e1 + i) + (int64_t)mp.ncp[0] * ((e2 + j) + (int64_t)mp.ncp[1] * (e3 + k));
                int cp = mp.cpmap[cpl];
                double w = wcp[cp];
                double N = Bv[0][i] * Bv[1][j] * Bv[2][k] * w;
                double d0 = dBv[0][i] * Bv[1][j] * Bv[2][k] * w;
                double d1 = Bv[0][i] * dBv[1][j] * Bv[2][k] * w;
                double d2 = Bv[0][i] * Bv[1][j] * dBv[2][k] * w;
                const double *x = coords + 3 * (int64_t)cp;
                s[0] += N; s[1] += d0; s[2] += d1; s[3] += d2;
                for (int r = 0; r < 3; ++r) {
                    s[4 + r] += x[r] * N;
                    s[7 + 3 * r] += x[r] * d0;
                    s[8 + 3 * r] += x[r] * d1;
                    s[9 + 3 * r] += x[r] * d2;
                }
            }
    const double W = s[0], Wi = 1.0 / W;
    const double dw[3] = {s[1] * Wi, s[2] * Wi, s[3] * Wi};
    double J[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[r][c] = (s[7 + 3 * r + c] - s[4 + r] * dw[c]) * Wi;
    const double m00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double m01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double m02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * m00 + J[0][1] * m01 + J[0][2] * m02;
    const double id = 1.0 / det;
    double Ji[9];   // Ji[3 m + k] = d xi_m / d x_k
    Ji[0] = m00 * id;
    Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[3] = m01 * id;
    Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[6] = m02 * id;
    Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const double hdet = half[3 * e1 + 0] * half[3 * e2 + 1] * half[3 * e3 + 2];
    // half[] is [e_d][3] per direction with the direction's element index
    const double om = g.w[g0] * g.w[g1] * g.w[g2] * fabs(det * hdet);
    double out[13];
    for (int c = 0; c < 9; ++c) out[c] = Ji[c] * Wi;
    out[9] = dw[0];
    out[10] = dw[1];
    out[11] = dw[2];
    out[12] = om;
    const int64_t o64 = sweep_geo_off<P, 8>(mp.gnb[0], e1, e2, e3, 0, g0, g1, g2);
    const int64_t o32 = sweep_geo_off<P, 4>(mp.gnb[1], e1, e2, e3, 0, g0, g1, g2);
    for (int c = 0; c < 13; ++c) {
        if (geo64) geo64[o64 + (int64_t)c * Sweep<P, 8>::NTQ] = out[c];
        if (geo32) geo32[o32 + (int64_t)c * Sweep<P, 4>::NTQ] = (float)out[c];
    }
}

// Brick edge in elements: 2(p+1) for the apply and the shape gradient (8 elements per colour),
// p+1 for the diagonal blocks (1 element per colour, keeps the 9-value accumulator within LDS).
// Orders an element's LDS stage writes before the next stage's reads: within a wave when the
// element's threads share one, else across the workgroup.
template <bool WAVE_LOCAL>
__device__ __forceinline__ void mf_stage_sync() {
    if constexpr (WAVE_LOCAL) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

template <int P, int MODE>
struct MfBrick {
    static constexpr int BR = MODE == 1 ? (P + 1) : 2 * (P + 1);
    static constexpr int BN = BR + P;
    static constexpr int K = BR / (P + 1);            // elements per colour per direction
    static constexpr int EPC = K * K * K;
};

// MODE 0: y = K x (3 values per node); MODE 1: 3x3 diagonal blocks (9 values per node);
// MODE 2: shape derivative of the strain energy, y_cj = -x^T (dK / dX_cj) x (3 values per node).
// For MODE 2 the quadrature-point tensor is M = 2 H^T sigma - (sigma : H) I in place of sigma:
// with d(dR_a/dx_k)/dX_cj = -dR_a/dx_j dR_c/dx_k and d(det J)/dX_cj = det J dR_c/dx_j,
// -x^T dK_e/dX_cj x = sum_q w|det J| sum_k M_jk dR_c/dx_k, the same backward contraction as K x.
// The work of one brick (`brick`, in the nb0 x nb1 x . brick grid of the patch): its elements'
// contributions summed into partial[brick]. Every thread of the workgroup calls it.
template <typename T, int P, int NG, int MODE>
__device__ __forceinline__ void mf_brick_body(MfSmem<T, P, NG, MODE == 1 ? 9 : 3, MfBrick<P, MODE>::BN> &sm,
                                              const MfPatch &mp, const T *__restrict__ geo, const T *__restrict__ x,
                                              const T *__restrict__ wcp, int32_t nb0, int32_t nb1,
                                              T *__restrict__ partial, int64_t brick) {
    using C = MfCfg<P, NG>;
    using Bk = MfBrick<P, MODE>;
    constexpr int NP1 = C::NP1, NQ = C::NQ, NN = C::NN, BN = Bk::BN, BR = Bk::BR, NC = MODE == 1 ? 9 : 3;
    const int tid = threadIdx.x;
    const int slot = tid / C::TPE, r = tid % C::TPE;
    const int b1 = (int)(brick % nb0), b2 = (int)((brick / nb0) % nb1), b3 = (int)(brick / ((int64_t)nb0 * nb1));
    for (int i = tid; i < BN * BN * BN * NC; i += 256) sm.acc[i] = T(0);
    const T lam = (T)mp.lam, mu = (T)mp.mu;
    // elements of colour c: e_d = b_d BR + c_d + NP1 t_d, t_d in [0, K): EPC per colour
    for (int color = 0; color < NP1 * NP1 * NP1; ++color) {
        const int c1 = color % NP1, c2 = (color / NP1) % NP1, c3 = color / (NP1 * NP1);
        for (int bt = 0; bt < Bk::EPC; bt += C::EPB) {
            // the previous colour's accumulation is complete (elements of one colour share no node;
            // the staging buffers of a slot are used by its own wave only)
            if constexpr (C::WAVE_LOCAL) {
                if (bt == 0) __syncthreads();
                else mf_stage_sync<true>();
            } else {
                __syncthreads();
            }
            const int tt = bt + slot;
            const bool sl = slot < C::EPB && tt < Bk::EPC;
            const int t1 = tt % Bk::K, t2 = (tt / Bk::K) % Bk::K, t3 = tt / (Bk::K * Bk::K);
            const int e1 = b1 * BR + c1 + NP1 * t1, e2 = b2 * BR + c2 + NP1 * t2, e3 = b3 * BR + c3 + NP1 * t3;
            const bool live = sl && e1 < mp.N[0] && e2 < mp.N[1] && e3 < mp.N[2];
            const int64_t e = e1 + (int64_t)mp.N[0] * (e2 + (int64_t)mp.N[1] * e3);
            T *b0 = sm.buf0[slot < C::EPB ? slot : 0];
            T *b1v = sm.buf1[slot < C::EPB ? slot : 0];
            // stage nodal data and 1D tables
            if (live && r < NN) {
                const int i = r % NP1, j = (r / NP1) % NP1, k = r / (NP1 * NP1);
                int64_t cpl = (e1 + i) + (int64_t)mp.ncp[0] * ((e2 + j) + (int64_t)mp.ncp[1] * (e3 + k));
                int cp = mp.cpmap[cpl];
                T w = wcp[cp];
                sm.w[slot][r] = w;
                if (MODE != 1)
                    for (int l = 0; l < 3; ++l) b0[l * NN + r] = w * x[3 * (int64_t)cp + l];
            }
            if (live)
                for (int idx = r; idx < 3 * NG * NP1; idx += C::TPE) {
                    const int d = idx / (NG * NP1), rem = idx % (NG * NP1), gq = rem / NP1, rr = rem % NP1;
                    const int ed = d == 0 ? e1 : d == 1 ? e2 : e3;
                    int64_t o = ((int64_t)(mp.boff[d] + ed) * NG + gq) * NP1 + rr;
                    sm.B[slot][d][gq][rr] = (T)mp.Btab[o];
                    sm.dB[slot][d][gq][rr] = (T)mp.dBtab[o];
                }
            mf_stage_sync<C::WAVE_LOCAL>();
            T yv[NC];
            for (int c = 0; c < NC; ++c) yv[c] = T(0);
            if (MODE != 1) {
                const auto &B = sm.B[slot < C::EPB ? slot : 0];
                const auto &dB = sm.dB[slot < C::EPB ? slot : 0];
                // stage 1: contract i -> g0. b1v[(f*2+v)*NG*NP1*NP1 + (k*NP1+j)*NG + g0]
                if (live && r < NG * NP1 * NP1) {
                    const int g0 = r % NG, jk = r / NG;
                    for (int f = 0; f < 3; ++f) {
                        T v = 0, d = 0;
                        for (int i = 0; i < NP1; ++i) {
                            T fv = b0[f * NN + jk * NP1 + i];
                            v += fv * B[0][g0][i];
                            d += fv * dB[0][g0][i];
                        }
                        b1v[(f * 2 + 0) * NG * NP1 * NP1 + r] = v;
                        b1v[(f * 2 + 1) * NG * NP1 * NP1 + r] = d;
                    }
                }
                mf_stage_sync<C::WAVE_LOCAL>();
                // stage 2: contract j -> g1. b0[(f*3+v)*NG*NG*NP1 + (k*NG+g1)*NG + g0], v: val, d0, d1
                if (live && r < NG * NG * NP1) {
                    const int g0 = r % NG, g1 = (r / NG) % NG, k = r / (NG * NG);
                    for (int f = 0; f < 3; ++f) {
                        T v = 0, d0 = 0, d1 = 0;
                        for (int j = 0; j < NP1; ++j) {
                            T sv = b1v[(f * 2 + 0) * NG * NP1 * NP1 + (k * NP1 + j) * NG + g0];
                            T sd = b1v[(f * 2 + 1) * NG * NP1 * NP1 + (k * NP1 + j) * NG + g0];
                            v += sv * B[1][g1][j];
                            d1 += sv * dB[1][g1][j];
                            d0 += sd * B[1][g1][j];
                        }
                        b0[(f * 3 + 0) * NG * NG * NP1 + r] = v;
                        b0[(f * 3 + 1) * NG * NG * NP1 + r] = d0;
                        b0[(f * 3 + 2) * NG * NG * NP1 + r] = d1;
                    }
                }
                mf_stage_sync<C::WAVE_LOCAL>();
                // stage 3 (k -> g2) and the quadrature-point work
                if (live && r < NQ) {
                    const int g0 = r % NG, g1 = (r / NG) % NG, g2 = r / (NG * NG);
                    T V[3], G[3][3];    // G[m][l] = sum w x_l dB/dxi_m
                    for (int f = 0; f < 3; ++f) {
                        T v = 0, d0 = 0, d1 = 0, d2 = 0;
                        for (int k = 0; k < NP1; ++k) {
                            const int o = (k * NG + g1) * NG + g0;
                            T sv = b0[(f * 3 + 0) * NG * NG * NP1 + o];
                            v += sv * B[2][g2][k];
                            d2 += sv * dB[2][g2][k];
                            d0 += b0[(f * 3 + 1) * NG * NG * NP1 + o] * B[2][g2][k];
                            d1 += b0[(f * 3 + 2) * NG * NG * NP1 + o] * B[2][g2][k];
                        }
                        V[f] = v;
                        G[0][f] = d0;
                        G[1][f] = d1;
                        G[2][f] = d2;
                    }
                    constexpr int CS = Sweep<P, sizeof(T)>::NTQ;   // stride between geometry values
                    const T *gq = geo + sweep_geo_off<P, sizeof(T)>(mp.gnb[sizeof(T) == 8 ? 0 : 1], e1, e2, e3, 0, g0,
                                                                     g1, g2);
                    T A[9], gw[3];
                    for (int c = 0; c < 9; ++c) A[c] = gq[c * CS];
                    gw[0] = gq[9 * CS];
                    gw[1] = gq[10 * CS];
                    gw[2] = gq[11 * CS];
                    const T om = gq[12 * CS];
                    // H[l][k] = d u_l / d x_k = sum_m (G[m][l] - gw[m] V[l]) A[m][k]
                    T H[3][3];
                    for (int l = 0; l < 3; ++l)
                        for (int k = 0; k < 3; ++k) {
                            T s = 0;
                            for (int m = 0; m < 3; ++m) s += (G[m][l] - gw[m] * V[l]) * A[3 * m + k];
                            H[l][k] = s;
                        }
                    const T tr = lam * (H[0][0] + H[1][1] + H[2][2]);
                    T S[3][3];
                    for (int k = 0; k < 3; ++k)
                        for (int l = 0; l < 3; ++l) S[k][l] = mu * (H[k][l] + H[l][k]) + (k == l ? tr : T(0));
                    if constexpr (MODE == 2) {
                        T en = 0;   // sigma : H
                        for (int k = 0; k < 3; ++k)
                            for (int l = 0; l < 3; ++l) en += S[k][l] * H[k][l];
                        T M[3][3];
                        for (int j = 0; j < 3; ++j)
                            for (int k = 0; k < 3; ++k) {
                                T m = 0;
                                for (int i = 0; i < 3; ++i) m += H[i][j] * S[i][k];
                                M[j][k] = 2 * m - (j == k ? en : T(0));
                            }
                        for (int j = 0; j < 3; ++j)
                            for (int k = 0; k < 3; ++k) S[j][k] = M[j][k];
                    }
                    // coefficients of dB/dxi_m (Tm) and of B (Tb) for each component k
                    for (int k = 0; k < 3; ++k) {
                        T tb = 0;
                        for (int m = 0; m < 3; ++m) {
                            T s = 0;
                            for (int l = 0; l < 3; ++l) s += S[k][l] * A[3 * m + l];
                            s *= om;
                            b1v[(k * 4 + 1 + m) * NQ + r] = s;
                            tb -= s * gw[m];
                        }
                        b1v[(k * 4 + 0) * NQ + r] = tb;
                    }
                }
                mf_stage_sync<C::WAVE_LOCAL>();
                // backward stage 3^T (g2 -> k): P, R0, R1 per component into b0
                if (live && r < NG * NG * NP1) {
                    const int g0 = r % NG, g1 = (r / NG) % NG, k = r / (NG * NG);
                    for (int c = 0; c < 3; ++c) {
                        T Pv = 0, R0 = 0, R1 = 0;
                        for (int g2 = 0; g2 < NG; ++g2) {
                            const int q = g0 + NG * (g1 + NG * g2);
                            Pv += b1v[(c * 4 + 0) * NQ + q] * B[2][g2][k] + b1v[(c * 4 + 3) * NQ + q] * dB[2][g2][k];
                            R0 += b1v[(c * 4 + 1) * NQ + q] * B[2][g2][k];
                            R1 += b1v[(c * 4 + 2) * NQ + q] * B[2][g2][k];
                        }
                        b0[(c * 3 + 0) * NG * NG * NP1 + r] = Pv;
                        b0[(c * 3 + 1) * NG * NG * NP1 + r] = R0;
                        b0[(c * 3 + 2) * NG * NG * NP1 + r] = R1;
                    }
                }
                mf_stage_sync<C::WAVE_LOCAL>();
                // backward stage 2^T (g1 -> j): P', R0' into b1v
                if (live && r < NG * NP1 * NP1) {
                    const int g0 = r % NG, j = (r / NG) % NP1, k = r / (NG * NP1);
                    for (int c = 0; c < 3; ++c) {
                        T Pv = 0, R0 = 0;
                        for (int g1 = 0; g1 < NG; ++g1) {
                            const int o = (k * NG + g1) * NG + g0;
                            Pv += b0[(c * 3 + 0) * NG * NG * NP1 + o] * B[1][g1][j] +
                                  b0[(c * 3 + 2) * NG * NG * NP1 + o] * dB[1][g1][j];
                            R0 += b0[(c * 3 + 1) * NG * NG * NP1 + o] * B[1][g1][j];
                        }
                        b1v[(c * 2 + 0) * NG * NP1 * NP1 + r] = Pv;
                        b1v[(c * 2 + 1) * NG * NP1 * NP1 + r] = R0;
                    }
                }
                mf_stage_sync<C::WAVE_LOCAL>();
                // backward stage 1^T (g0 -> i): node values
                if (live && r < NN) {
                    const int i = r % NP1, jk = r / NP1;
                    for (int c = 0; c < 3; ++c) {
                        T s = 0;
                        for (int g0 = 0; g0 < NG; ++g0) {
                            const int o = jk * NG + g0;
                            s += b1v[(c * 2 + 0) * NG * NP1 * NP1 + o] * B[0][g0][i] +
                                 b1v[(c * 2 + 1) * NG * NP1 * NP1 + o] * dB[0][g0][i];
                        }
                        yv[c] = s * sm.w[slot][r];
                    }
                }
            } else {
                // diagonal blocks: thread = node a, loop over quadrature points
                const auto &B = sm.B[slot < C::EPB ? slot : 0];
                const auto &dB = sm.dB[slot < C::EPB ? slot : 0];
                if (live && r < NN) {
                    const int i = r % NP1, j = (r / NP1) % NP1, k = r / (NP1 * NP1);
                    const T w = sm.w[slot][r];
                    constexpr int CS = Sweep<P, sizeof(T)>::NTQ;
                    for (int q = 0; q < NQ; ++q) {
                        const int g0 = q % NG, g1 = (q / NG) % NG, g2 = q / (NG * NG);
                        const T *gq = geo + sweep_geo_off<P, sizeof(T)>(mp.gnb[sizeof(T) == 8 ? 0 : 1], e1, e2, e3, 0,
                                                                         g0, g1, g2);
                        const T Bq = B[0][g0][i] * B[1][g1][j] * B[2][g2][k];
                        const T dq[3] = {dB[0][g0][i] * B[1][g1][j] * B[2][g2][k], B[0][g0][i] * dB[1][g1][j] * B[2][g2][k],
                                         B[0][g0][i] * B[1][g1][j] * dB[2][g2][k]};
                        T gx[3];
                        for (int kk = 0; kk < 3; ++kk) {
                            T s = 0;
                            for (int m = 0; m < 3; ++m) s += (dq[m] - Bq * gq[(9 + m) * CS]) * gq[(3 * m + kk) * CS];
                            gx[kk] = w * s;
                        }
                        const T om = gq[12 * CS];
                        const T nn2 = gx[0] * gx[0] + gx[1] * gx[1] + gx[2] * gx[2];
                        for (int kk = 0; kk < 3; ++kk)
                            for (int ll = 0; ll < 3; ++ll)
                                yv[3 * kk + ll] += om * ((lam + mu) * gx[kk] * gx[ll] + (kk == ll ? mu * nn2 : T(0)));
                    }
                }
            }
            // accumulate into the brick: elements of one colour share no node
            if (live && r < NN) {
                const int i = r % NP1, j = (r / NP1) % NP1, k = r / (NP1 * NP1);
                const int n1 = e1 - b1 * BR + i, n2 = e2 - b2 * BR + j, n3 = e3 - b3 * BR + k;
                T *dst = sm.acc + ((n3 * BN + n2) * BN + n1) * NC;
                for (int c = 0; c < NC; ++c) dst[c] += yv[c];
            }
        }
    }
    __syncthreads();
    T *out = partial + brick * (int64_t)(BN * BN * BN * NC);
    for (int i = tid; i < BN * BN * BN * NC; i += 256) out[i] = sm.acc[i];
}

template <typename T, int P, int NG, int MODE>
__global__ void __launch_bounds__(256) mf_brick_kernel(MfPatch mp, const T *__restrict__ geo, const T *__restrict__ x,
                                                       const T *__restrict__ wcp, const int32_t *__restrict__ nbrick,
                                                       T *__restrict__ partial, const int32_t *__restrict__ pred) {
    if (pred && *pred == 0) return;   // predicated launch (device-side solver control)
    __shared__ MfSmem<T, P, NG, MODE == 1 ? 9 : 3, MfBrick<P, MODE>::BN> sm;
    mf_brick_body<T, P, NG, MODE>(sm, mp, geo, x, wcp, nbrick[0], nbrick[1], partial, blockIdx.x);
}

// y[global cp] (+)= sum over the bricks covering patch CP t (fixed order); bricks of br[d]
// elements cover bn[d] = br[d] + p control points along direction d, in a grid of nb[d].
template <typename T, int NC>
__device__ __forceinline__ void mf_gather_point(const MfPatch &mp, const int32_t br[3], const int32_t bn[3],
                                                const int32_t nb[3], const T *__restrict__ partial, T *__restrict__ y,
                                                int accumulate, int64_t t) {
    const int c1 = (int)(t % mp.ncp[0]), c2 = (int)((t / mp.ncp[0]) % mp.ncp[1]),
              c3 = (int)(t / ((int64_t)mp.ncp[0] * mp.ncp[1]));
    const int cc[3] = {c1, c2, c3};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        // brick b covers CPs [b br, b br + bn)
        hi[d] = std::min(nb[d] - 1, cc[d] / br[d]);
        int l = (cc[d] - bn[d]) / br[d] + 1;
        if (cc[d] - bn[d] < 0) l = 0;
        lo[d] = std::max(0, l);
    }
    const int64_t bvol = (int64_t)bn[0] * bn[1] * bn[2] * NC;
    T s[NC];
    for (int c = 0; c < NC; ++c) s[c] = T(0);
    for (int b3 = lo[2]; b3 <= hi[2]; ++b3)
        for (int b2 = lo[1]; b2 <= hi[1]; ++b2)
            for (int b1 = lo[0]; b1 <= hi[0]; ++b1) {
                const int64_t brick = b1 + (int64_t)nb[0] * (b2 + (int64_t)nb[1] * b3);
                const int n1 = c1 - b1 * br[0], n2 = c2 - b2 * br[1], n3 = c3 - b3 * br[2];
                const T *src = partial + brick * bvol + ((n3 * bn[1] + n2) * bn[0] + n1) * NC;
                for (int c = 0; c < NC; ++c) s[c] += src[c];
            }
    const int cp = mp.cpmap[t];
    for (int c = 0; c < NC; ++c) {
        if (accumulate) y[(int64_t)NC * cp + c] += s[c];
        else y[(int64_t)NC * cp + c] = s[c];
    }
}

struct BrickGrid {
    int32_t br[3], bn[3], nb[3];
};

template <typename T, int NC>
__global__ void mf_gather_kernel(MfPatch mp, BrickGrid bg, const T *__restrict__ partial, T *__restrict__ y,
                                 int accumulate, const int32_t *__restrict__ pred) {
    if (pred && *pred == 0) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncp = (int64_t)mp.ncp[0] * mp.ncp[1] * mp.ncp[2];
    if (t >= ncp) return;
    mf_gather_point<T, NC>(mp, bg.br, bg.bn, bg.nb, partial, y, accumulate, t);
}

// The apply as brick sweeps: one workgroup of Sweep<P, sizeof(T)>::NT threads per brick.
template <typename T, int P>
__global__ void __launch_bounds__((Sweep<P, sizeof(T)>::NT)) __attribute__((amdgpu_waves_per_eu(Sweep<P, sizeof(T)>::WPE)))
mf_sweep_kernel(MfPatch mp, const T *__restrict__ geo,
                                                                const T *__restrict__ x, const T *__restrict__ wcp,
                                                                T *__restrict__ partial,
                                                                const int32_t *__restrict__ pred) {
    if (pred && *pred == 0) return;   // predicated launch (device-side solver control)
    __shared__ SweepSmem<T, P> sm;
    mf_sweep_body<T, P>(sm, mp, geo, x, wcp, partial, blockIdx.x);
}

template <typename T>
__global__ void fill_zero_pred_kernel(T *p, int64_t n, const int32_t *pred) {
    if (pred && *pred == 0) return;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

template <typename T>
__global__ void fill_kernel(T *p, int64_t n, T v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace yeti

// ------------------------------------------------------------------------------------------
namespace yeti {
struct MfPatchDev {
    int p, ng;
    MfPatch mp;
    int32_t nbrick[2][3];                 // coloured bricks: [0] strain-energy gradient, [1] diagonal blocks
    int32_t sweep_br[2][3];               // sweep brick edge per direction, [0] FP64, [1] FP32 (grids: mp.gnb)
    DevBuf<int32_t> cpmap, nbrick_dev;    // [2][3]
    DevBuf<double> Btab, dBtab, half, geo64;
    DevBuf<float> geo32;
    DevBuf<double> part64;
    DevBuf<float> part32;
    GaussRule g;
    // host copies for the fast-diagonalisation set-up
    std::vector<double> hB, hdB, hhalf, hgw;
    std::vector<int32_t> hcpmap;
    std::vector<double> hU[3];
};
}  // namespace yeti

struct yeti_mf {
    int32_t nb_cp;
    int precisions;     // bit 0: FP64, bit 1: FP32
    bool zero_first = false;   // several patches, or CPs in no patch: y is cleared, then every patch adds
    std::vector<yeti::MfPatchDev *> patches;
    yeti::DevBuf<double> coords, w64;
    yeti::DevBuf<float> w32;
    ~yeti_mf() {
        for (auto *p : patches) delete p;
    }
};

namespace yeti {

template <typename T, int P, int NG, int MODE>
static int launch_brick(const MfPatchDev &pd, const T *geo, const T *x, const T *w, T *part, hipStream_t st,
                        const int32_t *pred) {
    constexpr int BM = MODE == 1 ? 1 : 0;   // brick grid: apply-sized or diagonal-sized
    const int32_t *nbk = pd.nbrick[BM];
    int64_t nb = (int64_t)nbk[0] * nbk[1] * nbk[2];
    hipLaunchKernelGGL((mf_brick_kernel<T, P, NG, MODE>), dim3((unsigned)nb), dim3(256), 0, st, pd.mp, geo, x, w,
                       pd.nbrick_dev.p + 3 * BM, part, pred);
    YETI_HIP(hipGetLastError());
    return 0;
}

template <typename T, int P>
static int launch_sweep(const MfPatchDev &pd, const T *geo, const T *x, const T *w, T *part, hipStream_t st,
                        const int32_t *pred) {
    const int32_t *g = pd.mp.gnb[sizeof(T) == 8 ? 0 : 1];
    const int64_t nb = (int64_t)g[0] * g[1] * g[2];
    hipLaunchKernelGGL((mf_sweep_kernel<T, P>), dim3((unsigned)nb), dim3(Sweep<P, sizeof(T)>::NT), 0, st, pd.mp, geo, x, w,
                       part,
                       pred);
    YETI_HIP(hipGetLastError());
    return 0;
}

// MODE 0 (K x): brick sweeps; MODE 1 and 2: the element-wise coloured bricks.
template <typename T, int MODE, int P>
static int launch_mode(const MfPatchDev &pd, const T *geo, const T *x, const T *w, T *part, hipStream_t st,
                       const int32_t *pred) {
    if constexpr (MODE == 0) return launch_sweep<T, P>(pd, geo, x, w, part, st, pred);
    else return launch_brick<T, P, P + 1, MODE>(pd, geo, x, w, part, st, pred);
}

template <typename T, int MODE>
static int dispatch_brick(const MfPatchDev &pd, const T *geo, const T *x, const T *w, T *part, hipStream_t st,
                          const int32_t *pred) {
    switch (pd.p * 16 + pd.ng) {
    case 1 * 16 + 2: return launch_mode<T, MODE, 1>(pd, geo, x, w, part, st, pred);
    case 2 * 16 + 3: return launch_mode<T, MODE, 2>(pd, geo, x, w, part, st, pred);
    case 3 * 16 + 4: return launch_mode<T, MODE, 3>(pd, geo, x, w, part, st, pred);
    case 4 * 16 + 5: return launch_mode<T, MODE, 4>(pd, geo, x, w, part, st, pred);
    default: return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "matrix-free: no kernel for p=%d ng=%d", pd.p, pd.ng);
    }
}

// The bricks whose partial sums MODE's kernel writes (MODE 0 in precision pi: 0 FP64, 1 FP32).
static BrickGrid brick_grid(const MfPatchDev &pd, int mode, int pi) {
    BrickGrid bg;
    for (int d = 0; d < 3; ++d) {
        if (mode == 0) {
            bg.br[d] = pd.sweep_br[pi][d];
            bg.nb[d] = pd.mp.gnb[pi][d];
        } else {
            bg.br[d] = mode == 1 ? pd.p + 1 : 2 * (pd.p + 1);
            bg.nb[d] = pd.nbrick[mode == 1 ? 1 : 0][d];
        }
        bg.bn[d] = bg.br[d] + pd.p;
    }
    return bg;
}

template <typename T, int MODE>
static int mf_run(const yeti_mf *m, const T *x, T *y, hipStream_t st, const int32_t *pred = nullptr) {
    constexpr int NC = MODE == 1 ? 9 : 3;
    if (m->zero_first) {
        const int64_t n = (int64_t)NC * m->nb_cp;
        hipLaunchKernelGGL((fill_zero_pred_kernel<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, n, pred);
        YETI_HIP(hipGetLastError());
    }
    for (size_t ip = 0; ip < m->patches.size(); ++ip) {
        const MfPatchDev &pd = *m->patches[ip];
        const T *geo;
        const T *w;
        T *part;
        if constexpr (sizeof(T) == 8) {
            geo = (const T *)pd.geo64.p;
            w = (const T *)m->w64.p;
            part = (T *)pd.part64.p;
        } else {
            geo = (const T *)pd.geo32.p;
            w = (const T *)m->w32.p;
            part = (T *)pd.part32.p;
        }
        if (!geo || !part) return set_error(YETI_ERR_INVALID_ARGUMENT, "matrix-free: precision not prepared");
        int rc = dispatch_brick<T, MODE>(pd, geo, x, w, part, st, pred);
        if (rc) return rc;
        const BrickGrid bg = brick_grid(pd, MODE, sizeof(T) == 8 ? 0 : 1);
        int64_t ncp = (int64_t)pd.mp.ncp[0] * pd.mp.ncp[1] * pd.mp.ncp[2];
        hipLaunchKernelGGL((mf_gather_kernel<T, NC>), dim3((unsigned)((ncp + 255) / 256)), dim3(256), 0, st, pd.mp, bg,
                           part, y, (m->zero_first || ip > 0) ? 1 : 0, pred);
        YETI_HIP(hipGetLastError());
    }
    return 0;
}

// Predicated applies for the device-controlled solvers (nothing runs while *pred == 0).
int mf_apply_pred(const yeti_mf *m, const double *x, double *y, hipStream_t st, const int32_t *pred) {
    return mf_run<double, 0>(m, x, y, st, pred);
}
int mf_apply_pred(const yeti_mf *m, const float *x, float *y, hipStream_t st, const int32_t *pred) {
    return mf_run<float, 0>(m, x, y, st, pred);
}

// Quadrature-point geometry of one patch from the handle's current coordinates.
static int mf_geometry(const yeti_mf *m, const MfPatchDev &pd, hipStream_t st) {
    const int64_t nel = (int64_t)pd.mp.N[0] * pd.mp.N[1] * pd.mp.N[2];
    const int64_t nt = nel * pd.ng * pd.ng * pd.ng;
    const unsigned blocks = (unsigned)((nt + 255) / 256);
    if (nt == 0) return 0;
    switch (pd.p) {
    case 1: hipLaunchKernelGGL((mf_geometry_kernel<1, 2>), dim3(blocks), dim3(256), 0, st, pd.mp, m->coords.p, m->w64.p, pd.half.p, pd.g, pd.geo64.p, pd.geo32.p); break;
    case 2: hipLaunchKernelGGL((mf_geometry_kernel<2, 3>), dim3(blocks), dim3(256), 0, st, pd.mp, m->coords.p, m->w64.p, pd.half.p, pd.g, pd.geo64.p, pd.geo32.p); break;
    case 3: hipLaunchKernelGGL((mf_geometry_kernel<3, 4>), dim3(blocks), dim3(256), 0, st, pd.mp, m->coords.p, m->w64.p, pd.half.p, pd.g, pd.geo64.p, pd.geo32.p); break;
    default: hipLaunchKernelGGL((mf_geometry_kernel<4, 5>), dim3(blocks), dim3(256), 0, st, pd.mp, m->coords.p, m->w64.p, pd.half.p, pd.g, pd.geo64.p, pd.geo32.p); break;
    }
    YETI_HIP(hipGetLastError());
    return 0;
}

}  // namespace yeti


// ==========================================================================================
// Fast diagonalisation preconditioner (single-patch handles), the preconditioner family of
// YETI's matrix-free CG (iga_wq_mf/fsrc/tensor_algebra.f90:992-1161, after Sangalli and Tani,
// "Isogeometric preconditioners based on fast solvers for the Sylvester equation", 2016).
// Component c of the displacement is preconditioned by the separable operator
//   A_c = sum_d alpha_cd K_d (x) M_(other directions),
// K_d / M_d the 1D B-spline stiffness / mass Gram matrices of direction d on the free index range
// of component c, alpha_cd = |det J| (mu g_dd + (lam + mu) (J^-1)_dc^2) from the patch-mean
// Jacobian J (the d = e terms of the diagonal block K_cc of isotropic elasticity). With
// K_d U_d = M_d U_d Lambda_d and U_d^T M_d U_d = I,
//   A_c^-1 = (U_2 (x) U_1 (x) U_0) diag(1 / (alpha_c0 l0_i + alpha_c1 l1_j + alpha_c2 l2_k))
//            (U_2 (x) U_1 (x) U_0)^T,
// applied as three mode products, a scaling and three mode products per component. Fixed dofs
// must be whole patch faces per component (the clamped face of the solids here).
namespace yeti {


}  // namespace yeti

struct yeti_fdm {
    yeti::FdmArgs fa;
    int32_t ncp = 0;
    yeti::DevBuf<double> U64, Ut64, invD64, wa64, wb64;
    yeti::DevBuf<float> U32, Ut32, invD32, wa32, wb32;
};

namespace yeti {

// cyclic Jacobi eigensolver of a symmetric n x n matrix a (row-major, destroyed): eigenvalues
// in w, eigenvectors in the columns of v
static void jacobi_eigen(int n, std::vector<double> &a, std::vector<double> &w, std::vector<double> &v) {
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) (i == j ? dia : off) += a[(size_t)i * n + j] * a[(size_t)i * n + j];
        if (off <= 1e-30 * dia) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {   // columns p, q
                    const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - sn * akq;
                    a[(size_t)k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {   // rows p, q
                    const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - sn * aqk;
                    a[(size_t)q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
                    v[(size_t)k * n + p] = c * vkp - sn * vkq;
                    v[(size_t)k * n + q] = sn * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
}

// K U = M U Lambda, U^T M U = I (M = L L^T, C = L^-1 K L^-T = Q Lambda Q^T, U = L^-T Q)
static bool gen_eigen(int n, const std::vector<double> &K, const std::vector<double> &M, std::vector<double> &lam,
                      std::vector<double> &U) {
    std::vector<double> L((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double d = M[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(d > 0.0)) return false;
        L[(size_t)j * n + j] = std::sqrt(d);
        for (int i = j + 1; i < n; ++i) {
            double s = M[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = s / L[(size_t)j * n + j];
        }
    }
    // Y = L^-1 K (forward substitution on columns), C = Y L^-T (on rows)
    std::vector<double> Y = K;
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < n; ++i) {
            double s = Y[(size_t)i * n + c];
            for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * Y[(size_t)k * n + c];
            Y[(size_t)i * n + c] = s / L[(size_t)i * n + i];
        }
    std::vector<double> Cm((size_t)n * n);
    for (int r = 0; r < n; ++r)
        for (int j = 0; j < n; ++j) {
            double s = Y[(size_t)r * n + j];
            for (int k = 0; k < j; ++k) s -= Cm[(size_t)r * n + k] * L[(size_t)j * n + k];
            Cm[(size_t)r * n + j] = s / L[(size_t)j * n + j];
        }
    for (int i = 0; i < n; ++i)   // symmetrise rounding
        for (int j = i + 1; j < n; ++j) {
            const double a = 0.5 * (Cm[(size_t)i * n + j] + Cm[(size_t)j * n + i]);
            Cm[(size_t)i * n + j] = Cm[(size_t)j * n + i] = a;
        }
    std::vector<double> Q;
    jacobi_eigen(n, Cm, lam, Q);
    // U = L^-T Q (back substitution on columns)
    U.assign((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c)
        for (int i = n - 1; i >= 0; --i) {
            double s = Q[(size_t)i * n + c];
            for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * U[(size_t)k * n + c];
            U[(size_t)i * n + c] = s / L[(size_t)i * n + i];
        }
    return true;
}

// One mode product of the preconditioner, out(.., a, ..) = sum_b A[a][b] in(.., b, ..) along one
// axis for the three components, written as a batched strided GEMM C = L R: along axis 0,
// L = A (m0 x m0) and R = the tensor seen as m0 x (m1 m2); along axis 1, per k, L = the tensor
// slice (m0 x m1) and R = A^T; along axis 2, L = the tensor seen as (m0 m1) x m2 and R = A^T.
// Element (row, col) of operand X of component c, batch z, is at X[off[c] + z bs[c] + row rs[c]
// + col cs[c]].

// K depth of the preconditioner GEMM's LDS tiles: fewer, larger staging rounds, each a memory
// round trip before its matrix-core steps.
constexpr int FDM_KC = 16;

// One 16 x 16 x 4 matrix-core step, D = A B + D, in FP32 or FP64: lane l holds A[l & 15][l >> 4]
// and B[l >> 4][l & 15]; the 4 outputs of a lane are D[i][l & 15] with i = 4 (l >> 4) + r in
// FP32 and i = 4 r + (l >> 4) in FP64 (both checked on the device).
template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct MfmaAcc<double> { typedef double type __attribute__((ext_vector_type(4))); };
__device__ __forceinline__ MfmaAcc<float>::type mfma_16x16x4(float a, float b, MfmaAcc<float>::type c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ MfmaAcc<double>::type mfma_16x16x4(double a, double b, MfmaAcc<double>::type c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 64 x 64 tile of C per 256-thread workgroup, K in steps of FDM_KC through LDS, each wave a 32 x 32
// quarter on the matrix cores (2 x 2 blocks of 16 x 16), accumulated in T. grid = (N tiles,
// M tiles, 3 x batches).
// The operand loads of a K step are issued together from clamped addresses (no load behind a
// branch; what falls outside the operands is replaced by zero when it is written to LDS), and
// the next step's loads are in flight while the matrix cores work on the current one, so a tile
// costs about one memory round trip per K step rather than one per load.
template <typename T>
__device__ __forceinline__ void fdm_load_L(const FdmGemm &g, int c, int mt, int k0, const T *__restrict__ Lb,
                                           T (&lv)[FDM_KC / 4]) {
    const int M = g.M[c], K = g.K[c], t = threadIdx.x;
    const int64_t lrs = g.lrs[c], lcs = g.lcs[c];
#pragma unroll
    for (int q = 0; q < FDM_KC / 4; ++q) {   // L, 64 rows x FDM_KC: lanes along the rows
        const int row = mt + (t & 63), col = k0 + (t >> 6) + 4 * q;
        lv[q] = Lb[min(row, M - 1) * lrs + min(col, K - 1) * lcs];
    }
}

// R, FDM_KC x 64 columns, lanes along K: read as stored
template <typename T>
__device__ __forceinline__ void fdm_load_R(const FdmGemm &g, int c, int nt, int k0, const T *__restrict__ Rb,
                                           T (&rv)[FDM_KC / 4]) {
    const int N = g.N[c], K = g.K[c], t = threadIdx.x;
    const int rk = min(k0 + t % FDM_KC, K - 1);
    const int64_t rrs = g.rrs[c], rcs = g.rcs[c];
#pragma unroll
    for (int q = 0; q < FDM_KC / 4; ++q) rv[q] = Rb[rk * rrs + min(nt + t / FDM_KC + (256 / FDM_KC) * q, N - 1) * rcs];
}

// ... or gathered from the DOF vector r: the CP ids of a step are loaded one step ahead of its
// values, so the dependent load does not stall the step.
__device__ __forceinline__ void fdm_load_ids(const FdmGemm &g, const FdmArgs &fa, int c, int nt, int k0,
                                             int32_t (&id)[FDM_KC / 4]) {
    const int N = g.N[c], K = g.K[c], t = threadIdx.x;
    const int rk = min(k0 + t % FDM_KC, K - 1), m1c = fa.m[c][1];
#pragma unroll
    for (int q = 0; q < FDM_KC / 4; ++q) {
        const int col = min(nt + t / FDM_KC + (256 / FDM_KC) * q, N - 1);
        const int j = col % m1c, k = col / m1c;
        const int64_t gp = (fa.lo[c][0] + rk) +
                           (int64_t)fa.n[0] * ((fa.lo[c][1] + j) + (int64_t)fa.n[1] * (fa.lo[c][2] + k));
        id[q] = fa.cpmap[gp];
    }
}

template <typename T>
__device__ __forceinline__ void fdm_load_gathered(int c, const T *__restrict__ r, const int32_t (&id)[FDM_KC / 4],
                                                  T (&rv)[FDM_KC / 4]) {
#pragma unroll
    for (int q = 0; q < FDM_KC / 4; ++q) rv[q] = r[3 * (int64_t)id[q] + c];
}

// The step's operands into LDS, zero outside L (M x K) and R (K x N).
template <typename T>
__device__ __forceinline__ void fdm_store_step(int M, int N, int K, int mt, int nt, int k0, const T (&lv)[FDM_KC / 4],
                                               const T (&rv)[FDM_KC / 4], T (*Ls)[65], T (*Rs)[65]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < FDM_KC / 4; ++q) {
        const int kk = (t >> 6) + 4 * q;
        Ls[kk][t & 63] = (mt + (t & 63) < M && k0 + kk < K) ? lv[q] : T(0);
    }
#pragma unroll
    for (int q = 0; q < FDM_KC / 4; ++q) {
        const int kk = t % FDM_KC, nn = t / FDM_KC + (256 / FDM_KC) * q;
        Rs[kk][nn] = (k0 + kk < K && nt + nn < N) ? rv[q] : T(0);
    }
}

// One 64 x 64 tile (rows from mt, columns from nt) of component c, batch z; every thread of the
// workgroup calls it. GATHER = g.gather, a template argument so that the loads in flight across
// a K step are not merged from two code paths (which would wait for them).
template <typename T, bool GATHER>
__device__ __forceinline__ void fdm_gemm_tile(const FdmGemm &g, const FdmArgs &fa, int c, int z, int mt, int nt,
                                              const T *__restrict__ L, const T *__restrict__ R, T *__restrict__ Cout,
                                              const T *__restrict__ invD, const T *__restrict__ r, T (*Ls)[65],
                                              T (*Rs)[65]) {
    const int M = g.M[c], N = g.N[c], K = g.K[c];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int rbase = 32 * (wave >> 1), cbase = 32 * (wave & 1);
    const T *Lb = L + g.loff[c] + z * g.lbs[c];
    const T *Rb = R + g.roff[c] + z * g.rbs[c];
    typename MfmaAcc<T>::type acc[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = typename MfmaAcc<T>::type{T(0), T(0), T(0), T(0)};
    T lv[FDM_KC / 4], rv[FDM_KC / 4];
    int32_t idn[FDM_KC / 4];
    if (K > 0) {
        fdm_load_L<T>(g, c, mt, 0, Lb, lv);
        if (GATHER) {
            fdm_load_ids(g, fa, c, nt, 0, idn);
            fdm_load_gathered<T>(c, r, idn, rv);
            if (FDM_KC < K) fdm_load_ids(g, fa, c, nt, FDM_KC, idn);
        } else {
            fdm_load_R<T>(g, c, nt, 0, Rb, rv);
        }
    }
    for (int k0 = 0; k0 < K; k0 += FDM_KC) {
        fdm_store_step<T>(M, N, K, mt, nt, k0, lv, rv, Ls, Rs);
        __syncthreads();
        if (k0 + FDM_KC < K) {   // the next step's operands
            fdm_load_L<T>(g, c, mt, k0 + FDM_KC, Lb, lv);
            if (GATHER) {
                fdm_load_gathered<T>(c, r, idn, rv);
                if (k0 + 2 * FDM_KC < K) fdm_load_ids(g, fa, c, nt, k0 + 2 * FDM_KC, idn);
            } else {
                fdm_load_R<T>(g, c, nt, k0 + FDM_KC, Rb, rv);
            }
        }
        // wave w: rows 32 (w / 2) + 16 rb, columns 32 (w % 2) + 16 cb, MFMA K steps of 4
#pragma unroll
        for (int k4 = 0; k4 < FDM_KC; k4 += 4) {
            T a[2], b[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) a[rb] = Ls[k4 + (lane >> 4)][rbase + 16 * rb + (lane & 15)];
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) b[cb] = Rs[k4 + (lane >> 4)][cbase + 16 * cb + (lane & 15)];
            // C^T = R^T L^T on the matrix cores, so that a lane's outputs share a row of C and
            // consecutive lanes hold consecutive rows (C's rows are its contiguous index)
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = mfma_16x16x4(b[cb], a[rb], acc[rb][cb]);
        }
        __syncthreads();
    }
    // accumulator (rb, cb), register r of lane l: tile row 16 rb + (l & 15), column 16 cb + i, with
    // i = 4 (l >> 4) + r (FP32) or 4 r + (l >> 4) (FP64); the scaling loads are issued together
    // (clamped addresses) before the stores
    const int64_t cbs = g.cbs[c], crs = g.crs[c], ccs = g.ccs[c];
    T sc[2][2][4];
    const bool scale = g.scale;
    if (scale) {
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = sizeof(T) == 4 ? 4 * (lane >> 4) + r : 4 * r + (lane >> 4);
                    const int row = mt + rbase + 16 * rb + (lane & 15), col = nt + cbase + 16 * cb + i;
                    sc[rb][cb][r] = invD[fa.doff[c] + z * cbs + min(row, M - 1) * crs + min(col, N - 1) * ccs];
                }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = sizeof(T) == 4 ? 4 * (lane >> 4) + r : 4 * r + (lane >> 4);
                const int row = mt + rbase + 16 * rb + (lane & 15), col = nt + cbase + 16 * cb + i;
                if (row < M && col < N)
                    Cout[g.coff[c] + z * cbs + row * crs + col * ccs] =
                        scale ? acc[rb][cb][r] * sc[rb][cb][r] : acc[rb][cb][r];
            }
}

template <typename T>
__global__ void __launch_bounds__(256) fdm_gemm_kernel(FdmGemm g, FdmArgs fa, const T *__restrict__ L,
                                                       const T *__restrict__ R, T *__restrict__ Cout,
                                                       const T *__restrict__ invD, const T *__restrict__ r,
                                                       const int32_t *__restrict__ pred) {
    if (pred && *pred == 0) return;
    const int c = blockIdx.z % 3, z = blockIdx.z / 3;
    if (z >= g.nbatch[c]) return;
    const int mt = blockIdx.y * 64, nt = blockIdx.x * 64;
    if (mt >= g.M[c] || nt >= g.N[c]) return;
    __shared__ T Ls[FDM_KC][65];
    __shared__ T Rs[FDM_KC][65];
    if (g.gather)
        fdm_gemm_tile<T, true>(g, fa, c, z, mt, nt, L, R, Cout, invD, r, Ls, Rs);
    else
        fdm_gemm_tile<T, false>(g, fa, c, z, mt, nt, L, R, Cout, invD, r, Ls, Rs);
}

// GEMM descriptor of the mode product along `axis` with A[a][b] = P[b m + a] (transposed_store)
// or P[a m + b]; the tensor operands of component c start at c maxtot.
static FdmGemm fdm_gemm_op(const FdmArgs &fa, int axis, bool transposed_store) {
    FdmGemm g{};
    for (int c = 0; c < 3; ++c) {
        const int64_t m0 = fa.m[c][0], m1 = fa.m[c][1], m2 = fa.m[c][2], m = fa.m[c][axis];
        const int64_t toff = c * fa.maxtot, moff = fa.uoff[c][axis];
        const int64_t ars = transposed_store ? 1 : m, acs = transposed_store ? m : 1;   // A(a, b)
        g.coff[c] = toff;
        if (axis == 0) {
            g.M[c] = (int32_t)m0; g.N[c] = (int32_t)(m1 * m2); g.K[c] = (int32_t)m0; g.nbatch[c] = 1;
            g.loff[c] = moff; g.lrs[c] = ars; g.lcs[c] = acs; g.lbs[c] = 0;
            g.roff[c] = toff; g.rrs[c] = 1; g.rcs[c] = m0; g.rbs[c] = 0;
            g.crs[c] = 1; g.ccs[c] = m0; g.cbs[c] = 0;
        } else if (axis == 1) {
            g.M[c] = (int32_t)m0; g.N[c] = (int32_t)m1; g.K[c] = (int32_t)m1; g.nbatch[c] = (int32_t)m2;
            g.loff[c] = toff; g.lrs[c] = 1; g.lcs[c] = m0; g.lbs[c] = m0 * m1;
            g.roff[c] = moff; g.rrs[c] = acs; g.rcs[c] = ars; g.rbs[c] = 0;   // R(b, j) = A(j, b)
            g.crs[c] = 1; g.ccs[c] = m0; g.cbs[c] = m0 * m1;
        } else {
            g.M[c] = (int32_t)(m0 * m1); g.N[c] = (int32_t)m2; g.K[c] = (int32_t)m2; g.nbatch[c] = 1;
            g.loff[c] = toff; g.lrs[c] = 1; g.lcs[c] = m0 * m1; g.lbs[c] = 0;
            g.roff[c] = moff; g.rrs[c] = acs; g.rcs[c] = ars; g.rbs[c] = 0;
            g.crs[c] = 1; g.ccs[c] = m0 * m1; g.cbs[c] = 0;
        }
        (void)m;
    }
    return g;
}

template <typename T>
static void fdm_product(const FdmArgs &fa, int axis, const T *P, bool transposed_store, const T *in, T *out,
                        const T *r, const T *invD, hipStream_t st, const int32_t *pred) {
    FdmGemm g = fdm_gemm_op(fa, axis, transposed_store);
    g.gather = r != nullptr;
    g.scale = invD != nullptr;
    int maxM = 0, maxN = 0, maxB = 0;
    for (int c = 0; c < 3; ++c) {
        maxM = std::max(maxM, g.M[c]);
        maxN = std::max(maxN, g.N[c]);
        maxB = std::max(maxB, g.nbatch[c]);
    }
    const T *Lp = axis == 0 ? P : in;
    const T *Rp = axis == 0 ? in : P;
    dim3 grid((unsigned)((maxN + 63) / 64), (unsigned)((maxM + 63) / 64), (unsigned)(3 * maxB));
    hipLaunchKernelGGL(fdm_gemm_kernel<T>, grid, dim3(256), 0, st, g, fa, Lp, Rp, out, invD, r, pred);
}

// z at grid point g (0 outside the component's free box and on masked dofs), its terms of r.z
// added to acc in component order. The grid has fewer than 2^31 points (the set-up checks), so its index arithmetic is
// 32-bit; every load is issued from a clamped address and the value selected afterwards, so the
// three components' loads are in flight together.
template <typename T>
__device__ __forceinline__ void fdm_scatter_point(const FdmArgs &fa, int32_t g, const T *__restrict__ w,
                                                  const T *__restrict__ r, const uint8_t *__restrict__ mask,
                                                  T *__restrict__ z, double &acc) {
    const uint32_t n0 = fa.n[0], n1 = fa.n[1], q = (uint32_t)g / n0;
    const int i = (int)((uint32_t)g - q * n0), j = (int)(q % n1), k = (int)(q / n1);
    const int64_t cp = fa.cpmap[g];
    T wv[3], rv[3];
    uint8_t mk[3];
    bool in[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int ii = i - fa.lo[c][0], jj = j - fa.lo[c][1], kk = k - fa.lo[c][2];
        in[c] = ii >= 0 && ii < fa.m[c][0] && jj >= 0 && jj < fa.m[c][1] && kk >= 0 && kk < fa.m[c][2];
        const int ic = max(0, min(ii, fa.m[c][0] - 1)), jc = max(0, min(jj, fa.m[c][1] - 1)),
                  kc = max(0, min(kk, fa.m[c][2] - 1));
        wv[c] = w[c * fa.maxtot + ic + (int64_t)fa.m[c][0] * (jc + (int64_t)fa.m[c][1] * kc)];
        mk[c] = mask[3 * cp + c];
        rv[c] = r[3 * cp + c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T v = (in[c] && mk[c]) ? wv[c] : T(0);
        z[3 * cp + c] = v;
        acc += (double)rv[c] * (double)v;
    }
}

// z on every CP (0 outside the free boxes and on masked dofs); r.z reduced in the same launch
// and handed to rz_tail (the session's scalars: first direction, or beta and the iteration count)
template <typename T>
__global__ void __launch_bounds__(256) fdm_scatter_kernel(FdmArgs fa, int64_t ncp, const T *__restrict__ w,
                                                          const T *__restrict__ r, const uint8_t *__restrict__ mask,
                                                          T *__restrict__ z, double *__restrict__ part, RzTail tail,
                                                          const int32_t *__restrict__ pred) {
    if (pred && *pred == 0) return;
    __shared__ double sh[4];
    double acc = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ncp; g += (int64_t)gridDim.x * blockDim.x)
        fdm_scatter_point<T>(fa, (int32_t)g, w, r, mask, z, acc);
    double tot;
    if (last_block_total(block_sum256(acc, sh), part, tail.flags + F_COUNTER, sh, tot) && threadIdx.x == 0)
        rz_tail(tail, tot);
}

// 7 launches: three forward mode products (the first gathers r, the third scales by 1/D), three
// backward ones and the scatter with r.z.
template <typename T>
static int fdm_apply_t(const yeti_fdm *f, const T *U, const T *Ut, const T *invD, T *wa, T *wb, const T *r,
                       const uint8_t *mask, T *z, double *part, int nblocks, const RzTail &tail, hipStream_t st,
                       const int32_t *pred) {
    const FdmArgs &fa = f->fa;
    const T *nul = nullptr;
    // forward: A = U^T along axes 0, 1, 2 (U read transposed, or Ut read as stored)
    fdm_product<T>(fa, 0, U, true, nul, wb, r, nul, st, pred);
    fdm_product<T>(fa, 1, Ut, false, wb, wa, nul, nul, st, pred);
    fdm_product<T>(fa, 2, Ut, false, wa, wb, nul, invD, st, pred);
    // backward: A = U
    fdm_product<T>(fa, 0, Ut, true, wb, wa, nul, nul, st, pred);
    fdm_product<T>(fa, 1, U, false, wa, wb, nul, nul, st, pred);
    fdm_product<T>(fa, 2, U, false, wb, wa, nul, nul, st, pred);
    hipLaunchKernelGGL(fdm_scatter_kernel<T>, dim3(nblocks), dim3(256), 0, st, fa, (int64_t)f->ncp, wa, r, mask, z,
                       part, tail, pred);
    YETI_HIP(hipGetLastError());
    return 0;
}

int fdm_apply(const yeti_fdm *f, const double *r, const uint8_t *mask, double *z, double *part, int nblocks,
              const RzTail &tail, hipStream_t st, const int32_t *pred) {
    return fdm_apply_t<double>(f, f->U64.p, f->Ut64.p, f->invD64.p, f->wa64.p, f->wb64.p, r, mask, z, part, nblocks,
                               tail, st, pred);
}
int fdm_apply(const yeti_fdm *f, const float *r, const uint8_t *mask, float *z, double *part, int nblocks,
              const RzTail &tail, hipStream_t st, const int32_t *pred) {
    return fdm_apply_t<float>(f, f->U32.p, f->Ut32.p, f->invD32.p, f->wa32.p, f->wb32.p, r, mask, z, part, nblocks,
                              tail, st, pred);
}

// Host set-up. mask_host: free-dof mask (3 nb_cp). precision 32 or 64. Device memory comes from
// the current allocation scope.
// Upper bound of the device bytes of fdm_create (no face removed), for counting set-ups.
size_t fdm_bytes_bound(const yeti_mf *m, int precision) {
    if (m->patches.size() != 1) return 0;
    const MfPatchDev &pd = *m->patches[0];
    const size_t w = precision == 64 ? 8 : 4;
    const size_t n0 = pd.mp.ncp[0], n1 = pd.mp.ncp[1], n2 = pd.mp.ncp[2], tot = n0 * n1 * n2;
    auto al = [](size_t b) { return (b + 255) / 256 * 256 + 256; };
    const size_t mats = 3 * (n0 * n0 + n1 * n1 + n2 * n2) * w;
    return 2 * al(mats) + al(3 * tot * w) + 2 * al(3 * tot * w);
}

int fdm_create(const yeti_mf *m, const uint8_t *mask_host, int precision, hipStream_t st, yeti_fdm **out) {
    *out = nullptr;
    if (m->patches.size() != 1)
        return set_error(YETI_ERR_UNSUPPORTED_ELEMENT, "fast diagonalisation: single-patch handles only");
    const MfPatchDev &pd = *m->patches[0];
    const int p = pd.p, ng = pd.ng, np1 = p + 1;
    auto *f = new yeti_fdm();
    auto fail = [&](int code) {
        delete f;
        return code;
    };
    FdmArgs &fa = f->fa;
    for (int d = 0; d < 3; ++d) fa.n[d] = pd.mp.ncp[d];
    const int64_t ngrid = (int64_t)fa.n[0] * fa.n[1] * fa.n[2];
    if (ngrid > INT32_MAX)
        return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "fast diagonalisation: more than 2^31 - 1 control points"));
    f->ncp = (int32_t)ngrid;
    // fixed dofs: whole faces per component
    for (int c = 0; c < 3; ++c) {
        bool face_fixed[3][2];
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                bool all = true;
                for (int64_t g = 0; g < ngrid && all; ++g) {
                    const int idx[3] = {(int)(g % fa.n[0]), (int)((g / fa.n[0]) % fa.n[1]), (int)(g / ((int64_t)fa.n[0] * fa.n[1]))};
                    if (idx[d] != (side ? fa.n[d] - 1 : 0)) continue;
                    if (mask_host[3 * (int64_t)pd.hcpmap[g] + c]) all = false;
                }
                face_fixed[d][side] = all;
            }
        for (int d = 0; d < 3; ++d) {
            fa.lo[c][d] = face_fixed[d][0] ? 1 : 0;
            fa.m[c][d] = fa.n[d] - fa.lo[c][d] - (face_fixed[d][1] ? 1 : 0);
            if (fa.m[c][d] < 1) return fail(set_error(YETI_ERR_UNSUPPORTED_ELEMENT, "fast diagonalisation: no free dofs"));
        }
        for (int64_t g = 0; g < ngrid; ++g) {   // every fixed dof must lie on a fixed face
            const int idx[3] = {(int)(g % fa.n[0]), (int)((g / fa.n[0]) % fa.n[1]), (int)(g / ((int64_t)fa.n[0] * fa.n[1]))};
            bool inside = true;
            for (int d = 0; d < 3; ++d) inside &= idx[d] >= fa.lo[c][d] && idx[d] < fa.lo[c][d] + fa.m[c][d];
            if (inside && !mask_host[3 * (int64_t)pd.hcpmap[g] + c])
                return fail(set_error(YETI_ERR_UNSUPPORTED_ELEMENT,
                                      "fast diagonalisation: fixed dofs are not whole patch faces"));
        }
        fa.tot[c] = (int64_t)fa.m[c][0] * fa.m[c][1] * fa.m[c][2];
    }
    fa.maxtot = std::max(fa.tot[0], std::max(fa.tot[1], fa.tot[2]));
    // 1D Gram matrices per direction on the whole index range
    std::vector<double> K1[3], M1[3];
    for (int d = 0; d < 3; ++d) {
        const int n = fa.n[d], N = pd.mp.N[d];
        K1[d].assign((size_t)n * n, 0.0);
        M1[d].assign((size_t)n * n, 0.0);
        for (int e = 0; e < N; ++e) {
            const double h = pd.hhalf[3 * e + d];
            for (int g = 0; g < ng; ++g) {
                const size_t o = ((size_t)(pd.mp.boff[d] + e) * ng + g) * np1;
                for (int r1 = 0; r1 < np1; ++r1)
                    for (int r2 = 0; r2 < np1; ++r2) {
                        const int i = e + r1, j = e + r2;   // basis span - p + r, span = p + e
                        M1[d][(size_t)i * n + j] += h * pd.hgw[g] * pd.hB[o + r1] * pd.hB[o + r2];
                        K1[d][(size_t)i * n + j] += h * pd.hgw[g] * pd.hdB[o + r1] * pd.hdB[o + r2];
                    }
            }
        }
    }
    // patch-mean Jacobian from the face centroids of the control net
    const double *X = nullptr;
    std::vector<double> Xh((size_t)3 * m->nb_cp);
    if (hipMemcpyAsync(Xh.data(), m->coords.p, Xh.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "fast diagonalisation: coordinate copy failed"));
    X = Xh.data();
    double J[3][3];
    for (int d = 0; d < 3; ++d) {
        double c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
        int64_t n0 = 0, n1 = 0;
        for (int64_t g = 0; g < ngrid; ++g) {
            const int idx[3] = {(int)(g % fa.n[0]), (int)((g / fa.n[0]) % fa.n[1]), (int)(g / ((int64_t)fa.n[0] * fa.n[1]))};
            const double *x = X + 3 * (int64_t)pd.hcpmap[g];
            if (idx[d] == 0) {
                for (int k = 0; k < 3; ++k) c0[k] += x[k];
                ++n0;
            }
            if (idx[d] == fa.n[d] - 1) {
                for (int k = 0; k < 3; ++k) c1[k] += x[k];
                ++n1;
            }
        }
        const double plen = pd.hU[d].back() - pd.hU[d].front();
        for (int k = 0; k < 3; ++k) J[k][d] = (c1[k] / n1 - c0[k] / n0) / plen;
    }
    const double m00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], m01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                 m02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * m00 + J[0][1] * m01 + J[0][2] * m02;
    if (!(std::fabs(det) > 0.0)) return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "fast diagonalisation: degenerate patch"));
    double Ji[3][3];   // Ji[d][k] = d u_d / d x_k
    Ji[0][0] = m00 / det;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    Ji[1][0] = m01 / det;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    Ji[2][0] = m02 / det;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    const double lam = pd.mp.lam, mu = pd.mp.mu;
    // per (component, direction): restricted Gram matrices, generalized eigenpairs
    std::vector<double> Uall, Utall, Dall;
    std::vector<double> eig[3][3];
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
            const int n = fa.n[d], mm = fa.m[c][d], lo = fa.lo[c][d];
            std::vector<double> Ks((size_t)mm * mm), Ms((size_t)mm * mm);
            for (int i = 0; i < mm; ++i)
                for (int j = 0; j < mm; ++j) {
                    Ks[(size_t)i * mm + j] = K1[d][(size_t)(lo + i) * n + lo + j];
                    Ms[(size_t)i * mm + j] = M1[d][(size_t)(lo + i) * n + lo + j];
                }
            std::vector<double> U;
            if (!gen_eigen(mm, Ks, Ms, eig[c][d], U))
                return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "fast diagonalisation: singular mass matrix"));
            fa.uoff[c][d] = (int64_t)Uall.size();
            Uall.insert(Uall.end(), U.begin(), U.end());
            for (int i = 0; i < mm; ++i)
                for (int j = 0; j < mm; ++j) Utall.push_back(U[(size_t)j * mm + i]);
        }
    for (int c = 0; c < 3; ++c) {
        double alpha[3];
        for (int d = 0; d < 3; ++d) {
            const double gdd = Ji[d][0] * Ji[d][0] + Ji[d][1] * Ji[d][1] + Ji[d][2] * Ji[d][2];
            alpha[d] = std::fabs(det) * (mu * gdd + (lam + mu) * Ji[d][c] * Ji[d][c]);
        }
        fa.doff[c] = (int64_t)Dall.size();
        for (int k = 0; k < fa.m[c][2]; ++k)
            for (int j = 0; j < fa.m[c][1]; ++j)
                for (int i = 0; i < fa.m[c][0]; ++i)
                {
                    const double D = alpha[0] * eig[c][0][i] + alpha[1] * eig[c][1][j] + alpha[2] * eig[c][2][k];
                    Dall.push_back(D > 1e-300 ? 1.0 / D : 0.0);   // 0 on a rigid mode of an unclamped component
                }
    }
    fa.cpmap = pd.cpmap.p;
    hipError_t e = hipSuccess;
    if (precision == 64) {
        e = f->U64.upload(Uall.data(), Uall.size());
        if (e == hipSuccess) e = f->Ut64.upload(Utall.data(), Utall.size());
        if (e == hipSuccess) e = f->invD64.upload(Dall.data(), Dall.size());
        if (e == hipSuccess) e = f->wa64.alloc(3 * (size_t)fa.maxtot);
        if (e == hipSuccess) e = f->wb64.alloc(3 * (size_t)fa.maxtot);
    } else {
        std::vector<float> u32(Uall.begin(), Uall.end()), ut32(Utall.begin(), Utall.end()), d32(Dall.begin(), Dall.end());
        e = f->U32.upload(u32.data(), u32.size());
        if (e == hipSuccess) e = f->Ut32.upload(ut32.data(), ut32.size());
        if (e == hipSuccess) e = f->invD32.upload(d32.data(), d32.size());
        if (e == hipSuccess) e = f->wa32.alloc(3 * (size_t)fa.maxtot);
        if (e == hipSuccess) e = f->wb32.alloc(3 * (size_t)fa.maxtot);
    }
    if (e != hipSuccess) return fail(set_error(YETI_ERR_BUDGET, "fast diagonalisation: device memory"));
    *out = f;
    return 0;
}

void fdm_destroy(yeti_fdm *f) { delete f; }

// ---- batches: the operator and the preconditioner of many sessions per launch -------------
// Bricks of degree P of every member: the iteration apply p -> brick partials in the batch
// precision T (active members), or (RPL) the FP64 apply of the replacement x -> FP64 partials
// (members with a replacement due).
template <typename T, int P, bool RPL>
__global__ void __launch_bounds__((Sweep<P, sizeof(T)>::NT)) __attribute__((amdgpu_waves_per_eu(Sweep<P, sizeof(T)>::WPE)))
bt_sweep_kernel(BatchDev bd, int list) {
    __shared__ SweepSmem<T, P> sm;
    constexpr int pi = sizeof(T) == 8 ? 0 : 1;
    const int32_t cnt = bd.count[list];
    const BTask *tl = bd.list[list];
    for (int32_t t = blockIdx.x; t < cnt; t += gridDim.x) {
        const BTask tk = tl[t];
        const BReq &R = bd.req[tk.req];
        if (R.flags[RPL ? F_UPDATE : F_ACTIVE] == 0) continue;   // the same for every thread
        __syncthreads();   // the previous task's last reads of the shared tables are done
        const T *x = RPL ? (const T *)R.x : (const T *)R.vp;
        mf_sweep_body<T, P>(sm, R.mp, (const T *)R.geo[pi], x, (const T *)R.w[pi], (T *)R.part[pi], tk.a);
    }
}

// Node sums of the bricks: q (iteration) or y64 (RPL), BATCH_GATHER_BLOCK control points a task.
template <typename T, bool RPL>
__global__ void __launch_bounds__(256) bt_gather_kernel(BatchDev bd) {
    constexpr int pi = sizeof(T) == 8 ? 0 : 1;
    const int32_t cnt = bd.count[BT_GATHER];
    const BTask *tl = bd.list[BT_GATHER];
    for (int32_t t = blockIdx.x; t < cnt; t += gridDim.x) {
        const BTask tk = tl[t];
        const BReq &R = bd.req[tk.req];
        if (R.flags[RPL ? F_UPDATE : F_ACTIVE] == 0) continue;
        const int64_t cp = (int64_t)tk.a * BATCH_GATHER_BLOCK + threadIdx.x;
        if (cp >= R.ncp) continue;
        T *y = RPL ? (T *)R.y64 : (T *)R.vq;
        mf_gather_point<T, 3>(R.mp, R.br[pi], R.bn[pi], R.nbrick[pi], (const T *)R.part[pi], y, 0, cp);
    }
}

// Tiles of the preconditioner's mode product `prod` of every active member.
template <typename T>
__global__ void __launch_bounds__(256) bt_fdm_gemm_kernel(BatchDev bd, int prod) {
    __shared__ T Ls[FDM_KC][65];
    __shared__ T Rs[FDM_KC][65];
    const int list = BT_GEMM0 + prod;
    const int32_t cnt = bd.count[list];
    const BTask *tl = bd.list[list];
    for (int32_t t = blockIdx.x; t < cnt; t += gridDim.x) {
        const BTask tk = tl[t];
        const BReq &R = bd.req[tk.req];
        if (R.flags[F_ACTIVE] == 0) continue;
        const FdmGemm &g = R.g[prod];
        const int c = tk.a, z = tk.b;
        const int ntn = (g.N[c] + 63) / 64;
        __syncthreads();   // the previous tile's reads of Ls / Rs are done
        const int mt = (tk.c / ntn) * 64, nt = (tk.c % ntn) * 64;
        if (g.gather)
            fdm_gemm_tile<T, true>(g, R.fa, c, z, mt, nt, (const T *)R.L[prod], (const T *)R.R[prod],
                                   (T *)R.C[prod], (const T *)R.invD, (const T *)R.vr, Ls, Rs);
        else
            fdm_gemm_tile<T, false>(g, R.fa, c, z, mt, nt, (const T *)R.L[prod], (const T *)R.R[prod],
                                    (T *)R.C[prod], (const T *)R.invD, (const T *)R.vr, Ls, Rs);
    }
}

// z from the last product (0 outside the free boxes and on masked dofs) and r.z, handed to
// rz_tail by each member's last chunk (mode 0: first direction; 1: end of an iteration).
template <typename T>
__global__ void __launch_bounds__(256) bt_fdm_scatter_kernel(BatchDev bd, int mode) {
    __shared__ double sh[4];
    const int32_t cnt = bd.count[BT_CP];
    const BTask *tl = bd.list[BT_CP];
    for (int32_t t = blockIdx.x; t < cnt; t += gridDim.x) {
        const BTask tk = tl[t];
        const BReq &R = bd.req[tk.req];
        if (R.flags[F_ACTIVE] == 0) continue;
        const FdmArgs &fa = R.fa;
        const T *w = (const T *)R.wout;
        const T *r = (const T *)R.vr;
        T *z = (T *)R.vz;
        const int64_t g0 = (int64_t)tk.a * R.cchunk, g1 = std::min<int64_t>(R.ncp, g0 + R.cchunk);
        double acc = 0.0;
        for (int64_t g = g0 + threadIdx.x; g < g1; g += blockDim.x)
            fdm_scatter_point<T>(fa, (int32_t)g, w, r, R.mask, z, acc);
        double tot;
        if (member_total(block_sum256(acc, sh), R.red, R.flags + F_COUNTER, R.ncchunk, tk.a, sh, tot) &&
            threadIdx.x == 0) {
            RzTail tail;
            tail.sc = R.sc;
            tail.flags = R.flags;
            tail.mode = mode;
            tail.max_iters = R.max_iters;
            rz_tail(tail, tot);
        }
    }
}

constexpr unsigned BT_GRID_BRICK = 1024, BT_GRID = 2048;

// `lists`: BT_BRICK1 (bricks of the batch precision) or BT_RBRICK1 (FP64 bricks of an FP32 batch)
template <typename T, bool RPL>
static int batch_apply_t(const BatchDev &bd, int lists, hipStream_t st) {
    hipLaunchKernelGGL((bt_sweep_kernel<T, 1, RPL>), dim3(BT_GRID_BRICK), dim3(Sweep<1, sizeof(T)>::NT), 0, st, bd, lists);
    hipLaunchKernelGGL((bt_sweep_kernel<T, 2, RPL>), dim3(BT_GRID_BRICK), dim3(Sweep<2, sizeof(T)>::NT), 0, st, bd,
                       lists + 1);
    hipLaunchKernelGGL((bt_sweep_kernel<T, 3, RPL>), dim3(BT_GRID_BRICK), dim3(Sweep<3, sizeof(T)>::NT), 0, st, bd,
                       lists + 2);
    hipLaunchKernelGGL((bt_sweep_kernel<T, 4, RPL>), dim3(BT_GRID_BRICK), dim3(Sweep<4, sizeof(T)>::NT), 0, st, bd,
                       lists + 3);
    hipLaunchKernelGGL((bt_gather_kernel<T, RPL>), dim3(BT_GRID), dim3(BATCH_GATHER_BLOCK), 0, st, bd);
    YETI_HIP(hipGetLastError());
    return 0;
}

int batch_mf_apply(const BatchDev &bd, int prec64, bool rpl, hipStream_t st) {
    if (rpl) return batch_apply_t<double, true>(bd, prec64 ? (int)BT_BRICK1 : (int)BT_RBRICK1, st);
    return prec64 ? batch_apply_t<double, false>(bd, BT_BRICK1, st) : batch_apply_t<float, false>(bd, BT_BRICK1, st);
}

template <typename T>
static int batch_fdm_t(const BatchDev &bd, int mode, hipStream_t st) {
    for (int prod = 0; prod < 6; ++prod)
        hipLaunchKernelGGL(bt_fdm_gemm_kernel<T>, dim3(BT_GRID), dim3(256), 0, st, bd, prod);
    hipLaunchKernelGGL(bt_fdm_scatter_kernel<T>, dim3(BT_GRID), dim3(256), 0, st, bd, mode);
    YETI_HIP(hipGetLastError());
    return 0;
}

int batch_fdm_apply(const BatchDev &bd, int prec64, int mode, hipStream_t st) {
    return prec64 ? batch_fdm_t<double>(bd, mode, st) : batch_fdm_t<float>(bd, mode, st);
}

// The six mode products of fdm_apply_t: (axis, matrix, transposed store, input, output).
int batch_fill_operator(const yeti_mf *m, const yeti_fdm *f, int prec64, BReq *R) {
    if (!m || !f) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: session without fast diagonalisation");
    if (m->patches.size() != 1 || m->zero_first)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: single-patch full-grid operators only");
    const MfPatchDev &pd = *m->patches[0];
    if (pd.p < 1 || pd.p > 4 || pd.ng != pd.p + 1)
        return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "batch: degree %d with %d points", pd.p, pd.ng);
    if (!prec64 && (!pd.geo32.p || !pd.part32.p || !m->w32.p))
        return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: FP32 operator not prepared");
    R->mp = pd.mp;
    R->p = pd.p;
    R->ng = pd.ng;
    for (int k = 0; k < 2; ++k)
        for (int d = 0; d < 3; ++d) {
            R->nbrick[k][d] = pd.mp.gnb[k][d];
            R->br[k][d] = pd.sweep_br[k][d];
            R->bn[k][d] = pd.sweep_br[k][d] + pd.p;
        }
    R->ncp = m->nb_cp;
    R->n = 3 * (int64_t)m->nb_cp;
    R->geo[0] = pd.geo64.p;
    R->geo[1] = pd.geo32.p;
    R->w[0] = m->w64.p;
    R->w[1] = m->w32.p;
    R->part[0] = pd.part64.p;
    R->part[1] = pd.part32.p;
    R->fa = f->fa;
    const void *U = prec64 ? (const void *)f->U64.p : (const void *)f->U32.p;
    const void *Ut = prec64 ? (const void *)f->Ut64.p : (const void *)f->Ut32.p;
    void *wa = prec64 ? (void *)f->wa64.p : (void *)f->wa32.p;
    void *wb = prec64 ? (void *)f->wb64.p : (void *)f->wb32.p;
    R->invD = prec64 ? (const void *)f->invD64.p : (const void *)f->invD32.p;
    if (!U || !Ut || !wa || !wb || !R->invD) return set_error(YETI_ERR_INVALID_ARGUMENT, "batch: preconditioner precision");
    struct Prod { int axis; const void *P; bool tr; const void *in; void *out; bool gather, scale; };
    const Prod prods[6] = {{0, U, true, nullptr, wb, true, false},  {1, Ut, false, wb, wa, false, false},
                           {2, Ut, false, wa, wb, false, true},     {0, Ut, true, wb, wa, false, false},
                           {1, U, false, wa, wb, false, false},     {2, U, false, wb, wa, false, false}};
    for (int k = 0; k < 6; ++k) {
        FdmGemm g = fdm_gemm_op(f->fa, prods[k].axis, prods[k].tr);
        g.gather = prods[k].gather ? 1 : 0;
        g.scale = prods[k].scale ? 1 : 0;
        R->g[k] = g;
        R->L[k] = prods[k].axis == 0 ? prods[k].P : prods[k].in;
        R->R[k] = prods[k].axis == 0 ? prods[k].in : prods[k].P;
        R->C[k] = prods[k].out;
    }
    R->wout = wa;
    return 0;
}

// Sweep brick grid, brick edges, geometry size and partial-sum size of a degree-p patch.
template <int P, int W>
static void sweep_dims_t(const int32_t N[3], int32_t gnb[3], int32_t br[3], int64_t &geo_n, int64_t &part_n) {
    using S = Sweep<P, W>;
    sweep_grid<P, W>(N, gnb);
    br[0] = S::BRX;
    br[1] = S::BRY;
    br[2] = S::BZ;
    geo_n = sweep_geo_size<P, W>(gnb);
    part_n = (int64_t)gnb[0] * gnb[1] * gnb[2] * S::PART;
}
// W: bytes per value (8 FP64, 4 FP32)
static void sweep_dims(int p, int W, const int32_t N[3], int32_t gnb[3], int32_t br[3], int64_t &geo_n,
                       int64_t &part_n) {
    switch (p * 16 + W) {
    case 1 * 16 + 8: sweep_dims_t<1, 8>(N, gnb, br, geo_n, part_n); break;
    case 2 * 16 + 8: sweep_dims_t<2, 8>(N, gnb, br, geo_n, part_n); break;
    case 3 * 16 + 8: sweep_dims_t<3, 8>(N, gnb, br, geo_n, part_n); break;
    case 4 * 16 + 8: sweep_dims_t<4, 8>(N, gnb, br, geo_n, part_n); break;
    case 1 * 16 + 4: sweep_dims_t<1, 4>(N, gnb, br, geo_n, part_n); break;
    case 2 * 16 + 4: sweep_dims_t<2, 4>(N, gnb, br, geo_n, part_n); break;
    case 3 * 16 + 4: sweep_dims_t<3, 4>(N, gnb, br, geo_n, part_n); break;
    default: sweep_dims_t<4, 4>(N, gnb, br, geo_n, part_n); break;
    }
}

}  // namespace yeti

using namespace yeti;

extern "C" {

static int32_t mf_create(const yeti_system_inputs *in, int32_t precisions, hipStream_t st, yeti_mf **out) {
    if (!in || !out) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (!(precisions & 3)) return set_error(YETI_ERR_INVALID_ARGUMENT, "no precision requested");
    // ids, spans and lengths checked on the host before any device call
    if (int rc0 = validate_system_inputs(in, true)) return rc0;
    auto *m = new yeti_mf();
    auto fail = [&](int code) {
        delete m;
        return code;
    };
    m->nb_cp = in->nb_cp;
    m->precisions = precisions;
    // per-CP weights (the inputs repeat them per element) and coordinates
    std::vector<double> wcp((size_t)in->nb_cp, 1.0);
    int64_t kv = 0, woff = 0, ieoff = 0, eloff = 0;
    for (int ip = 0; ip < in->nb_patch; ++ip) {
        int nn = in->NNODE[ip];
        for (int64_t e = 0; e < in->nb_elem_patch[ip]; ++e)
            for (int a = 0; a < nn; ++a) wcp[in->IEN[ieoff + nn * e + a] - 1] = in->weight[woff + nn * e + a];
        woff += (int64_t)nn * in->nb_elem_patch[ip];
        ieoff += (int64_t)nn * in->nb_elem_patch[ip];
    }
    if (m->coords.upload(in->COORDS3D, 3 * (size_t)in->nb_cp) != hipSuccess ||
        m->w64.upload(wcp.data(), wcp.size()) != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "matrix-free: upload failed"));
    if (precisions & 2) {
        std::vector<float> w32(wcp.begin(), wcp.end());
        if (m->w32.upload(w32.data(), w32.size()) != hipSuccess)
            return fail(set_error(YETI_ERR_DEVICE, "matrix-free: upload failed"));
    }
    woff = ieoff = 0;
    for (int ip = 0; ip < in->nb_patch; ++ip) {
        auto *pd = new MfPatchDev();
        m->patches.push_back(pd);
        const int p0 = in->Jpqr[3 * ip], p1 = in->Jpqr[3 * ip + 1], p2 = in->Jpqr[3 * ip + 2];
        const int nn = in->NNODE[ip];
        const int ne = in->nb_elem_patch[ip];
        int nkv[3] = {in->Nkv[3 * ip], in->Nkv[3 * ip + 1], in->Nkv[3 * ip + 2]};
        const double *U[3] = {in->Ukv + kv, in->Ukv + kv + nkv[0], in->Ukv + kv + nkv[0] + nkv[1]};
        kv += nkv[0] + nkv[1] + nkv[2];
        if (p0 != p1 || p1 != p2) return fail(set_error(YETI_ERR_UNSUPPORTED_DEGREE, "matrix-free: degrees differ per direction"));
        int nbint = in->NBINT[ip];
        int ng = (int)std::pow((double)nbint, 1.0 / 3.0);
        if (ng * ng * ng < nbint) ++ng;
        pd->p = p0;
        pd->ng = ng;
        if (!((p0 == 1 && ng == 2) || (p0 == 2 && ng == 3) || (p0 == 3 && ng == 4) || (p0 == 4 && ng == 5)))
            return fail(set_error(YETI_ERR_UNSUPPORTED_DEGREE, "matrix-free: p=%d with %d Gauss points", p0, ng));
        // element grid from the spans; require every span between distinct knots to hold one element
        int N[3], ncp[3];
        for (int d = 0; d < 3; ++d) {
            ncp[d] = nkv[d] - p0 - 1;
            N[d] = ncp[d] - p0;
            for (int s = p0; s < p0 + N[d]; ++s)
                if (!(U[d][s + 1] > U[d][s]))
                    return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "matrix-free: repeated interior knots are not supported"));
        }
        if ((int64_t)N[0] * N[1] * N[2] != ne)
            return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "matrix-free: patch %d is not a full element grid", ip + 1));
        std::vector<int32_t> cpmap((size_t)ncp[0] * ncp[1] * ncp[2], -1);
        for (int e = 0; e < ne; ++e) {
            int ed[3];
            for (int d = 0; d < 3; ++d) ed[d] = in->Nijk[3 * (int64_t)(eloff + e) + d] - 1 - p0;
            for (int a = 0; a < nn; ++a) {
                int i = a % (p0 + 1), j = (a / (p0 + 1)) % (p0 + 1), k = a / ((p0 + 1) * (p0 + 1));
                // local node a <-> basis index p - i in each direction (shap.f90:94-119)
                int64_t c = (ed[0] + p0 - i) + (int64_t)ncp[0] * ((ed[1] + p0 - j) + (int64_t)ncp[1] * (ed[2] + p0 - k));
                cpmap[c] = in->IEN[ieoff + (int64_t)nn * e + a] - 1;
            }
        }
        for (int32_t v : cpmap)
            if (v < 0) return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "matrix-free: connectivity is not a tensor grid"));
        // 1D tables per direction and element, and half spans
        double gx[MAX_NG], gwt[MAX_NG];
        gauss_rule_host(ng, gx, gwt);
        for (int i = 0; i < ng; ++i) {
            pd->g.x[i] = gx[i];
            pd->g.w[i] = gwt[i];
        }
        const int np1 = p0 + 1;
        std::vector<double> Bt, dBt;
        int nmax = std::max(N[0], std::max(N[1], N[2]));
        std::vector<double> half((size_t)nmax * 3, 0.0);
        for (int d = 0; d < 3; ++d) {
            pd->mp.boff[d] = (int32_t)(Bt.size() / (ng * np1));
            for (int e = 0; e < N[d]; ++e) {
                int s = p0 + e;
                double u0 = U[d][s], u1 = U[d][s + 1];
                half[3 * e + d] = 0.5 * (u1 - u0);
                for (int g = 0; g < ng; ++g) {
                    double u = ((u1 - u0) * gx[g] + (u1 + u0)) * 0.5;
                    double Nv[16], dNv[16];
                    host_bspline(s, p0, u, U[d], Nv, dNv);
                    for (int r = 0; r < np1; ++r) {
                        Bt.push_back(Nv[r]);
                        dBt.push_back(dNv[r]);
                    }
                }
            }
        }
        pd->hB = Bt;
        pd->hdB = dBt;
        pd->hhalf = half;
        pd->hgw.assign(gwt, gwt + ng);
        pd->hcpmap = cpmap;
        for (int d = 0; d < 3; ++d) pd->hU[d].assign(U[d], U[d] + nkv[d]);
        for (int d = 0; d < 3; ++d) {
            pd->mp.N[d] = N[d];
            pd->mp.ncp[d] = ncp[d];
            pd->nbrick[0][d] = (N[d] + 2 * np1 - 1) / (2 * np1);
            pd->nbrick[1][d] = (N[d] + np1 - 1) / np1;
        }
        int64_t geo_n64 = 0, part_n64 = 0, geo_n32 = 0, part_n32 = 0;
        sweep_dims(p0, 8, N, pd->mp.gnb[0], pd->sweep_br[0], geo_n64, part_n64);
        sweep_dims(p0, 4, N, pd->mp.gnb[1], pd->sweep_br[1], geo_n32, part_n32);
        double E = in->MATERIAL_PROPERTIES[(int64_t)in->n_mat_props_max * ip];
        double nu = in->MATERIAL_PROPERTIES[(int64_t)in->n_mat_props_max * ip + 1];
        pd->mp.lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        pd->mp.mu = E / (2.0 * (1.0 + nu));
        hipError_t er = pd->cpmap.upload(cpmap.data(), cpmap.size());
        if (er == hipSuccess) er = pd->Btab.upload(Bt.data(), Bt.size());
        if (er == hipSuccess) er = pd->dBtab.upload(dBt.data(), dBt.size());
        if (er == hipSuccess) er = pd->half.upload(half.data(), half.size());
        if (er == hipSuccess) er = pd->nbrick_dev.upload(&pd->nbrick[0][0], 6);
        if (er != hipSuccess) return fail(set_error(YETI_ERR_DEVICE, "matrix-free: upload failed"));
        pd->mp.cpmap = pd->cpmap.p;
        pd->mp.Btab = pd->Btab.p;
        pd->mp.dBtab = pd->dBtab.p;
        const int64_t nel = (int64_t)ne;
        const int nq = ng * ng * ng;
        const int bn0 = 2 * np1 + p0, bn1 = np1 + p0;
        const int64_t nb0 = (int64_t)pd->nbrick[0][0] * pd->nbrick[0][1] * pd->nbrick[0][2];
        const int64_t nb1 = (int64_t)pd->nbrick[1][0] * pd->nbrick[1][1] * pd->nbrick[1][2];
        const int64_t part_n = std::max(part_n64, std::max(nb0 * bn0 * bn0 * bn0 * 3, nb1 * bn1 * bn1 * bn1 * 9));
        // geometry in FP64 always (it seeds the FP32 copy), in whole bricks of the sweep; brick
        // partials per precision
        (void)nel;
        (void)nq;
        er = pd->geo64.alloc(geo_n64);
        if (er == hipSuccess && (precisions & 2)) er = pd->geo32.alloc(geo_n32);
        // the FP64 partials also serve the (FP64) diagonal-block and strain-energy gradient passes
        if (er == hipSuccess) er = pd->part64.alloc(part_n);
        if (er == hipSuccess && (precisions & 2)) er = pd->part32.alloc(part_n32);
        if (er != hipSuccess) return fail(set_error(YETI_ERR_DEVICE, "matrix-free: allocation failed (out of memory?)"));
        if (!counting()) {
            const bool arena = alloc_ctx().arena != nullptr;
            if (mf_geometry(m, *pd, st) != 0 ||
                (arena ? hipStreamSynchronize(st) : hipDeviceSynchronize()) != hipSuccess)
                return fail(set_error(YETI_ERR_DEVICE, "matrix-free: geometry kernel failed"));
        }
        // the FP64 geometry is kept: the diagonal blocks and the FP64 apply read it
        woff += (int64_t)nn * ne;
        ieoff += (int64_t)nn * ne;
        eloff += ne;
    }
    // a CP shared by two patches gets both patches' sums; a CP in no patch gets 0
    std::vector<char> covered((size_t)m->nb_cp, 0);
    for (auto *pd : m->patches)
        for (int32_t v : pd->hcpmap)
            if (v >= 0 && v < m->nb_cp) covered[(size_t)v] = 1;
    m->zero_first = m->patches.size() > 1;
    for (char c : covered) m->zero_first |= c == 0;
    *out = m;
    return 0;
}

int32_t yeti_mf_create(const yeti_system_inputs *in, int32_t precisions, yeti_mf **out) {
    return mf_create(in, precisions, nullptr, out);
}

int32_t yeti_mf_create_in(const yeti_system_inputs *in, int32_t precisions, void *arena, size_t arena_bytes,
                          void *stream, yeti_mf **out, size_t *used) {
    if (!out) return set_error(YETI_ERR_INVALID_ARGUMENT, "null handle pointer");
    *out = nullptr;
    Arena a;
    a.base = (char *)arena;
    a.size = arena ? arena_bytes : 0;
    AllocScope scope(&a, (hipStream_t)stream);
    yeti_mf *m = nullptr;
    int rc = mf_create(in, precisions, (hipStream_t)stream, &m);
    if (used) *used = a.used;
    if (rc) return rc;
    if (!arena) {
        delete m;   // counting only
        return 0;
    }
    *out = m;
    return 0;
}

int32_t yeti_mf_apply_f64(const yeti_mf *m, const double *x, double *y, void *stream) {
    if (!m || !x || !y) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return mf_run<double, 0>(m, x, y, (hipStream_t)stream);
}

int32_t yeti_mf_apply_f32(const yeti_mf *m, const float *x, float *y, void *stream) {
    if (!m || !x || !y) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return mf_run<float, 0>(m, x, y, (hipStream_t)stream);
}

int32_t yeti_mf_diag_blocks_f64(const yeti_mf *m, double *blocks, void *stream) {
    if (!m || !blocks) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return mf_run<double, 1>(m, nullptr, blocks, (hipStream_t)stream);
}

int32_t yeti_mf_strain_energy_grad_f64(const yeti_mf *m, const double *u, double *g, void *stream) {
    if (!m || !u || !g) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    if (!(m->precisions & 1)) return set_error(YETI_ERR_INVALID_ARGUMENT, "matrix-free: FP64 not prepared");
    return mf_run<double, 2>(m, u, g, (hipStream_t)stream);
}

int32_t yeti_mf_update_coords(yeti_mf *m, const double *coords, void *stream) {
    if (!m || !coords) return set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t st = (hipStream_t)stream;
    YETI_HIP(hipMemcpyAsync(m->coords.p, coords, 3 * sizeof(double) * (size_t)m->nb_cp, hipMemcpyDeviceToDevice, st));
    for (auto *pd : m->patches) {
        int rc = mf_geometry(m, *pd, st);
        if (rc) return rc;
    }
    return 0;
}

int32_t yeti_mf_nb_cp(const yeti_mf *m) { return m ? m->nb_cp : -1; }

void yeti_mf_destroy(yeti_mf *m) { delete m; }

}  // extern "C"
