// elem_kernel.hpp -- the FP64 element stiffness kernel (included by assembly.hip, which defines
// GaussRule and bspline_ders before including it).
//
// Persistent, warp-specialised pipeline: one workgroup of 12 waves per CU walks its share of the
// patch's elements. Each element is cut into steps (pass, Gauss-point chunk). In step s
//   * 4 producer waves build step s+1: per Gauss point the reductions W, dW, sum x N,
//     sum x (x) dN over the element's nodes -> J^{-1}, w|det J| (shap.f90:94-178), then
//     G_c[a][q] = sqrt(w|det J|) dR_a/dx_c into one half of a double-buffered LDS array;
//   * 8 consumer waves accumulate S^{kl}_{ab} = sum_q w|J| dR_a/dx_k dR_b/dx_l over step s's
//     chunk on v_mfma_f64_16x16x4f64 and, after an element's last chunk, contract its COO
//     values into an LDS stage, from which all 12 waves write them (and the row / column
//     indices, computed from the position) with coalesced stores;
// one __syncthreads() ends each step. Producers never need each other: each producer wave owns
// a quarter of every chunk's Gauss points and reduces them inside the wave.
//
// The S tiles are split into units that are independent in the epilogue:
//   D unit of node-tile pair (I, J): S^{00}, S^{11}, S^{22}       -> K_ab[k][k]
//   P unit (k, l) of (I, J):          S^{kl}, S^{lk}  (k < l)      -> K_ab[k][l], K_ab[l][k]
// with K_ab[k][l] = lam S^{kl}_ab + mu S^{lk}_ab + mu d_kl tr S_ab, the isotropic form of
// stiffmatrix_byCP's BoI^T D BoJ (stiffmatrix.F90:175-216). A compile-time longest-first
// schedule deals the units to the consumer waves so every SIMD gets the same MFMA work.
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
    static constexpr int DCAP = 2;                      // D units per consumer wave per pass
    static constexpr int PCAP = 4;                      // P units per consumer wave per pass
    static constexpr int MAXPASS = 8;
    static constexpr int MAXR = 8;                      // staging rounds per pass
    static constexpr int GPW = QC / PW;                 // Gauss points per producer wave per chunk
    static constexpr int LPG = 64 / GPW;                // lanes per Gauss point in the reductions
    static constexpr int NPL = (NN + LPG - 1) / LPG;    // nodes per lane in the reductions
    static constexpr int64_t NDE = 9LL * NN * (NN - 1) / 2 + 6LL * NN;   // COO entries per element
    // LDS staging of the COO values (coalesced stores): what is left of ~150 KiB after G
    static constexpr int64_t GBYTES = 2LL * 3 * NNP * QCP * 8;
    static constexpr int64_t STGMAX = (150 * 1024 - GBYTES - 6 * 1024) / 8;
    static constexpr int STG = (int)(NDE < STGMAX ? NDE : STGMAX);
    static_assert(QC % PW == 0 && 64 % GPW == 0, "chunk must split over producer waves");
};

__host__ __device__ constexpr int64_t coo_start(int b) { return (9LL * b * b + 3LL * b) / 2; }

// Static schedule of D and P units over consumer waves and passes. A pass takes whole tile
// columns J (all pairs I <= J), so the COO entries of b in [16 J0, 16 J1) are complete after it;
// inside a pass the units go longest-first to the least-loaded wave. Each pass's b-range is
// emitted in rounds whose entries fit the LDS staging buffer.
template <int NT, int WAVES, int DCAP, int PCAP, int MAXPASS, int MAXR>
struct UnitSchedule {
    int npass;
    int nd[MAXPASS][WAVES];
    int np[MAXPASS][WAVES];
    int d[MAXPASS][WAVES][DCAP];      // tile-pair index
    int p[MAXPASS][WAVES][PCAP];      // tile-pair index * 3 + (0: (0,1), 1: (0,2), 2: (1,2))
    int nround[MAXPASS];
    int rb[MAXPASS][MAXR + 1];        // round boundaries in b
    bool ok;
};

template <int NT, int WAVES, int DCAP, int PCAP, int MAXPASS, int MAXR>
constexpr UnitSchedule<NT, WAVES, DCAP, PCAP, MAXPASS, MAXR> make_schedule(int nn, int stg) {
    UnitSchedule<NT, WAVES, DCAP, PCAP, MAXPASS, MAXR> s{};
    s.ok = true;
    int J = 0, pass = 0;
    while (J < NT && pass < MAXPASS) {
        // largest run of columns starting at J whose units fit the per-pass capacity
        int J1 = J, nd = 0, np = 0;
        while (J1 < NT) {
            int add = J1 + 1;   // pairs in column J1
            if (nd + add > DCAP * WAVES || np + 3 * add > PCAP * WAVES) break;
            nd += add;
            np += 3 * add;
            ++J1;
        }
        if (J1 == J) {
            s.ok = false;
            return s;
        }
        int load[WAVES] = {};
        for (int w = 0; w < WAVES; ++w) s.nd[pass][w] = s.np[pass][w] = 0;
        for (int pi = J * (J + 1) / 2; pi < J1 * (J1 + 1) / 2; ++pi) {
            int best = -1;
            for (int w = 0; w < WAVES; ++w)
                if (s.nd[pass][w] < DCAP && (best < 0 || load[w] < load[best])) best = w;
            if (best < 0) {
                s.ok = false;
                return s;
            }
            s.d[pass][best][s.nd[pass][best]++] = pi;
            load[best] += 3;
        }
        for (int pi = J * (J + 1) / 2; pi < J1 * (J1 + 1) / 2; ++pi)
            for (int which = 0; which < 3; ++which) {
                int best = -1;
                for (int w = 0; w < WAVES; ++w)
                    if (s.np[pass][w] < PCAP && (best < 0 || load[w] < load[best])) best = w;
                if (best < 0) {
                    s.ok = false;
                    return s;
                }
                s.p[pass][best][s.np[pass][best]++] = 3 * pi + which;
                load[best] += 2;
            }
        // rounds over b in [16 J, min(16 J1, nn))
        int b0 = 16 * J, bend = 16 * J1 < nn ? 16 * J1 : nn, r = 0;
        s.rb[pass][0] = b0;
        while (b0 < bend) {
            int b1 = b0 + 1;
            while (b1 < bend && coo_start(b1 + 1) - coo_start(b0) <= stg) ++b1;
            if (coo_start(b1) - coo_start(b0) > stg || r >= MAXR) {
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
    static constexpr auto S = make_schedule<C::NT, C::CW, C::DCAP, C::PCAP, C::MAXPASS, C::MAXR>(C::NN, C::STG);
    static constexpr int NPASS = S.npass;
    static_assert(S.ok, "unit schedule does not fit the wave / staging capacities");
};

struct PatchArgs {
    const double *coords;     // (3, nb_cp)
    const int32_t *ien;       // (nnode, nelem_patch), 1-based
    const double *weight;     // (nnode, nelem_patch)
    const int32_t *nijk;      // (3, nelem_patch), 1-based spans
    const double *U0, *U1, *U2;  // knot vectors
    const int32_t *elems;     // active elements (patch-local 0-based)
    const int64_t *offs;      // COO offset per active element
    int32_t n_active;
    double lam, mu;
    GaussRule g;
};

template <int P, int NG>
struct ElemSmem {
    using C = ElemCfg<P, NG>;
    // sqrt(w |det J|) dR_a/dx_c at the chunk's Gauss points, double buffered
    __attribute__((aligned(16))) double G[2][3][C::NNP][C::QCP];
    double B[C::PW][3][NG][C::NP1];     // per producer wave: B(reversed local index) at the abscissae
    double dB[C::PW][3][NG][C::NP1];
    double half[C::PW][3];
    double stage[C::STG];             // COO values of one round, element-local order
    int32_t dof[C::NN];               // 3 (IEN - 1) of the element being emitted
};

__device__ constexpr int p_k(int which) { return which == 2 ? 1 : 0; }
__device__ constexpr int p_l(int which) { return which == 0 ? 1 : 2; }
__device__ constexpr int pair_J(int pi) {
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= pi) ++j;
    return j;
}
__device__ constexpr int pair_I(int pi) { return pi - pair_J(pi) * (pair_J(pi) + 1) / 2; }

__device__ __forceinline__ double bcast(double v, int src) {
    // wave-uniform source lane -> scalar broadcast
    int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------------------
// Producer: per-element set-up (1D bases in the wave's own LDS table, node data in registers).
template <int P, int NG>
struct ProducerElem {
    using C = ElemCfg<P, NG>;
    double x[C::NPL][3];   // coordinates of the reduction nodes a = sub + LPG * m
    double w[C::NPL];      // their weights
    double wb[(C::NNP + 63) / 64];   // weight of the G-row nodes a = lane + 64 m
};

template <int P, int NG>
__device__ __forceinline__ void producer_load_element(ElemSmem<P, NG> &sm, const PatchArgs &pa, int e, int pw,
                                                      ProducerElem<P, NG> &pe) {
    using C = ElemCfg<P, NG>;
    constexpr int NP1 = C::NP1, NN = C::NN;
    const int lane = threadIdx.x & 63;
    const int sub = lane % C::LPG;
    const int32_t *ien = pa.ien + (int64_t)NN * e;
    const double *wt = pa.weight + (int64_t)NN * e;
#pragma unroll
    for (int m = 0; m < C::NPL; ++m) {
        const int a = sub + C::LPG * m;
        if (a < NN) {
            const double *xp = pa.coords + 3 * (int64_t)(ien[a] - 1);
            pe.x[m][0] = xp[0];
            pe.x[m][1] = xp[1];
            pe.x[m][2] = xp[2];
            pe.w[m] = wt[a];
        } else {
            pe.x[m][0] = pe.x[m][1] = pe.x[m][2] = 0.0;
            pe.w[m] = 0.0;
        }
    }
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
        if (g == 0) sm.half[pw][d] = 0.5 * (u1 - u0);
    }
    // the table is read back by this wave only: order the LDS writes before the reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Producer: Gauss points [qw, qw + GPW) of the chunk starting at q0 into buffer `buf`.
template <int P, int NG>
__device__ __forceinline__ void producer_chunk(ElemSmem<P, NG> &sm, const PatchArgs &pa, int pw, int buf,
                                               int q0, const ProducerElem<P, NG> &pe) {
    using C = ElemCfg<P, NG>;
    constexpr int NP1 = C::NP1, NN = C::NN, NNP = C::NNP;
    const int lane = threadIdx.x & 63;
    const int gq = lane / C::LPG;             // Gauss point of this lane within the wave's share
    const int sub = lane % C::LPG;
    const int qw = pw * C::GPW;               // first chunk-local Gauss point of this wave
    const int q = q0 + qw + gq;
    const bool live = q < C::NQ;
    const int g0 = q % NG, g1 = (q / NG) % NG, g2 = (q / (NG * NG)) % NG;
    // ---- A. reductions over the element's nodes -----------------------------------------
    double s[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) s[m] = 0.0;
    if (live) {
#pragma unroll
        for (int m = 0; m < C::NPL; ++m) {
            const int a = sub + C::LPG * m;
            if (a < NN) {
                const int i = a % NP1, j = (a / NP1) % NP1, k = a / (NP1 * NP1);
                const double b0 = sm.B[pw][0][g0][i], b1 = sm.B[pw][1][g1][j], b2 = sm.B[pw][2][g2][k];
                const double w = pe.w[m];
                const double b12 = b1 * b2 * w;
                const double Na = b0 * b12;
                const double d0 = sm.dB[pw][0][g0][i] * b12;
                const double d1 = b0 * sm.dB[pw][1][g1][j] * b2 * w;
                const double d2 = b0 * b1 * sm.dB[pw][2][g2][k] * w;
                const double x0 = pe.x[m][0], x1 = pe.x[m][1], x2 = pe.x[m][2];
                s[0] += Na;
                s[1] += d0; s[2] += d1; s[3] += d2;
                s[4] += x0 * Na; s[5] += x1 * Na; s[6] += x2 * Na;
                s[7] += x0 * d0; s[8] += x0 * d1; s[9] += x0 * d2;
                s[10] += x1 * d0; s[11] += x1 * d1; s[12] += x1 * d2;
                s[13] += x2 * d0; s[14] += x2 * d1; s[15] += x2 * d2;
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 16; ++m)
#pragma unroll
        for (int o = 1; o < C::LPG; o <<= 1) s[m] += __shfl_xor(s[m], o, C::LPG);
    // every lane of the group now holds the sums; form 1/W, dW/W, J^{-1}, w|det J|
    double gpv[14];      // Wi, dw0..2, Ji[c][r] (9), sqrt(w |det J|)
    {
        const double Wi = live ? 1.0 / s[0] : 0.0;
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
        const double id = live ? 1.0 / det : 0.0;
        gpv[0] = Wi;
        gpv[1] = dw[0];
        gpv[2] = dw[1];
        gpv[3] = dw[2];
        gpv[4] = m00 * id;
        gpv[5] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
        gpv[6] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
        gpv[7] = m01 * id;
        gpv[8] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
        gpv[9] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
        gpv[10] = m02 * id;
        gpv[11] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
        gpv[12] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
        gpv[13] = 0.0;
        if (live) {
            const double detJ = det * sm.half[pw][0] * sm.half[pw][1] * sm.half[pw][2];
            gpv[13] = sqrt(pa.g.w[g0] * pa.g.w[g1] * pa.g.w[g2] * fabs(detJ));
        }
    }
    // ---- B. G_c[a][q]: lane = node, loop over the wave's Gauss points ------------------
#pragma unroll
    for (int m = 0; m < (NNP + 63) / 64; ++m) {
        const int a = lane + 64 * m;
        const int i = a % NP1, j = (a / NP1) % NP1, k = a / (NP1 * NP1);
        const double wa = pe.wb[m];
#pragma unroll
        for (int t = 0; t < C::GPW; ++t) {
            const int src = t * C::LPG;                  // a lane of Gauss point t's group
            const int qt = q0 + qw + t;
            double v[14];
#pragma unroll
            for (int c = 0; c < 14; ++c) v[c] = bcast(gpv[c], src);
            double g0v = 0.0, g1v = 0.0, g2v = 0.0;
            if (a < NN && qt < C::NQ) {
                const int h0 = qt % NG, h1 = (qt / NG) % NG, h2 = qt / (NG * NG);
                const double b0 = sm.B[pw][0][h0][i], b1 = sm.B[pw][1][h1][j], b2 = sm.B[pw][2][h2][k];
                const double w = wa * v[0];
                const double b12 = b1 * b2 * w;
                const double R = b0 * b12;
                const double r0 = sm.dB[pw][0][h0][i] * b12 - R * v[1];
                const double r1 = b0 * sm.dB[pw][1][h1][j] * b2 * w - R * v[2];
                const double r2 = b0 * b1 * sm.dB[pw][2][h2][k] * w - R * v[3];
                g0v = (r0 * v[4] + r1 * v[7] + r2 * v[10]) * v[13];
                g1v = (r0 * v[5] + r1 * v[8] + r2 * v[11]) * v[13];
                g2v = (r0 * v[6] + r1 * v[9] + r2 * v[12]) * v[13];
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
// Consumer: MFMA over one chunk for consumer wave W of pass PASS (unit indices compile-time).
// K step 2t uses Gauss points 8t + 2g, step 2t+1 uses 8t + 2g + 1 (g = lane K group): the order
// of the K sum is free as long as A and B use the same permutation.
template <int P, int NG, int PASS, int W>
__device__ __forceinline__ void mfma_chunk(const ElemSmem<P, NG> &sm, int buf,
                                           f64x4 (&accD)[ElemCfg<P, NG>::DCAP][3],
                                           f64x4 (&accP)[ElemCfg<P, NG>::PCAP][2]) {
    using C = ElemCfg<P, NG>;
    constexpr int nd = Sched<P, NG>::S.nd[PASS][W], np = Sched<P, NG>::S.np[PASS][W];
    const int lane = threadIdx.x & 63;
    const int nl = lane & 15, kg = lane >> 4;
#pragma unroll 1
    for (int t = 0; t < C::QC / 8; ++t) {
        const int qo = 8 * t + 2 * kg;
        static_for<nd>([&](auto ic) {
            constexpr int u = decltype(ic)::value;
            constexpr int pi = Sched<P, NG>::S.d[PASS][W][u];
            constexpr int I = pair_I(pi), J = pair_J(pi);
            const int ra = 16 * I + nl, rb = 16 * J + nl;
            f64x2 A[3], B[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                A[c] = *reinterpret_cast<const f64x2 *>(&sm.G[buf][c][ra][qo]);
                B[c] = *reinterpret_cast<const f64x2 *>(&sm.G[buf][c][rb][qo]);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                accD[u][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[c].x, B[c].x, accD[u][c], 0, 0, 0);
                accD[u][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[c].y, B[c].y, accD[u][c], 0, 0, 0);
            }
        });
        static_for<np>([&](auto ic) {
            constexpr int u = decltype(ic)::value;
            constexpr int code = Sched<P, NG>::S.p[PASS][W][u];
            constexpr int pi = code / 3, which = code % 3;
            constexpr int k = p_k(which), l = p_l(which);
            constexpr int I = pair_I(pi), J = pair_J(pi);
            const int ra = 16 * I + nl, rb = 16 * J + nl;
            const f64x2 Ak = *reinterpret_cast<const f64x2 *>(&sm.G[buf][k][ra][qo]);
            const f64x2 Al = *reinterpret_cast<const f64x2 *>(&sm.G[buf][l][ra][qo]);
            const f64x2 Bk = *reinterpret_cast<const f64x2 *>(&sm.G[buf][k][rb][qo]);
            const f64x2 Bl = *reinterpret_cast<const f64x2 *>(&sm.G[buf][l][rb][qo]);
            accP[u][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ak.x, Bl.x, accP[u][0], 0, 0, 0);
            accP[u][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ak.y, Bl.y, accP[u][0], 0, 0, 0);
            accP[u][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(Al.x, Bk.x, accP[u][1], 0, 0, 0);
            accP[u][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(Al.y, Bk.y, accP[u][1], 0, 0, 0);
        });
    }
}

// Consumer epilogue, first half: isotropic contraction of the accumulators whose column node b
// lies in round R of pass PASS, written to the LDS stage in element-local COO order.
// Accumulator lane layout (v_mfma_f64_16x16x4f64): register r of lane l holds row (l >> 4) + 4 r
// and column l & 15 of the tile; rows are nodes a of tile I, columns nodes b of tile J. COO
// position of block (a, b), a < b: start(b) + 9 a + kk + 3 ll with start(b) = 9 b (b-1)/2 + 6 b
// (build_system_elemStorage.f90:260-293); diagonal block: start(b) + 9 b + {0..5} holding
// (0,0) (0,1) (1,1) (0,2) (1,2) (2,2), diagonal halved.
template <int P, int NG, int PASS, int W, int R>
__device__ __forceinline__ void stage_round(ElemSmem<P, NG> &sm, double lam, double mu,
                                            const f64x4 (&accD)[ElemCfg<P, NG>::DCAP][3],
                                            const f64x4 (&accP)[ElemCfg<P, NG>::PCAP][2]) {
    constexpr auto &S = Sched<P, NG>::S;
    constexpr int nd = S.nd[PASS][W], np = S.np[PASS][W];
    constexpr int b0 = S.rb[PASS][R], b1 = S.rb[PASS][R + 1];
    (void)S;
    constexpr int s0 = (int)coo_start(b0);
    const int lane = threadIdx.x & 63;
    const int nl = lane & 15, kg = lane >> 4;
    static_for<nd>([&](auto ic) {
        constexpr int u = decltype(ic)::value;
        constexpr int pi = Sched<P, NG>::S.d[PASS][W][u];
        constexpr int I = pair_I(pi), J = pair_J(pi);
        if constexpr (16 * J + 15 >= b0 && 16 * J < b1) {
            const int b = 16 * J + nl;
            if (b >= b0 && b < b1) {
                double *st = sm.stage + ((int)coo_start(b) - s0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 16 * I + kg + 4 * r;
                    if (I == J && a > b) continue;
                    const double c0 = accD[u][0][r], c1 = accD[u][1][r], c2 = accD[u][2][r];
                    const double tr = mu * (c0 + c1 + c2);
                    const double k00 = (lam + mu) * c0 + tr, k11 = (lam + mu) * c1 + tr, k22 = (lam + mu) * c2 + tr;
                    if (a < b) {
                        st[9 * a + 0] = k00;
                        st[9 * a + 4] = k11;
                        st[9 * a + 8] = k22;
                    } else {
                        st[9 * b + 0] = 0.5 * k00;
                        st[9 * b + 2] = 0.5 * k11;
                        st[9 * b + 5] = 0.5 * k22;
                    }
                }
            }
        }
    });
    static_for<np>([&](auto ic) {
        constexpr int u = decltype(ic)::value;
        constexpr int code = Sched<P, NG>::S.p[PASS][W][u];
        constexpr int pi = code / 3, which = code % 3;
        constexpr int k = p_k(which), l = p_l(which);
        constexpr int I = pair_I(pi), J = pair_J(pi);
        constexpr int mdiag = (k == 0 && l == 1) ? 1 : (k == 0 && l == 2) ? 3 : 4;
        if constexpr (16 * J + 15 >= b0 && 16 * J < b1) {
            const int b = 16 * J + nl;
            if (b >= b0 && b < b1) {
                double *st = sm.stage + ((int)coo_start(b) - s0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 16 * I + kg + 4 * r;
                    if (I == J && a > b) continue;
                    const double ckl = accP[u][0][r], clk = accP[u][1][r];
                    const double kkl = lam * ckl + mu * clk;     // K_ab[k][l]
                    const double klk = lam * clk + mu * ckl;     // K_ab[l][k]
                    if (a < b) {
                        st[9 * a + k + 3 * l] = kkl;
                        st[9 * a + l + 3 * k] = klk;
                    } else {
                        st[9 * b + mdiag] = kkl;   // diagonal block keeps the upper entry (k, l), k < l
                    }
                }
            }
        }
    });
}

// Epilogue, second half (all 768 threads): round R of pass PASS leaves the LDS stage with
// consecutive lanes on consecutive COO entries. The (row, col) of an entry follow from its
// element-local position L alone: b = largest node with start(b) <= L, then the 9-entry block
// index a and (kk, ll) inside it, or the 6-entry diagonal pattern.
template <int P, int NG, bool WRITE_IDX, int PASS, int R>
__device__ __forceinline__ void copy_round(const ElemSmem<P, NG> &sm, int64_t base, double *__restrict__ Kdata,
                                           int32_t *__restrict__ Krow, int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    constexpr auto &S = Sched<P, NG>::S;
    constexpr int b0 = S.rb[PASS][R], b1 = S.rb[PASS][R + 1];
    constexpr int s0 = (int)coo_start(b0);
    constexpr int n = (int)coo_start(b1) - s0;
    const int64_t g0 = base + s0;
    for (int j = threadIdx.x; j < n; j += C::NTH) {
        Kdata[g0 + j] = sm.stage[j];
        if (WRITE_IDX) {
            const int L = s0 + j;
            int b = (int)((__builtin_sqrtf((float)(72 * L + 9)) - 3.0f) * (1.0f / 18.0f));
            b = b < b0 ? b0 : b >= b1 ? b1 - 1 : b;
            if ((int)coo_start(b + 1) <= L) ++b;
            if ((int)coo_start(b) > L) --b;
            const int off = L - (int)coo_start(b);
            int a, kk, ll;
            if (off < 9 * b) {
                a = off / 9;
                const int m = off - 9 * a;
                ll = m / 3;
                kk = m - 3 * ll;
            } else {
                const int d = off - 9 * b;
                a = b;
                kk = (0x210100 >> (4 * d)) & 15;
                ll = (0x222110 >> (4 * d)) & 15;
            }
            Krow[g0 + j] = sm.dof[a] + kk;
            Kcol[g0 + j] = sm.dof[b] + ll;
        }
    }
}

// Rounds of pass PASS for consumer wave W: stage, barrier, copy, (barrier before the next round).
template <int P, int NG, bool WRITE_IDX, int PASS, int W, int R = 0>
__device__ __forceinline__ void consumer_epilogue(ElemSmem<P, NG> &sm, double lam, double mu, int64_t base,
                                                  const f64x4 (&accD)[ElemCfg<P, NG>::DCAP][3],
                                                  const f64x4 (&accP)[ElemCfg<P, NG>::PCAP][2],
                                                  double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                                  int32_t *__restrict__ Kcol) {
    constexpr int NR = Sched<P, NG>::S.nround[PASS];
    stage_round<P, NG, PASS, W, R>(sm, lam, mu, accD, accP);
    __syncthreads();
    copy_round<P, NG, WRITE_IDX, PASS, R>(sm, base, Kdata, Krow, Kcol);
    if constexpr (R + 1 < NR) {
        __syncthreads();
        consumer_epilogue<P, NG, WRITE_IDX, PASS, W, R + 1>(sm, lam, mu, base, accD, accP, Kdata, Krow, Kcol);
    }
}

// The producers' side of the same rounds (same barriers, they help with the copies).
template <int P, int NG, bool WRITE_IDX, int PASS, int R = 0>
__device__ __forceinline__ void producer_epilogue_pass(const ElemSmem<P, NG> &sm, int64_t base,
                                                       double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                                       int32_t *__restrict__ Kcol) {
    constexpr int NR = Sched<P, NG>::S.nround[PASS];
    __syncthreads();
    copy_round<P, NG, WRITE_IDX, PASS, R>(sm, base, Kdata, Krow, Kcol);
    if constexpr (R + 1 < NR) {
        __syncthreads();
        producer_epilogue_pass<P, NG, WRITE_IDX, PASS, R + 1>(sm, base, Kdata, Krow, Kcol);
    }
}

template <int P, int NG, bool WRITE_IDX, int PASS = 0>
__device__ __forceinline__ void producer_epilogue(const ElemSmem<P, NG> &sm, int pass, int64_t base,
                                                  double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                                  int32_t *__restrict__ Kcol) {
    if (pass == PASS) {
        producer_epilogue_pass<P, NG, WRITE_IDX, PASS>(sm, base, Kdata, Krow, Kcol);
    } else if constexpr (PASS + 1 < Sched<P, NG>::NPASS) {
        producer_epilogue<P, NG, WRITE_IDX, PASS + 1>(sm, pass, base, Kdata, Krow, Kcol);
    }
}

template <int P, int NG>
struct ConsumerAcc {
    using C = ElemCfg<P, NG>;
    f64x4 D[C::DCAP][3];
    f64x4 Pp[C::PCAP][2];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int u = 0; u < C::DCAP; ++u)
#pragma unroll
            for (int c = 0; c < 3; ++c) D[u][c] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int u = 0; u < C::PCAP; ++u)
#pragma unroll
            for (int c = 0; c < 2; ++c) Pp[u][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    }
};

// One consumer step for wave W: the pass is dispatched to compile-time code.
template <int P, int NG, bool WRITE_IDX, int W, int PASS = 0>
__device__ __forceinline__ void consumer_step(ElemSmem<P, NG> &sm, const PatchArgs &pa, int buf, int pass, int ch,
                                              int e, int64_t base, ConsumerAcc<P, NG> &acc,
                                              double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                              int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    if (pass == PASS) {
        if (ch == 0) acc.zero();
        mfma_chunk<P, NG, PASS, W>(sm, buf, acc.D, acc.Pp);
        if (ch == C::NCH - 1) {
            if (WRITE_IDX && W < 2) {
                const int t = threadIdx.x;
                if (t < C::NN) sm.dof[t] = 3 * (pa.ien[(int64_t)C::NN * e + t] - 1);
            }
            consumer_epilogue<P, NG, WRITE_IDX, PASS, W>(sm, pa.lam, pa.mu, base, acc.D, acc.Pp, Kdata, Krow, Kcol);
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
elem_stiffness_kernel(PatchArgs pa, double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                      int32_t *__restrict__ Kcol) {
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
    // the first step, one after each step), so each role keeps only its own registers live.
    if (producer) {
        ProducerElem<P, NG> pe;
        int pe_elem = -1;
        if (nsteps > 0) {
            producer_load_element<P, NG>(sm, pa, pa.elems[blockIdx.x], pw, pe);
            pe_elem = 0;
            producer_chunk<P, NG>(sm, pa, pw, 0, 0, pe);
        }
        __syncthreads();
        for (int s = 0; s < nsteps; ++s) {
            const int s1 = s + 1;
            if (s1 < nsteps) {
                const int el = s1 / SPE, ch = (s1 % SPE) % C::NCH;
                if (el != pe_elem) {
                    producer_load_element<P, NG>(sm, pa, pa.elems[blockIdx.x + (int64_t)el * gridDim.x], pw, pe);
                    pe_elem = el;
                }
                producer_chunk<P, NG>(sm, pa, pw, s1 & 1, ch * C::QC, pe);
            }
            const int rem = s % SPE;
            if (rem % C::NCH == C::NCH - 1) {
                const int64_t gi = blockIdx.x + (int64_t)(s / SPE) * gridDim.x;
                producer_epilogue<P, NG, WRITE_IDX>(sm, rem / C::NCH, pa.offs[gi], Kdata, Krow, Kcol);
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
