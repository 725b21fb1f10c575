// elem_kernel.hpp -- the FP64 element stiffness kernels (included by assembly.hip, which defines
// GaussRule and bspline_ders before including it).
//
// 1. elem_geo_kernel, one lane per (element, Gauss point): the Jacobian of the NURBS map
//    (shap.f90:94-178) reduced to 12 numbers per point, A[m][c] = sqrt(w |det J|) / W
//    dxi_m/dx_c and dW/dxi_m / W, written to HBM (12 doubles per Gauss point).
// 2. elem_stiffness_kernel, persistent and warp-specialised: one workgroup of 12 waves per CU
//    walks its share of the elements; each element is cut into steps (pass, Gauss-point chunk).
//    In step s
//      * 4 producer waves build step s+1: for every node a and Gauss point q of the chunk,
//        G_c[a][q] = sqrt(w|det J|) dR_a/dx_c = w_a sum_m (dN_a/dxi_m - N_a dW_m/W) A[m][c]
//        into one half of a double-buffered LDS array;
//      * 8 consumer waves accumulate S^{kl}_{ab} = sum_q G_k[a][q] G_l[b][q] over the chunk on
//        v_mfma_f64_16x16x4f64. The 16x16 tiles of a 16-node tile pair (I, J) are grouped in
//        units that contract independently: D = {S^00, S^11, S^22} -> K_ab[k][k], and
//        P(k,l) = {S^kl, S^lk} -> K_ab[k][l], K_ab[l][k] (k < l; on a diagonal pair only S^kl,
//        whose transpose is S^lk); units are dealt to the waves in pair-major runs (about 10
//        tiles per wave at p = 3), so a wave loads each distinct operand fragment once per two K
//        steps and most of its units share fragments;
//    one __syncthreads() ends each step. After an element's last chunk of a pass the consumers
//    contract their units in registers, K_ab[k][l] = lam S^{kl}_ab + mu S^{lk}_ab +
//    mu d_kl tr S_ab (the isotropic form of stiffmatrix_byCP's BoI^T D BoJ,
//    stiffmatrix.F90:175-216), into an LDS stage laid out as the COO entries, and the waves copy
//    it out with coalesced stores, computing the row and column indices from the position. The
//    stage holds one round of node columns b at a time; the last (largest) round of a pass is
//    copied in slices during the following steps, beside the MFMA work and the chunk builds.
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
    // waves that copy the staged COO values out: consumers and producers. (The producers' loads
    // then queue behind stores of their own, the vector memory counter covering both, but with
    // 12 waves on the copy the blocking round ends sooner: 6.42 -> 6.25 ms per launch at p = 3
    // with indices, values only unchanged.)
    static constexpr bool PROD_COPY = true;
    static constexpr int CPW = PROD_COPY ? CW + PW : CW;
    // The deferred slices of the last round, copied beside the next element's steps, are left to
    // the consumers: in those steps the producers' chunk building is the longer task (measured
    // per wave with the shader clock at p = 3: producers busy, consumers waiting at the step
    // barrier), while the blocking rounds keep all 12 waves.
    static constexpr bool PROD_DEFER_COPY = false;
    static constexpr int DPW = PROD_DEFER_COPY ? CW + PW : CW;
    static constexpr int MAXTILE = NNP >= 128 ? 9 : 12; // accumulator tiles per consumer wave per pass
    // Values-only runs use the instantiation that writes indices, with null index pointers:
    // the values-only instantiation at p = 4 (NNP = 128, ~620 bytes of register spills per lane)
    // wrote wrong values for the second and later elements of a workgroup (a defect not traced
    // to the source); the with-indices instantiation agrees with the oracle in both uses.
    static constexpr bool VALUES_VIA_IDX = NNP >= 128;
    // consumers read one MFMA K step of their fragments per LDS read instead of two (the A and B
    // fragments then take half the registers), where the accumulators leave no room for more
    static constexpr bool SPLIT_K = NNP >= 128;
    static constexpr int MAXU = 8;                      // units per consumer wave per pass
    static constexpr int MAXF = 12;                     // distinct A (or B) fragments per wave
    static constexpr int MAXPASS = 8;
    static constexpr int MAXR = 8;                      // staging rounds per pass
    static constexpr int GPW = QC / PW;                 // Gauss points per producer wave per chunk
    // per-element record written by the geometry pass: 12 geometry values per Gauss point
    // ([v][q]), the 1D basis tables B, dB ([which][d][g][i]), the node weights and the nodes'
    // dof bases 3 (IEN - 1) (int32, from double offset DOFOFF)
    static constexpr int NTAB = 2 * 3 * NG * NP1;
    static constexpr int64_t DOFOFF = (int64_t)ELEM_NGEO * NQ + NTAB + NN;
    static constexpr int64_t REC = DOFOFF + (NN + 1) / 2;
    static constexpr int64_t NDE = 9LL * NN * (NN - 1) / 2 + 6LL * NN;   // COO entries per element
    // LDS stage of one round of COO values: what is left of ~150 KiB after G
    static constexpr int64_t GBYTES = 2LL * 3 * NNP * QCP * 8;
    static constexpr int64_t STGMAX = (150 * 1024 - GBYTES - 4 * 1024) / 8;
    static constexpr int STG = (int)(NDE < STGMAX ? NDE : STGMAX);
    static_assert(QC % PW == 0, "chunk must split over producer waves");
};

__host__ __device__ constexpr int64_t coo_start(int b) { return (9LL * b * b + 3LL * b) / 2; }

// Static schedule of consumer units over waves and passes. Unit code = pair * 4 + kind, kind 0:
// D (tiles S^00, S^11, S^22), 1: P(0,1), 2: P(0,2), 3: P(1,2) (tiles S^kl, S^lk; on a diagonal
// pair I = J only S^kl, since S^lk_ab = S^kl_ba lies in the same tile). A pass takes
// whole tile columns J (all pairs I <= J), so the COO entries of nodes b in [16 J0, 16 J1) are
// complete after it; its units are dealt in pair-major runs, balanced over the 4 SIMDs (waves w
// and w + 4 of the workgroup share one) and then over the two waves of each SIMD. Each pass's
// b-range is emitted in rounds that fit the LDS stage.
__host__ __device__ constexpr int pair_J(int pi) {
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= pi) ++j;
    return j;
}
__host__ __device__ constexpr int pair_I(int pi) { return pi - pair_J(pi) * (pair_J(pi) + 1) / 2; }
__host__ __device__ constexpr int unit_tiles(int kind, int pi) {
    return kind == 0 ? 3 : pair_I(pi) == pair_J(pi) ? 1 : 2;
}
__host__ __device__ constexpr int unit_k(int kind, int t) { return kind == 0 ? t : t == 0 ? (kind == 3 ? 1 : 0) : (kind == 1 ? 1 : 2); }
__host__ __device__ constexpr int unit_l(int kind, int t) { return kind == 0 ? t : t == 1 ? (kind == 3 ? 1 : 0) : (kind == 1 ? 1 : 2); }

template <int NT, int WAVES, int MAXU, int MAXTILE, int MAXF, int MAXPASS, int MAXR>
struct UnitSchedule {
    int npass;
    int nu[MAXPASS][WAVES];                 // units of a wave
    int unit[MAXPASS][WAVES][MAXU];         // pair * 4 + kind
    int toff[MAXPASS][WAVES][MAXU];         // accumulator slot of the unit's first tile
    int nt[MAXPASS][WAVES];                 // tiles of a wave
    int na[MAXPASS][WAVES], nb[MAXPASS][WAVES];
    int afr[MAXPASS][WAVES][MAXF];          // distinct A fragments: I * 3 + k
    int bfr[MAXPASS][WAVES][MAXF];          // distinct B fragments: J * 3 + l
    int aidx[MAXPASS][WAVES][MAXTILE];      // tile -> A fragment slot
    int bidx[MAXPASS][WAVES][MAXTILE];      // tile -> B fragment slot
    int nround[MAXPASS];
    int rb[MAXPASS][MAXR + 1];              // round boundaries in b
    bool ok;
};

template <int NT, int WAVES, int MAXU, int MAXTILE, int MAXF, int MAXPASS, int MAXR>
constexpr UnitSchedule<NT, WAVES, MAXU, MAXTILE, MAXF, MAXPASS, MAXR> make_schedule(int nn, int stg) {
    UnitSchedule<NT, WAVES, MAXU, MAXTILE, MAXF, MAXPASS, MAXR> s{};
    s.ok = true;
    int J = 0, pass = 0;
    while (J < NT && pass < MAXPASS) {
        int J1 = J, tiles = 0;
        while (J1 < NT && tiles + 9 * J1 + 6 <= MAXTILE * WAVES) {   // column J1: J1 pairs of 9 tiles, one of 6
            tiles += 9 * J1 + 6;
            ++J1;
        }
        if (J1 == J) {
            s.ok = false;
            return s;
        }
        const int u0 = 4 * (J * (J + 1) / 2), u1 = 4 * (J1 * (J1 + 1) / 2);
        // rounds over b in [16 J, min(16 J1, nn)), built from the end so that the last round
        // (whose copy-out is deferred into the next step) is the largest that fits
        const int bstart = 16 * J, bend = 16 * J1 < nn ? 16 * J1 : nn;
        int cuts[MAXR + 1] = {}, r = 0, b1 = bend;
        cuts[0] = bend;
        while (b1 > bstart) {
            int b0 = b1 - 1;
            while (b0 > bstart && coo_start(b1) - coo_start(b0 - 1) <= stg) --b0;
            if (coo_start(b1) - coo_start(b0) > stg || r >= MAXR) {
                s.ok = false;
                return s;
            }
            cuts[++r] = b0;
            b1 = b0;
        }
        for (int x = 0; x <= r; ++x) s.rb[pass][x] = cuts[r - x];
        s.nround[pass] = r;
        // Staging cost of unit u in round q (the entries it writes there; the three 1-tile units of
        // a diagonal pair also pass a wave barrier and add to LDS, about 3x per entry).
        auto stage_cost = [&](int u, int q) {
            const int pi = u / 4, kind = u % 4, Ju = pair_J(pi);
            const int lo0 = 16 * Ju > s.rb[pass][q] ? 16 * Ju : s.rb[pass][q];
            const int hi0 = 16 * Ju + 16 < s.rb[pass][q + 1] ? 16 * Ju + 16 : s.rb[pass][q + 1];
            const int nbq = hi0 > lo0 ? hi0 - lo0 : 0;
            return unit_tiles(kind, pi) * nbq * (kind != 0 && pair_I(pi) == Ju ? 3 : 1);
        };
        // Waves w and w + 4 share a SIMD's matrix core, so every SIMD gets at most T tiles (the
        // smallest T that places all). With one staging round the units are cut into 4 pair-major
        // runs, one per SIMD, each split between its two waves at the unit boundary that balances
        // their tiles (neighbouring units share fragments). With several rounds every unit is
        // dealt, largest staging cost first, to the wave that keeps the sum over rounds of the
        // slowest wave's staging smallest (each round waits for its slowest wave), then the one
        // that adds the fewest fragments: pair-major runs left the diagonal pairs' staging (one-
        // tile units that pass a wave barrier and add into LDS) on a few waves, one of them busy
        // 6.7k cycles in a round where others had nothing (64^3 p = 3: 6.16 -> 5.82 ms per launch
        // with indices, 5.04 -> 4.59 ms values only; 32^3 p = 4: 41.3 -> 36.9 ms; same output).
        bool placed = false;
        for (int T = (tiles + 3) / 4; T <= 2 * MAXTILE && !placed; ++T) {
            int r0[4] = {}, r1[4] = {}, sidx = 0, fill = 0;
            bool fits = WAVES == 8;
            r0[0] = u0;
            for (int u = u0; u < u1 && fits; ++u) {
                const int nt = unit_tiles(u % 4, u / 4);
                if (fill + nt > T) {
                    r1[sidx] = u;
                    if (++sidx >= 4) {
                        fits = false;
                        break;
                    }
                    r0[sidx] = u;
                    fill = 0;
                }
                fill += nt;
            }
            if (!fits) continue;
            r1[sidx] = u1;
            for (int x = sidx + 1; x < 4; ++x) r0[x] = r1[x] = u1;
            for (int w2 = 0; w2 < WAVES; ++w2) s.nu[pass][w2] = s.nt[pass][w2] = 0;
            for (int sm = 0; sm < 4 && fits && r <= 1; ++sm) {   // one round: contiguous, tile-balanced halves
                int best = -1, bestmax = 1 << 30;
                for (int m = r0[sm]; m <= r1[sm]; ++m) {
                    int ta = 0, tb = 0;
                    for (int u = r0[sm]; u < m; ++u) ta += unit_tiles(u % 4, u / 4);
                    for (int u = m; u < r1[sm]; ++u) tb += unit_tiles(u % 4, u / 4);
                    if (ta > MAXTILE || tb > MAXTILE || m - r0[sm] > MAXU || r1[sm] - m > MAXU) continue;
                    const int mx = ta > tb ? ta : tb;
                    if (mx < bestmax) {
                        bestmax = mx;
                        best = m;
                    }
                }
                if (best < 0) {
                    fits = false;
                    break;
                }
                const int wv[2] = {sm, sm + 4}, ua[2] = {r0[sm], best}, ub[2] = {best, r1[sm]};
                for (int h = 0; h < 2; ++h) {
                    int cnt = 0, f = 0;
                    for (int u = ua[h]; u < ub[h]; ++u) {
                        s.unit[pass][wv[h]][cnt] = u;
                        s.toff[pass][wv[h]][cnt] = f;
                        ++cnt;
                        f += unit_tiles(u % 4, u / 4);
                    }
                    s.nu[pass][wv[h]] = cnt;
                    s.nt[pass][wv[h]] = f;
                }
            }
            if (fits && r > 1) {   // several rounds: every unit dealt to a wave by staging cost
                int cost[WAVES][MAXR] = {}, cnt[WAVES] = {}, f[WAVES] = {}, simd[4] = {};
                int akey[WAVES][MAXF] = {}, bkey[WAVES][MAXF] = {}, na[WAVES] = {}, nb[WAVES] = {};
                bool taken[4 * MAXU * WAVES] = {};
                const int nun = u1 - u0;
                if (nun > 4 * MAXU * WAVES) fits = false;
                for (int step = 0; step < nun && fits; ++step) {
                    int pick = -1, pc = -1, pt = -1;   // largest staging cost first, then most tiles
                    for (int x = 0; x < nun; ++x) {
                        if (taken[x]) continue;
                        int c = 0;
                        for (int q = 0; q < r; ++q) c = stage_cost(u0 + x, q) > c ? stage_cost(u0 + x, q) : c;
                        const int tx = unit_tiles((u0 + x) % 4, (u0 + x) / 4);
                        if (c > pc || (c == pc && tx > pt)) {
                            pc = c;
                            pt = tx;
                            pick = x;
                        }
                    }
                    const int u = u0 + pick, pi = u / 4, kind = u % 4, nt = unit_tiles(kind, pi);
                    int wbest = -1;
                    long long best = 0;
                    for (int w = 0; w < WAVES; ++w) {
                        if (cnt[w] >= MAXU || f[w] + nt > MAXTILE || simd[w % 4] + nt > T) continue;
                        int newa = 0, newb = 0;   // fragments the unit adds to the wave
                        for (int tt = 0; tt < nt; ++tt) {
                            const int ak = pair_I(pi) * 3 + unit_k(kind, tt), bk = pair_J(pi) * 3 + unit_l(kind, tt);
                            bool ha = false, hb = false;
                            for (int y = 0; y < na[w]; ++y) ha |= akey[w][y] == ak;
                            for (int y = 0; y < nb[w]; ++y) hb |= bkey[w][y] == bk;
                            for (int t2 = 0; t2 < tt; ++t2) {
                                ha |= pair_I(pi) * 3 + unit_k(kind, t2) == ak;
                                hb |= pair_J(pi) * 3 + unit_l(kind, t2) == bk;
                            }
                            newa += ha ? 0 : 1;
                            newb += hb ? 0 : 1;
                        }
                        if (na[w] + newa > MAXF || nb[w] + newb > MAXF) continue;
                        long long crit = 0;   // sum over rounds of the slowest wave's staging
                        for (int q = 0; q < r; ++q) {
                            int mx = 0;
                            for (int w2 = 0; w2 < WAVES; ++w2) {
                                const int c2 = cost[w2][q] + (w2 == w ? stage_cost(u, q) : 0);
                                mx = c2 > mx ? c2 : mx;
                            }
                            crit += mx;
                        }
                        const long long sc = crit * 4096 + (newa + newb) * 64 + f[w];
                        if (wbest < 0 || sc < best) {
                            best = sc;
                            wbest = w;
                        }
                    }
                    if (wbest < 0) {
                        fits = false;
                        break;
                    }
                    taken[pick] = true;
                    for (int q = 0; q < r; ++q) cost[wbest][q] += stage_cost(u, q);
                    for (int tt = 0; tt < nt; ++tt) {
                        const int ak = pair_I(pi) * 3 + unit_k(kind, tt), bk = pair_J(pi) * 3 + unit_l(kind, tt);
                        bool ha = false, hb = false;
                        for (int y = 0; y < na[wbest]; ++y) ha |= akey[wbest][y] == ak;
                        for (int y = 0; y < nb[wbest]; ++y) hb |= bkey[wbest][y] == bk;
                        if (!ha) akey[wbest][na[wbest]++] = ak;
                        if (!hb) bkey[wbest][nb[wbest]++] = bk;
                    }
                    s.unit[pass][wbest][cnt[wbest]] = u;
                    s.toff[pass][wbest][cnt[wbest]] = f[wbest];
                    ++cnt[wbest];
                    f[wbest] += nt;
                    simd[wbest % 4] += nt;
                }
                for (int w = 0; w < WAVES; ++w) {
                    s.nu[pass][w] = cnt[w];
                    s.nt[pass][w] = f[w];
                }
            }
            placed = fits;
        }
        if (!placed) {
            s.ok = false;
            return s;
        }
        // distinct fragments per wave
        for (int w = 0; w < WAVES; ++w) {
            s.na[pass][w] = s.nb[pass][w] = 0;
            for (int x = 0; x < s.nu[pass][w]; ++x) {
                const int u = s.unit[pass][w][x], pi = u / 4, kind = u % 4;
                for (int t = 0; t < unit_tiles(kind, pi); ++t) {
                    const int akey = pair_I(pi) * 3 + unit_k(kind, t), bkey = pair_J(pi) * 3 + unit_l(kind, t);
                    int ai = -1, bi = -1;
                    for (int y = 0; y < s.na[pass][w]; ++y)
                        if (s.afr[pass][w][y] == akey) ai = y;
                    for (int y = 0; y < s.nb[pass][w]; ++y)
                        if (s.bfr[pass][w][y] == bkey) bi = y;
                    if (ai < 0) {
                        if (s.na[pass][w] == MAXF) {
                            s.ok = false;
                            return s;
                        }
                        ai = s.na[pass][w]++;
                        s.afr[pass][w][ai] = akey;
                    }
                    if (bi < 0) {
                        if (s.nb[pass][w] == MAXF) {
                            s.ok = false;
                            return s;
                        }
                        bi = s.nb[pass][w]++;
                        s.bfr[pass][w][bi] = bkey;
                    }
                    s.aidx[pass][w][s.toff[pass][w][x] + t] = ai;
                    s.bidx[pass][w][s.toff[pass][w][x] + t] = bi;
                }
            }
        }
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
    static constexpr auto S =
        make_schedule<C::NT, C::CW, C::MAXU, C::MAXTILE, C::MAXF, C::MAXPASS, C::MAXR>(C::NN, C::STG);
    static constexpr int NPASS = S.npass;
    static_assert(S.ok, "unit schedule does not fit the wave / staging capacities");
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
// Geometry pass: per active element gi a record of REC doubles at geo + gi REC: for Gauss point
// q, [v NQ + q] with v = 3 m + c: A[m][c] = sqrt(w |det J|) / W dxi_m/dx_c; v = 9 + m:
// dW/dxi_m / W; then the element's 1D basis tables (values and derivatives at the abscissae,
// reversed local index) and its node weights, which the element kernel's producers copy with
// independent loads. J is the
// derivative with respect to the knot parameters (shap.f90:133-153); w |det J| includes the
// parent-to-parameter factors (shap.f90:156-178) and the Gauss weights (UELMAT.f90:118-119).
// One wave per element, 4 per workgroup. The 16 sums a Gauss point needs (W, dW/dxi_m, and
// sum w x_c R and its xi derivatives, before the division by W) are tensor-product contractions
// of the four nodal quantities (w, w x, w y, w z), evaluated by sum factorisation through LDS:
// along xi_0 (values and derivatives at the NG abscissae), then xi_1, then xi_2, one lane per
// Gauss point in the last stage (NP1^3 NG + ... multiply-adds per quantity instead of NP1^3 NG^3).
template <int NQ>
constexpr int geo_epw() { return 1; }   // elements per wave

template <int P, int NG>
__global__ void __launch_bounds__(256) elem_geo_kernel(PatchArgs pa, double *__restrict__ geo) {
    using C = ElemCfg<P, NG>;
    constexpr int NP1 = C::NP1, NN = C::NN, NQ = C::NQ;
    constexpr int EPB = 4;                              // elements per workgroup, one per wave
    constexpr int NT1 = 4 * 2 * NG * NP1 * NP1;         // stage 0: [quantity][B | dB][g0][k][j]
    constexpr int NT2 = 4 * 3 * NG * NG * NP1;          // stage 1: [quantity][v][g0][g1][k]
    __shared__ double cp[EPB][4][NN];                   // (w, w x, w y, w z) per node
    __shared__ double tab[EPB][3][2][NG][NP1];          // [d][B | dB][g][i], i <-> basis P - i
    __shared__ double hh[EPB][3];                       // half span per direction
    __shared__ double T1[EPB][NT1];
    __shared__ double T2[EPB][NT2];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t gi = (int64_t)blockIdx.x * EPB + wv;
    const bool act = gi < pa.n_active;   // wave-uniform; inactive waves still meet the barriers
    const int e = act ? pa.elems[gi] : 0;
    double *rec = geo + gi * C::REC;
    if (act) {
        for (int a = lane; a < NN; a += 64) {
            const int32_t cp1 = pa.ien[(int64_t)NN * e + a];
            const double *xp = pa.coords + 3 * (int64_t)(cp1 - 1);
            const double w = pa.weight[(int64_t)NN * e + a];
            reinterpret_cast<int32_t *>(rec + C::DOFOFF)[a] = 3 * (cp1 - 1);
            cp[wv][0][a] = w;
            cp[wv][1][a] = w * xp[0];
            cp[wv][2][a] = w * xp[1];
            cp[wv][3][a] = w * xp[2];
            rec[ELEM_NGEO * NQ + C::NTAB + a] = w;
        }
        if (lane < 3 * NG) {   // 1D bases of direction d at abscissa g
            const int d = lane / NG, g = lane - d * NG;
            const double *U = d == 0 ? pa.U0 : d == 1 ? pa.U1 : pa.U2;
            const int span1 = pa.nijk[3 * (int64_t)e + d];
            const double u0 = U[span1 - 1], u1 = U[span1];
            const double u = ((u1 - u0) * pa.g.x[g] + (u1 + u0)) * 0.5;
            double N[NP1], dN[NP1];
            bspline_ders<P>(span1 - 1, u, U, N, dN);
            double *tb = rec + ELEM_NGEO * NQ + (d * NG + g) * NP1;
#pragma unroll
            for (int i = 0; i < NP1; ++i) {   // local index i <-> basis function P - i
                tab[wv][d][0][g][i] = N[P - i];
                tab[wv][d][1][g][i] = dN[P - i];
                tb[i] = N[P - i];
                tb[3 * NG * NP1 + i] = dN[P - i];
            }
            if (g == 0) hh[wv][d] = 0.5 * (u1 - u0);
        }
    }
    __syncthreads();
    if (act) {   // stage 0: contract i (xi_0)
        for (int o = lane; o < NT1; o += 64) {
            const int j = o % NP1, k = (o / NP1) % NP1, g0 = (o / (NP1 * NP1)) % NG, v = (o / (NP1 * NP1 * NG)) % 2,
                      qn = o / (NP1 * NP1 * NG * 2);
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < NP1; ++i) s += tab[wv][0][v][g0][i] * cp[wv][qn][i + NP1 * (j + NP1 * k)];
            T1[wv][o] = s;
        }
    }
    __syncthreads();
    if (act) {   // stage 1: contract j (xi_1); v = 0: B0 B1, 1: dB0 B1, 2: B0 dB1
        for (int o = lane; o < NT2; o += 64) {
            const int k = o % NP1, g1 = (o / NP1) % NG, g0 = (o / (NP1 * NG)) % NG, v = (o / (NP1 * NG * NG)) % 3,
                      qn = o / (NP1 * NG * NG * 3);
            const double *t1 = &T1[wv][(((qn * 2 + (v == 1 ? 1 : 0)) * NG + g0) * NP1 + k) * NP1];
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NP1; ++j) s += t1[j] * tab[wv][1][v == 2 ? 1 : 0][g1][j];
            T2[wv][o] = s;
        }
    }
    __syncthreads();
    if (!act) return;
    // stage 2: contract k (xi_2), one lane per Gauss point, and the point's Jacobian work
    for (int q = lane; q < NQ; q += 64) {
        const int g0 = q % NG, g1 = (q / NG) % NG, g2 = q / (NG * NG);
        double s[16];   // 0: W, 1-3: dW/dxi_m, 4 + c: sum w x_c R, 7 + 3 c + m: its xi_m derivative
#pragma unroll
        for (int qn = 0; qn < 4; ++qn) {
            double val = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
            const double *t2 = &T2[wv][((qn * 3) * NG * NG + g0 * NG + g1) * NP1];
#pragma unroll
            for (int k = 0; k < NP1; ++k) {
                const double bz = tab[wv][2][0][g2][k], dz = tab[wv][2][1][g2][k];
                const double t0 = t2[k];
                val += t0 * bz;
                d2 += t0 * dz;
                d0 += t2[NG * NG * NP1 + k] * bz;
                d1 += t2[2 * NG * NG * NP1 + k] * bz;
            }
            if (qn == 0) {
                s[0] = val;
                s[1] = d0;
                s[2] = d1;
                s[3] = d2;
            } else {
                s[3 + qn] = val;
                s[7 + 3 * (qn - 1) + 0] = d0;
                s[7 + 3 * (qn - 1) + 1] = d1;
                s[7 + 3 * (qn - 1) + 2] = d2;
            }
        }
        const double Wi = 1.0 / s[0];
        const double dw[3] = {s[1] * Wi, s[2] * Wi, s[3] * Wi};
        double J[3][3];
#pragma unroll
        for (int r2 = 0; r2 < 3; ++r2)
#pragma unroll
            for (int c = 0; c < 3; ++c) J[r2][c] = (s[7 + 3 * r2 + c] - s[4 + r2] * dw[c]) * Wi;
        const double m00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double m01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double m02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * m00 + J[0][1] * m01 + J[0][2] * m02;
        const double om = pa.g.w[g0] * pa.g.w[g1] * pa.g.w[g2] * fabs(det * hh[wv][0] * hh[wv][1] * hh[wv][2]);
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
        double *out = rec + q;
#pragma unroll
        for (int v = 0; v < 9; ++v) out[v * NQ] = A[v];
#pragma unroll
        for (int m = 0; m < 3; ++m) out[(9 + m) * NQ] = dw[m];
    }
}

// ---------------------------------------------------------------------------------------
template <int P, int NG>
struct ElemSmem {
    using C = ElemCfg<P, NG>;
    // G_c[a][q] = sqrt(w |det J|) dR_a/dx_c at the chunk's Gauss points, double buffered
    __attribute__((aligned(16))) double G[2][3][C::NNP][C::QCP];
    double B[C::PW][3][NG][C::NP1];     // per producer wave: B(reversed local index) at the abscissae
    double dB[C::PW][3][NG][C::NP1];
    double stage[C::STG];               // COO values of one round, element-local order
    // dof bases 3 (IEN - 1) and COO offset of the elements being emitted, by element parity
    // (written by the producers one step before the element's first step)
    int32_t dof[2][C::NN];
    int64_t ebase[2];
};

// Producer: per-element set-up (1D bases in the wave's own LDS table, node weights in registers),
// fetched from the element's record into registers one step before the element starts, so the
// loads overlap the step in between.
template <int P, int NG>
struct ProducerElem {
    using C = ElemCfg<P, NG>;
    double wb[(C::NNP + 63) / 64];   // weight of the G-row nodes a = lane + 64 m
};

template <int P, int NG>
struct ProducerFetch {
    using C = ElemCfg<P, NG>;
    double tab[(C::NTAB + 63) / 64];   // table entry lane + 64 m
    double w[(C::NNP + 63) / 64];
    int32_t dof[(C::NNP + 63) / 64];
    int64_t base;                      // COO offset of the element
};

template <int P, int NG>
__device__ __forceinline__ void producer_fetch_element(const double *__restrict__ rec, const int64_t *offs_gi,
                                                       ProducerFetch<P, NG> &f) {
    using C = ElemCfg<P, NG>;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < (C::NTAB + 63) / 64; ++m) {
        const int t = lane + 64 * m;
        f.tab[m] = t < C::NTAB ? rec[ELEM_NGEO * C::NQ + t] : 0.0;
    }
#pragma unroll
    for (int m = 0; m < (C::NNP + 63) / 64; ++m) {
        const int a = lane + 64 * m;
        f.w[m] = a < C::NN ? rec[ELEM_NGEO * C::NQ + C::NTAB + a] : 0.0;
        f.dof[m] = a < C::NN ? reinterpret_cast<const int32_t *>(rec + C::DOFOFF)[a] : 0;
    }
    f.base = *offs_gi;
}

template <int P, int NG>
__device__ __forceinline__ void producer_store_element(ElemSmem<P, NG> &sm, int pw, int slot,
                                                       const ProducerFetch<P, NG> &f, ProducerElem<P, NG> &pe) {
    using C = ElemCfg<P, NG>;
    constexpr int HALF = 3 * NG * C::NP1;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < (C::NTAB + 63) / 64; ++m) {
        const int t = lane + 64 * m;
        if (t < C::NTAB) {
            const int u = t < HALF ? t : t - HALF;
            const int d = u / (NG * C::NP1), g = (u / C::NP1) % NG, i = u % C::NP1;
            if (t < HALF) sm.B[pw][d][g][i] = f.tab[m];
            else sm.dB[pw][d][g][i] = f.tab[m];
        }
    }
#pragma unroll
    for (int m = 0; m < (C::NNP + 63) / 64; ++m) pe.wb[m] = f.w[m];
    if (pw == 0) {   // the element's dof bases and COO offset, for the copy-out of its values
#pragma unroll
        for (int m = 0; m < (C::NNP + 63) / 64; ++m)
            if (lane + 64 * m < C::NN) sm.dof[slot][lane + 64 * m] = f.dof[m];
        if (lane == 0) sm.ebase[slot] = f.base;
    }
    // the table is read back by this wave only: order the LDS writes before the reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Producer: the geometry of this wave's Gauss points of a chunk, loaded one step before it is
// used. g[t] holds Gauss point q0 + qw + t: lane 16 r + v of every 16-lane row r holds its value
// v, so one DPP row broadcast per value (no lane reads into scalar registers) hands it to all
// 64 node lanes.
template <int P, int NG>
struct ProducerGeo {
    double g[ElemCfg<P, NG>::GPW];
};

template <int P, int NG>
__device__ __forceinline__ ProducerGeo<P, NG> producer_prefetch(const double *__restrict__ geo_e, int pw, int q0) {
    using C = ElemCfg<P, NG>;
    const int v = threadIdx.x & 15;
    ProducerGeo<P, NG> r;
#pragma unroll
    for (int t = 0; t < C::GPW; ++t) {
        const int q = q0 + pw * C::GPW + t;
        r.g[t] = (v < ELEM_NGEO && q < C::NQ) ? geo_e[v * C::NQ + q] : 0.0;
    }
    return r;
}

// value of lane N of each 16-lane row, in every lane of that row (DPP row_newbcast)
template <int N>
__device__ __forceinline__ double row_bcast(double v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + N, 0xF, 0xF, true);
}

// Producer: Gauss points [q0 + qw, q0 + qw + GPW) of the chunk into buffer `buf`; lane = node.
template <int P, int NG>
__device__ __forceinline__ void producer_chunk(ElemSmem<P, NG> &sm, const ProducerGeo<P, NG> &gv, int pw, int buf,
                                               int q0, const ProducerElem<P, NG> &pe) {
    using C = ElemCfg<P, NG>;
    constexpr int NP1 = C::NP1, NN = C::NN, NNP = C::NNP, NQ = C::NQ;
    static_assert(ELEM_NGEO <= 16, "geometry of a Gauss point must fit a 16-lane row");
    const int lane = threadIdx.x & 63;
    const int qw = pw * C::GPW;
#pragma unroll
    for (int t = 0; t < C::GPW; ++t) {
        const int qt = q0 + qw + t;   // wave-uniform
        double A[9], dw[3];
        static_for<9>([&](auto v) { A[v] = row_bcast<decltype(v)::value>(gv.g[t]); });
        static_for<3>([&](auto m) { dw[m] = row_bcast<9 + decltype(m)::value>(gv.g[t]); });
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
    f64x4 S[C::MAXTILE];
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int u = 0; u < C::MAXTILE; ++u) S[u] = f64x4{0.0, 0.0, 0.0, 0.0};
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
        if constexpr (C::SPLIT_K) {
            // one MFMA K step per LDS read: half the fragment registers
#pragma unroll 1
            for (int t = 0; t < C::QC / 4; ++t) {
                const int qo = 8 * (t >> 1) + 2 * kg + (t & 1);
                double Af[na], Bf[nb];
                static_for<na>([&](auto ic) {
                    constexpr int x = decltype(ic)::value;
                    constexpr int key = Sched<P, NG>::S.afr[PASS][W][x];
                    Af[x] = sm.G[buf][key % 3][16 * (key / 3) + nl][qo];
                });
                static_for<nb>([&](auto ic) {
                    constexpr int x = decltype(ic)::value;
                    constexpr int key = Sched<P, NG>::S.bfr[PASS][W][x];
                    Bf[x] = sm.G[buf][key % 3][16 * (key / 3) + nl][qo];
                });
                static_for<nt>([&](auto ic) {
                    constexpr int u = decltype(ic)::value;
                    constexpr int ai = Sched<P, NG>::S.aidx[PASS][W][u], bi = Sched<P, NG>::S.bidx[PASS][W][u];
                    acc.S[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(Af[ai], Bf[bi], acc.S[u], 0, 0, 0);
                });
            }
            return;
        }
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
                constexpr int ai = Sched<P, NG>::S.aidx[PASS][W][u], bi = Sched<P, NG>::S.bidx[PASS][W][u];
                acc.S[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(Af[ai].x, Bf[bi].x, acc.S[u], 0, 0, 0);
            });
            static_for<nt>([&](auto ic) {
                constexpr int u = decltype(ic)::value;
                constexpr int ai = Sched<P, NG>::S.aidx[PASS][W][u], bi = Sched<P, NG>::S.bidx[PASS][W][u];
                acc.S[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(Af[ai].y, Bf[bi].y, acc.S[u], 0, 0, 0);
            });
        }
    }
}

// Epilogue, first half: wave W contracts its units whose column node b lies in round R of pass
// PASS and writes the K values at their COO positions (relative to the round) in the stage.
// Accumulator lane layout (v_mfma_f64_16x16x4f64): register r of lane l holds row (l >> 4) + 4 r
// and column l & 15 of the tile; rows are nodes a of tile I, columns nodes b of tile J. COO
// position of block (a, b), a < b: start(b) + 9 a + kk + 3 ll with start(b) = 9 b (b-1)/2 + 6 b;
// diagonal block: start(b) + 9 b + {0..5} holding (0,0) (0,1) (1,1) (0,2) (1,2) (2,2), diagonal
// halved (build_system_elemStorage.f90:260-293).
template <int P, int NG, int PASS, int W, int R>
__device__ __forceinline__ void stage_round(ElemSmem<P, NG> &sm, double lam, double mu, const ConsumerAcc<P, NG> &acc) {
    constexpr auto &S = Sched<P, NG>::S;
    constexpr int nu = S.nu[PASS][W];
    constexpr int b0 = S.rb[PASS][R], b1 = S.rb[PASS][R + 1];
    constexpr int c0 = (int)coo_start(b0);
    (void)S;
    const int lane = threadIdx.x & 63;
    const int nl = lane & 15, kg = lane >> 4;
    static_for<nu>([&](auto ic) {
        constexpr int x = decltype(ic)::value;
        constexpr int code = Sched<P, NG>::S.unit[PASS][W][x];
        constexpr int off = Sched<P, NG>::S.toff[PASS][W][x];
        constexpr int pi = code / 4, kind = code % 4;
        constexpr int I = pair_I(pi), J = pair_J(pi);
        if constexpr (16 * J + 15 >= b0 && 16 * J < b1) {
            const int b = 16 * J + nl;
            if (b >= b0 && b < b1) {
                double *st = sm.stage + ((int)coo_start(b) - c0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 16 * I + kg + 4 * r;
                    if (I == J && a > b) continue;
                    if constexpr (kind == 0) {
                        const double s0 = acc.S[off][r], s1 = acc.S[off + 1][r], s2 = acc.S[off + 2][r];
                        const double tr = mu * (s0 + s1 + s2);
                        const double k00 = (lam + mu) * s0 + tr, k11 = (lam + mu) * s1 + tr, k22 = (lam + mu) * s2 + tr;
                        if (a < b) {
                            st[9 * a + 0] = k00;
                            st[9 * a + 4] = k11;
                            st[9 * a + 8] = k22;
                        } else {
                            st[9 * b + 0] = 0.5 * k00;
                            st[9 * b + 2] = 0.5 * k11;
                            st[9 * b + 5] = 0.5 * k22;
                        }
                    } else if constexpr (I != J) {
                        constexpr int k = kind == 3 ? 1 : 0, l = kind == 1 ? 1 : 2;
                        const double skl = acc.S[off][r], slk = acc.S[off + 1][r];
                        st[9 * a + k + 3 * l] = lam * skl + mu * slk;   // K_ab[k][l]
                        st[9 * a + l + 3 * k] = lam * slk + mu * skl;   // K_ab[l][k]
                    } else {
                        // diagonal pair, tile s = S^kl: K_ab[k][l] = lam s_ab + mu s_ba and
                        // K_ab[l][k] = lam s_ba + mu s_ab; here the s_ab (a <= b) parts, the s_ba
                        // parts are added below by the lanes holding the lower triangle
                        constexpr int k = kind == 3 ? 1 : 0, l = kind == 1 ? 1 : 2;
                        constexpr int mdiag = kind == 1 ? 1 : kind == 2 ? 3 : 4;
                        const double sab = acc.S[off][r];
                        if (a < b) {
                            st[9 * a + k + 3 * l] = lam * sab;
                            st[9 * a + l + 3 * k] = mu * sab;
                        } else {
                            st[9 * b + mdiag] = (lam + mu) * sab;   // diagonal block keeps the upper entry (k, l)
                        }
                    }
                }
            }
            if constexpr (kind != 0 && I == J) {
                // lower triangle of the diagonal-pair tile: lane (a, b), a > b, holds s_ab = S^kl_ab,
                // the s_ba part of block (b, a), whose row in the COO belongs to node a
                constexpr int k = kind == 3 ? 1 : 0, l = kind == 1 ? 1 : 2;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int bb = 16 * J + nl;   // the lane's column node (the block's row node here)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 16 * I + kg + 4 * r;
                    if (a > bb && a >= b0 && a < b1) {
                        double *sta = sm.stage + ((int)coo_start(a) - c0);
                        const double s_ab = acc.S[off][r];
                        sta[9 * bb + k + 3 * l] += mu * s_ab;
                        sta[9 * bb + l + 3 * k] += lam * s_ab;
                    }
                }
            }
        }
    });
}

// Epilogue, second half (waves 0 .. NW-1): the COO entries of round R of pass PASS, row by row
// (row b = the contiguous entries [start(b), start(b+1))). The rows are cut into NPART slices;
// in slice `part` wave w takes rows b0 + w + NW part, then every (NW NPART)-th row, and copies
// them with consecutive lanes on consecutive entries. Entry j of a row: j < 9 b is block
// a = j / 9, (kk, ll) = (j mod 3, j mod 9 / 3); the last 6 are the diagonal block.
// row = 3 (IEN_a - 1) + kk, col = 3 (IEN_b - 1) + ll.
template <int P, int NG, bool WRITE_IDX, int PASS, int R, int NW, int NPART = 1>
__device__ __forceinline__ void copy_round(const ElemSmem<P, NG> &sm, int slot, double *__restrict__ Kdata,
                                           int32_t *__restrict__ Krow, int32_t *__restrict__ Kcol, int part = 0) {
    constexpr auto &S = Sched<P, NG>::S;
    constexpr int b0 = S.rb[PASS][R], b1 = S.rb[PASS][R + 1];
    constexpr int c0 = (int)coo_start(b0);
    (void)S;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t base = sm.ebase[slot];
    const int32_t *dof = sm.dof[slot];
    for (int b = b0 + wave + NW * part; b < b1; b += NW * NPART) {
        const int len = 9 * b + 6;
        const int64_t gb = base + coo_start(b);
        const double *st_b = sm.stage + ((int)coo_start(b) - c0);
        const int32_t dofb = WRITE_IDX ? dof[b] : 0;
        for (int j0 = 0; j0 < len; j0 += 128) {
            double v[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) v[h] = st_b[min(j0 + 64 * h + lane, len - 1)];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int j = j0 + 64 * h + lane;
                if (j < len) {
                    Kdata[gb + j] = v[h];
                    if (WRITE_IDX && Krow) {   // null: a values-only run of this instantiation
                        int a, kk, ll;
                        if (j < 9 * b) {
                            a = j / 9;
                            const int m = j - 9 * a;
                            ll = m / 3;
                            kk = m - 3 * ll;
                        } else {
                            const int d = j - 9 * b;
                            a = b;
                            kk = (0x210100 >> (4 * d)) & 15;
                            ll = (0x222110 >> (4 * d)) & 15;
                        }
                        Krow[gb + j] = dof[a] + kk;
                        Kcol[gb + j] = dofb + ll;
                    }
                }
            }
        }
    }
}

// Whether the last round's copy-out is deferred: it is then split into NCH - 1 slices of rows that
// all waves copy during the following NCH - 1 steps (which have no epilogue), beside the MFMA work
// and the chunk builds, instead of stalling the epilogue step. Needs at least two chunks per pass.
template <int P, int NG>
constexpr bool elem_defer_last() { return ElemCfg<P, NG>::NCH >= 2; }

// Rounds of pass PASS for consumer wave W: stage, barrier, copy (all waves), barrier before the
// next round; the last round is only staged when its copy is deferred.
template <int P, int NG, bool WRITE_IDX, int PASS, int W, int R = 0>
__device__ __forceinline__ void consumer_epilogue(ElemSmem<P, NG> &sm, double lam, double mu, int slot,
                                                  const ConsumerAcc<P, NG> &acc, double *__restrict__ Kdata,
                                                  int32_t *__restrict__ Krow, int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    constexpr int NR = Sched<P, NG>::S.nround[PASS];
    stage_round<P, NG, PASS, W, R>(sm, lam, mu, acc);
    if constexpr (R + 1 == NR && elem_defer_last<P, NG>()) return;
    __syncthreads();
    copy_round<P, NG, WRITE_IDX, PASS, R, C::CPW>(sm, slot, Kdata, Krow, Kcol);
    if constexpr (R + 1 < NR) {
        __syncthreads();
        consumer_epilogue<P, NG, WRITE_IDX, PASS, W, R + 1>(sm, lam, mu, slot, acc, Kdata, Krow, Kcol);
    }
}

// The producers' side of the same rounds (same barriers; they help with the copies when
// PROD_COPY).
template <int P, int NG, bool WRITE_IDX, int PASS, int R = 0>
__device__ __forceinline__ void producer_epilogue_pass(ElemSmem<P, NG> &sm, int slot, double *__restrict__ Kdata,
                                                       int32_t *__restrict__ Krow, int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    constexpr int NR = Sched<P, NG>::S.nround[PASS];
    if constexpr (R + 1 == NR && elem_defer_last<P, NG>()) return;
    __syncthreads();
    if constexpr (C::PROD_COPY) copy_round<P, NG, WRITE_IDX, PASS, R, C::CPW>(sm, slot, Kdata, Krow, Kcol);
    if constexpr (R + 1 < NR) {
        __syncthreads();
        producer_epilogue_pass<P, NG, WRITE_IDX, PASS, R + 1>(sm, slot, Kdata, Krow, Kcol);
    }
}

template <int P, int NG, bool WRITE_IDX, int PASS = 0>
__device__ __forceinline__ void producer_epilogue(ElemSmem<P, NG> &sm, int pass, int slot,
                                                  double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                                  int32_t *__restrict__ Kcol) {
    if (pass == PASS) {
        producer_epilogue_pass<P, NG, WRITE_IDX, PASS>(sm, slot, Kdata, Krow, Kcol);
    } else if constexpr (PASS + 1 < Sched<P, NG>::NPASS) {
        producer_epilogue<P, NG, WRITE_IDX, PASS + 1>(sm, pass, slot, Kdata, Krow, Kcol);
    }
}

// Slice `part` of the deferred copy-out of the last round of pass `pass` (by the copying waves;
// part = -1 copies the whole round).
template <int P, int NG, bool WRITE_IDX, int PASS = 0>
__device__ __forceinline__ void deferred_copy(const ElemSmem<P, NG> &sm, int pass, int part, int slot,
                                              double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                              int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    constexpr int NW = C::DPW;
    if (pass == PASS) {
        constexpr int R = Sched<P, NG>::S.nround[PASS] - 1;
        if (part < 0)
            copy_round<P, NG, WRITE_IDX, PASS, R, NW>(sm, slot, Kdata, Krow, Kcol);
        else
            copy_round<P, NG, WRITE_IDX, PASS, R, NW, C::NCH - 1>(sm, slot, Kdata, Krow, Kcol, part);
    } else if constexpr (PASS + 1 < Sched<P, NG>::NPASS) {
        deferred_copy<P, NG, WRITE_IDX, PASS + 1>(sm, pass, part, slot, Kdata, Krow, Kcol);
    }
}

// One consumer step for wave W: the pass is dispatched to compile-time code.
template <int P, int NG, bool WRITE_IDX, int W, int PASS = 0>
__device__ __forceinline__ void consumer_step(ElemSmem<P, NG> &sm, const PatchArgs &pa, int buf, int pass, int ch,
                                              int slot, ConsumerAcc<P, NG> &acc,
                                              double *__restrict__ Kdata, int32_t *__restrict__ Krow,
                                              int32_t *__restrict__ Kcol) {
    using C = ElemCfg<P, NG>;
    if (pass == PASS) {
        if (ch == 0) acc.zero();
        mfma_chunk<P, NG, PASS, W>(sm, buf, acc);
        if (ch == C::NCH - 1) {
            consumer_epilogue<P, NG, WRITE_IDX, PASS, W>(sm, pa.lam, pa.mu, slot, acc, Kdata, Krow, Kcol);
        }
    } else if constexpr (PASS + 1 < Sched<P, NG>::NPASS) {
        consumer_step<P, NG, WRITE_IDX, W, PASS + 1>(sm, pa, buf, pass, ch, slot, acc, Kdata, Krow, Kcol);
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
    int pend_pass = -1;     // pass whose last round waits in the stage for its copy-out
    int pend_slot = 0;
    // No global loads in this loop (the element's COO offset and dof bases come from the
    // producers through LDS): a wait for a load would also wait for this wave's stores.
    for (int s = 0; s < nsteps; ++s) {
        const int el = s / SPE, rem = s % SPE;
        const int pass = rem / C::NCH, ch = rem % C::NCH;
        const int slot = el & 1;
        consumer_step<P, NG, WRITE_IDX, W>(sm, pa, s & 1, pass, ch, slot, acc, Kdata, Krow, Kcol);
        if constexpr (elem_defer_last<P, NG>()) {
            if (ch == C::NCH - 1) {
                pend_pass = pass;
                pend_slot = slot;
            } else if (pend_pass >= 0) {
                deferred_copy<P, NG, WRITE_IDX>(sm, pend_pass, ch, pend_slot, Kdata, Krow, Kcol);
            }
        }
        __syncthreads();
    }
    // an unfinished deferred round (the stage is stable after the last barrier)
    if constexpr (elem_defer_last<P, NG>())
        if (pend_pass >= 0) deferred_copy<P, NG, WRITE_IDX>(sm, pend_pass, -1, pend_slot, Kdata, Krow, Kcol);
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
        // Step s: issue the geometry load for step s + 2, build the chunk of step s + 1 from the
        // load of step s - 1, then (at an element start) fetch the tables of step s + 2's
        // element, so each load has a whole step to arrive.
        ProducerElem<P, NG> pe;
        ProducerFetch<P, NG> fe;
        const double *geo_b = geo + (int64_t)blockIdx.x * C::REC;
        const int64_t rec_stride = (int64_t)gridDim.x * C::REC;
        auto step_geo = [&](int st) {   // geometry of step st for this wave's lanes
            return producer_prefetch<P, NG>(geo_b + (int64_t)(st / SPE) * rec_stride, pw, ((st % SPE) % C::NCH) * C::QC);
        };
        auto fetch = [&](int el) {
            producer_fetch_element<P, NG>(geo_b + (int64_t)el * rec_stride, pa.offs + blockIdx.x + (int64_t)el * gridDim.x,
                                          fe);
        };
        ProducerGeo<P, NG> gv{};   // geometry of the chunk built next
        if (nsteps > 0) {
            fetch(0);
            gv = step_geo(0);
            producer_store_element<P, NG>(sm, pw, 0, fe, pe);
            producer_chunk<P, NG>(sm, gv, pw, 0, 0, pe);
            if (nsteps > 1) {
                gv = step_geo(1);
                if (SPE == 1) fetch(1);
            }
        }
        __syncthreads();
        int pend_pass = -1;     // pass whose last round waits in the stage for its copy-out
        int pend_slot = 0;
        for (int s = 0; s < nsteps; ++s) {
            const int s1 = s + 1, s2 = s + 2;
            const int slot = (s / SPE) & 1;
            const ProducerGeo<P, NG> gcur = gv;
            // the element of step s1 starts: its slot's last reader, the element two back, was
            // copied out by the end of the previous step
            if (s1 < nsteps && s1 % SPE == 0) producer_store_element<P, NG>(sm, pw, (s1 / SPE) & 1, fe, pe);
            if (s2 < nsteps) gv = step_geo(s2);   // geometry for step s2, before this step's chunk
            if (s1 < nsteps) producer_chunk<P, NG>(sm, gcur, pw, s1 & 1, ((s1 % SPE) % C::NCH) * C::QC, pe);
            if (s2 < nsteps && s2 % SPE == 0) fetch(s2 / SPE);   // after the chunk's wait on gcur
            const int rem = s % SPE, ch = rem % C::NCH;
            if (ch == C::NCH - 1) {
                producer_epilogue<P, NG, WRITE_IDX>(sm, rem / C::NCH, slot, Kdata, Krow, Kcol);
                if constexpr (elem_defer_last<P, NG>() && C::PROD_DEFER_COPY) {
                    pend_pass = rem / C::NCH;
                    pend_slot = slot;
                }
            } else if (elem_defer_last<P, NG>() && C::PROD_DEFER_COPY && pend_pass >= 0) {
                deferred_copy<P, NG, WRITE_IDX>(sm, pend_pass, ch, pend_slot, Kdata, Krow, Kcol);
            }
            __syncthreads();
        }
        // an unfinished deferred round (the stage is stable after the last barrier)
        if constexpr (elem_defer_last<P, NG>() && C::PROD_DEFER_COPY)
            if (pend_pass >= 0) deferred_copy<P, NG, WRITE_IDX>(sm, pend_pass, -1, pend_slot, Kdata, Krow, Kcol);
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
