// mf_sweep.hpp -- the matrix-free apply y = K x as sum-factorised sweeps over bricks of elements
// (included by mf.hip, after MfPatch).
//
// Same operator as the element-wise contraction of mf.hip (mf_iga_get_ku_3d,
// iga_wq_mf/fsrc/mf_iga_solver.f90:292-555, generalised to 3-component NURBS elasticity), but the
// 1D contractions run over a whole brick of BRX x BRY x BZ elements instead of one element at a
// time: control points shared by neighbouring elements of the brick are contracted once, and the
// sums over the elements that share a control point happen inside the backward contractions
// instead of through a coloured scatter.
//
// A brick is swept through its node layers in z. For node layer n3 the brick's control points are
// read once (w x, 3 components), contracted along x (stage A: nodes -> Gauss positions
// a = e1 NG + g0, value and x-derivative) and along y (stage B: -> Gauss-point columns
// (a, b = e2 NG + g1): value, x- and y-derivative of the 3 components, 9 numbers per column) into
// a ring of the last P+1 node layers in LDS. Once the P+1 node layers of element layer e3 are in
// the ring, the Gauss points of the layer are visited: the z contraction (stage C), the
// quadrature-point work (the same 13 stored geometry numbers per point as the element-wise
// kernel) and the transposed z contraction into P+1 back-contraction registers (one per node
// layer e3 .. e3+P, shifted down after every element layer). Node layer e3 is then complete: its
// back-contractions, summed over g2, are contracted back along y (B^T) and x (A^T), scaled by
// the weights and written to the brick's partial sums; bricks overlap by P node planes, and a
// gather sums the bricks covering each control point in a fixed order. No floating-point
// atomics: the apply is bit-reproducible.
//
// Threads (Sweep below): at P = 1 one thread per Gauss point, the NG points g2 of a column in
// adjacent lanes (the sum over g2 by lane shuffles); at P >= 2 one thread per column, looping
// over g2 (the sum over g2 in registers). The stored geometry is laid out per (brick, element
// layer, value) as [column][g2], so each value is one contiguous read per wave either way.
#pragma once

namespace yeti {

// Brick shape and thread mapping per degree P and value width W (4: FP32, 8: FP64 bytes).
// P = 1: the NG Gauss points g2 of a column in adjacent lanes (CPW = floor(64 / NG) columns per
// wave, back-contractions summed over g2 with lane shuffles). P >= 2: one thread per column,
// looping over g2; bricks sized so that the columns fill whole waves (P = 2: 7 x 3 elements,
// 189 columns in 3 waves; P = 3: 2 x 2, one wave; P = 4: 2 x 2, 100 columns in 2 waves), so
// several bricks share a CU. (With lanes per g2 a P = 4 brick needs 9 waves, one workgroup per
// CU, and every barrier of a node layer stalls the CU.)
template <int P, int W>
struct Sweep {
    static constexpr int NP1 = P + 1, NG = P + 1;
    static constexpr bool PER_COLUMN = P >= 2;
    static constexpr int BRX = P == 1 ? 8 : P == 2 ? 7 : 2;    // brick edge in x (elements)
    static constexpr int BRY = P == 1 ? 8 : P == 2 ? 3 : 2;    // brick edge in y (elements)
    static constexpr int BZ = 8;                               // brick depth in z (elements)
    static constexpr int BNX = BRX + P, BNY = BRY + P;         // control points per brick row in x, y
    static constexpr int NZ = BZ + P;                          // node layers per brick
    static constexpr int NAX = BRX * NG, NAY = BRY * NG;       // Gauss positions per brick row in x, y
    static constexpr int NCOL = NAX * NAY;                     // Gauss-point columns per brick
    static constexpr int NTQ = NG * NCOL;                      // Gauss points per element layer
    static constexpr int CPW = PER_COLUMN ? 64 : 64 / NG;      // columns per wave
    static constexpr int NWAVE = (NCOL + CPW - 1) / CPW;
    static constexpr int NT = 64 * NWAVE;                      // threads per brick
    static constexpr int NGEO = 13;                            // stored geometry values per point
    static constexpr int PART = NZ * BNY * BNX * 3;            // partial sums per brick
    // waves per SIMD the register allocation is held to (the bricks' LDS allows more)
    static constexpr int WPE = P == 1 ? (W == 4 ? 4 : 3) : P == 2 ? (W == 4 ? 4 : 2) : P == 3 ? (W == 4 ? 3 : 2)
                                                                                               : (W == 4 ? 3 : 2);
    static_assert(BZ >= BRX && BZ >= BRY, "one table size for all directions");
    static_assert(NT <= 1024, "brick too large for one workgroup");
};

// Brick grid of a patch for the sweep (elements per direction over the brick edges).
template <int P, int W>
__host__ __device__ inline void sweep_grid(const int32_t N[3], int32_t nb[3]) {
    using S = Sweep<P, W>;
    nb[0] = (N[0] + S::BRX - 1) / S::BRX;
    nb[1] = (N[1] + S::BRY - 1) / S::BRY;
    nb[2] = (N[2] + S::BZ - 1) / S::BZ;
}

// Offset of geometry value c of Gauss point (g0, g1, g2) of element (e1, e2, e3):
// [brick][element layer in the brick][c][column][g2], column = a + NAX b.
template <int P, int W>
__host__ __device__ inline int64_t sweep_geo_off(const int32_t nb[3], int e1, int e2, int e3, int c, int g0, int g1,
                                                 int g2) {
    using S = Sweep<P, W>;
    const int b1 = e1 / S::BRX, b2 = e2 / S::BRY, b3 = e3 / S::BZ;
    const int64_t brick = b1 + (int64_t)nb[0] * (b2 + (int64_t)nb[1] * b3);
    const int col = (e1 - b1 * S::BRX) * S::NG + g0 + S::NAX * ((e2 - b2 * S::BRY) * S::NG + g1);
    return (((brick * S::BZ + (e3 - b3 * S::BZ)) * S::NGEO + c) * S::NCOL + col) * S::NG + g2;
}

template <int P, int W>
inline int64_t sweep_geo_size(const int32_t nb[3]) {
    using S = Sweep<P, W>;
    return (int64_t)nb[0] * nb[1] * nb[2] * S::BZ * S::NGEO * S::NTQ;
}

template <typename T, int P>
struct SweepSmem {
    using S = Sweep<P, sizeof(T)>;
    T tab[3][2][S::BZ][S::NG][S::NP1];   // [direction][B | dB][element of the brick][g][i]
    T X[3][S::BNY][S::BNX];              // node layer: [component][n2][n1] = w x
    // weights of the node layers, [slot][n2][n1]: a ring of P + 2 slots, so that the loader's
    // write of layer n3 + 1 cannot meet another thread's read of layer n3 - P in A^T
    T Wn[S::NP1 + 1][S::BNY][S::BNX];
    T Y[3][2][S::BNY][S::NAX];           // stage A: [component][value | d/dxi0][n2][a]
    // ring of node layers after stage B, [slot][3 f + var][column]: shared only at P = 1, where a
    // column's g2 points are separate threads; with one thread per column it lives in registers
    T Z[S::PER_COLUMN ? 1 : S::NP1][9][S::PER_COLUMN ? 1 : S::NCOL];
    T Ws[9][S::NCOL];                    // back-contractions of the completed node layer, summed over g2
    T Pb[3][2][S::BNY][S::NAX];          // stage B^T: [component][P' | R0'][n2][a]
};

// One brick: partial[brick] = its elements' contribution to K x at its (NZ x BNY x BNX) nodes.
// Every thread of the workgroup (Sweep<P, sizeof(T)>::NT threads) calls it.
template <typename T, int P>
__device__ __forceinline__ void mf_sweep_body(SweepSmem<T, P> &sm, const MfPatch &mp, const T *__restrict__ geo,
                                              const T *__restrict__ x, const T *__restrict__ wcp,
                                              T *__restrict__ partial, int64_t brick) {
    using S = Sweep<P, sizeof(T)>;
    const int32_t *gnb = mp.gnb[sizeof(T) == 8 ? 0 : 1];
    constexpr int NP1 = S::NP1, NG = S::NG, BRX = S::BRX, BRY = S::BRY, BNX = S::BNX, BNY = S::BNY, NAX = S::NAX,
                  NCOL = S::NCOL;
    constexpr bool PC = S::PER_COLUMN;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int col = PC ? tid : wave * S::CPW + lane / NG;
    const int g2l = PC ? 0 : lane % NG;               // the thread's g2 (lanes per g2 only)
    const bool valid = PC ? col < NCOL : (lane < S::CPW * NG && col < NCOL);
    const int a = col % NAX, b = col / NAX;        // Gauss positions (x, y) of the column
    const int e1l = a / NG, g0 = a % NG, e2l = b / NG, g1 = b % NG;
    const int b1 = (int)(brick % gnb[0]), b2 = (int)((brick / gnb[0]) % gnb[1]),
              b3 = (int)(brick / ((int64_t)gnb[0] * gnb[1]));
    const int E1 = b1 * BRX, E2 = b2 * BRY, E3 = b3 * S::BZ;
    const int ne1 = min(BRX, mp.N[0] - E1), ne2 = min(BRY, mp.N[1] - E2), ne3 = min(S::BZ, mp.N[2] - E3);
    const int nz = ne3 + P;
    const bool inside = valid && e1l < ne1 && e2l < ne2;
    const T lam = (T)mp.lam, mu = (T)mp.mu;
    T *out = partial + brick * (int64_t)S::PART;
    const T *geo_b = geo + brick * (int64_t)S::BZ * S::NGEO * S::NTQ + (int64_t)col * NG + g2l;

    // 1D tables of the brick's elements; zero past the patch, so contractions there give 0
    {
        const int Eb[3] = {E1, E2, E3}, nb[3] = {ne1, ne2, ne3};
        constexpr int NTAB = 3 * 2 * S::BZ * NG * NP1;
        for (int i = tid; i < NTAB; i += S::NT) {
            const int r = i % NP1, g = (i / NP1) % NG, le = (i / (NP1 * NG)) % S::BZ, wh = (i / (NP1 * NG * S::BZ)) % 2,
                      d = i / (NP1 * NG * S::BZ * 2);
            T v = T(0);
            if (le < nb[d]) {
                const int64_t o = ((int64_t)(mp.boff[d] + Eb[d] + le) * NG + g) * NP1 + r;
                v = (T)(wh ? mp.dBtab[o] : mp.Btab[o]);
            }
            sm.tab[d][wh][le][g][r] = v;
        }
    }
    T W[NP1][9];   // back-contractions for node layers e3 .. e3 + P of the thread's Gauss point
    // one thread per column: its stage-B outputs of node layers n3 - P .. n3 (Zr[k] = layer
    // n3 - P + k), shifted down as a layer comes in
    T Zr[PC ? NP1 : 1][9];
#pragma unroll
    for (int k = 0; k < NP1; ++k)
#pragma unroll
        for (int o = 0; o < 9; ++o) W[k][o] = T(0);
#pragma unroll
    for (int k = 0; k < (PC ? NP1 : 1); ++k)
#pragma unroll
        for (int o = 0; o < 9; ++o) Zr[k][o] = T(0);

    // back-contract node layer n3 from W[0] (all threads), write it, shift W down
    auto backward = [&](int n3) {
        // sum over the NG lanes of the column (adjacent lanes of one wave), or the thread's own
        // sum over g2 (one thread per column)
        T s[9];
#pragma unroll
        for (int o = 0; o < 9; ++o) {
            T v = W[0][o];
            if constexpr (!PC) {
#pragma unroll
                for (int k = 1; k < NG; ++k) v += __shfl(W[0][o], lane + k, 64);
            }
            s[o] = v;
        }
        if (valid && g2l == 0)
#pragma unroll
            for (int o = 0; o < 9; ++o) sm.Ws[o][col] = s[o];
#pragma unroll
        for (int k = 0; k < P; ++k)
#pragma unroll
            for (int o = 0; o < 9; ++o) W[k][o] = W[k + 1][o];
#pragma unroll
        for (int o = 0; o < 9; ++o) W[P][o] = T(0);
        __syncthreads();
        // B^T: (a, n2, f) <- columns (a, e2 NG + g1) with n2 = e2 + j
        for (int it = tid; it < NAX * BNY * 3; it += S::NT) {
            const int aa = it % NAX, n2 = (it / NAX) % BNY, f = it / (NAX * BNY);
            T pv = T(0), r0 = T(0);
            for (int j = 0; j < NP1; ++j) {
                const int e2 = n2 - j;
                if (e2 < 0 || e2 >= BRY) continue;
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    const int cc = aa + NAX * (e2 * NG + g);
                    const T bv = sm.tab[1][0][e2][g][j], dv = sm.tab[1][1][e2][g][j];
                    pv += sm.Ws[f * 3 + 0][cc] * bv + sm.Ws[f * 3 + 2][cc] * dv;
                    r0 += sm.Ws[f * 3 + 1][cc] * bv;
                }
            }
            sm.Pb[f][0][n2][aa] = pv;
            sm.Pb[f][1][n2][aa] = r0;
        }
        __syncthreads();
        // A^T: component f of node (n1, n2) of layer n3 <- positions a = e1 NG + g0 with
        // n1 = e1 + i (one thread per node and component)
        for (int it = tid; it < BNX * BNY * 3; it += S::NT) {
            const int f = it / (BNX * BNY), nd = it - f * (BNX * BNY);
            const int n1 = nd % BNX, n2 = nd / BNX;
            if (n1 >= ne1 + P || n2 >= ne2 + P) continue;
            T y = T(0);
            for (int i = 0; i < NP1; ++i) {
                const int e1 = n1 - i;
                if (e1 < 0 || e1 >= BRX) continue;
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    const T bv = sm.tab[0][0][e1][g][i], dv = sm.tab[0][1][e1][g][i];
                    y += sm.Pb[f][0][n2][e1 * NG + g] * bv + sm.Pb[f][1][n2][e1 * NG + g] * dv;
                }
            }
            const T w = sm.Wn[n3 % (NP1 + 1)][n2][n1];
            out[((int64_t)(n3 * BNY + n2) * BNX + n1) * 3 + f] = y * w;
        }
    };

    // Node-layer loads run one layer ahead (the CP ids two layers ahead), the geometry of the
    // next element layer one layer ahead, so their latency overlaps the contractions in between.
    // Thread t < BNX BNY loads node (t mod BNX, t / BNX) of each layer.
    static_assert(BNX * BNY <= S::NT, "one thread per node of a layer");
    const int ln1 = tid % BNX, ln2 = tid / BNX;
    const bool loader = tid < BNX * BNY && ln1 < ne1 + P && ln2 < ne2 + P;
    const int64_t cpl0 = (E1 + ln1) + (int64_t)mp.ncp[0] * ((E2 + ln2) + (int64_t)mp.ncp[1] * E3);
    const int64_t cpl_dz = (int64_t)mp.ncp[0] * mp.ncp[1];
    int cp_next = 0;          // CP of the loader's node in layer n3 + 1
    T xr[3] = {T(0), T(0), T(0)}, wr = T(0);   // its w and x in layer n3
    if (loader) {
        const int cp0 = mp.cpmap[cpl0];
        wr = wcp[cp0];
        xr[0] = x[3 * (int64_t)cp0];
        xr[1] = x[3 * (int64_t)cp0 + 1];
        xr[2] = x[3 * (int64_t)cp0 + 2];
        if (nz > 1) cp_next = mp.cpmap[cpl0 + cpl_dz];
    }
    // Geometry of the thread's Gauss point in element layer n3 - P: loaded at the start of node
    // layer n3, so the loads have the node-layer stages to arrive; the registers are live only
    // where the quadrature-point work needs them anyway (loading the next layer's instead
    // would keep 13 more registers live there and cost a wave per SIMD).
    T gv[S::NGEO];

    for (int n3 = 0; n3 < nz; ++n3) {
        const int slot = n3 % NP1;
        if (!PC && n3 >= P && inside) {
            const T *gq = geo_b + (int64_t)(n3 - P) * S::NGEO * S::NTQ;
#pragma unroll
            for (int c = 0; c < S::NGEO; ++c) gv[c] = gq[c * S::NTQ];
        }
        // ---- forward: node layer n3 into ring slot n3 mod (P+1) ----
        if (tid < BNX * BNY) {
            sm.X[0][ln2][ln1] = wr * xr[0];
            sm.X[1][ln2][ln1] = wr * xr[1];
            sm.X[2][ln2][ln1] = wr * xr[2];
            sm.Wn[n3 % (NP1 + 1)][ln2][ln1] = wr;
        }
        if (loader && n3 + 1 < nz) {
            const int cp = cp_next;
            wr = wcp[cp];
            xr[0] = x[3 * (int64_t)cp];
            xr[1] = x[3 * (int64_t)cp + 1];
            xr[2] = x[3 * (int64_t)cp + 2];
            if (n3 + 2 < nz) cp_next = mp.cpmap[cpl0 + (int64_t)(n3 + 2) * cpl_dz];
        }
        __syncthreads();
        // stage A: (a, n2, f) = sum_i X[f][n2][e1 + i] (B | dB)[e1][g0][i]
        for (int it = tid; it < NAX * BNY * 3; it += S::NT) {
            const int aa = it % NAX, n2 = (it / NAX) % BNY, f = it / (NAX * BNY);
            const int e1 = aa / NG, g = aa % NG;
            T v = T(0), d = T(0);
#pragma unroll
            for (int i = 0; i < NP1; ++i) {
                const T xv = sm.X[f][n2][e1 + i];
                v += xv * sm.tab[0][0][e1][g][i];
                d += xv * sm.tab[0][1][e1][g][i];
            }
            sm.Y[f][0][n2][aa] = v;
            sm.Y[f][1][n2][aa] = d;
        }
        __syncthreads();
        // stage B: column (a, b), outputs o = 3 f + var (value, d/dxi0, d/dxi1): into the
        // thread's registers (one thread per column), or dealt over the column's NG lanes
        if constexpr (PC) {
#pragma unroll
            for (int k = 0; k < P; ++k)
#pragma unroll
                for (int o = 0; o < 9; ++o) Zr[k][o] = Zr[k + 1][o];
            if (valid) {
#pragma unroll
                for (int o = 0; o < 9; ++o) {
                    const int f = o / 3, var = o % 3;
                    const int src = var == 1 ? 1 : 0;
                    T s = T(0);
#pragma unroll
                    for (int j = 0; j < NP1; ++j)
                        s += sm.Y[f][src][e2l + j][a] * sm.tab[1][var == 2 ? 1 : 0][e2l][g1][j];
                    Zr[P][o] = s;
                }
            }
            // no barrier: the ring is the thread's own, and Y is rewritten only after the next
            // layer's first barrier
        } else {
            if (valid) {
                for (int o = g2l; o < 9; o += NG) {
                    const int f = o / 3, var = o % 3;
                    const int src = var == 1 ? 1 : 0;
                    T s = T(0);
#pragma unroll
                    for (int j = 0; j < NP1; ++j)
                        s += sm.Y[f][src][e2l + j][a] * sm.tab[1][var == 2 ? 1 : 0][e2l][g1][j];
                    sm.Z[slot][o][col] = s;
                }
            }
            __syncthreads();
        }
        const int e3 = n3 - P;
        if (e3 < 0) continue;
        // ---- element layer e3: the thread's Gauss point (column, g2), or each g2 of its column ----
        if (inside) {
#pragma unroll 1
          for (int g2 = g2l; g2 < (PC ? NG : g2l + 1); ++g2) {
            if constexpr (PC) {
                const T *gq = geo_b + (int64_t)e3 * S::NGEO * S::NTQ + g2;
#pragma unroll
                for (int c = 0; c < S::NGEO; ++c) gv[c] = gq[c * S::NTQ];
            }
            T V[3], G[3][3];   // G[m][f]: d/dxi_m of component f
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                T v = T(0), d0 = T(0), d1 = T(0), d2 = T(0);
#pragma unroll
                for (int k = 0; k < NP1; ++k) {
                    const int sk = (e3 + k) % NP1;
                    const T bz = sm.tab[2][0][e3][g2][k], dz = sm.tab[2][1][e3][g2][k];
                    const T zv = PC ? Zr[k][f * 3 + 0] : sm.Z[sk][f * 3 + 0][col];
                    v += zv * bz;
                    d2 += zv * dz;
                    d0 += (PC ? Zr[k][f * 3 + 1] : sm.Z[sk][f * 3 + 1][col]) * bz;
                    d1 += (PC ? Zr[k][f * 3 + 2] : sm.Z[sk][f * 3 + 2][col]) * bz;
                }
                V[f] = v;
                G[0][f] = d0;
                G[1][f] = d1;
                G[2][f] = d2;
            }
            T A[9], gw[3];
#pragma unroll
            for (int c = 0; c < 9; ++c) A[c] = gv[c];
            gw[0] = gv[9];
            gw[1] = gv[10];
            gw[2] = gv[11];
            const T om = gv[12];
            // H[l][k] = d u_l / d x_k = sum_m (G[m][l] - gw[m] V[l]) A[m][k]
            T H[3][3];
#pragma unroll
            for (int l = 0; l < 3; ++l)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    T s = T(0);
#pragma unroll
                    for (int m = 0; m < 3; ++m) s += (G[m][l] - gw[m] * V[l]) * A[3 * m + k];
                    H[l][k] = s;
                }
            const T tr = lam * (H[0][0] + H[1][1] + H[2][2]);
            // coefficients of dB/dxi_m (Tm) and of B (tb) per component, into the z back-contraction
            T Tm[3][3], tb[3];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                T sg[3];
#pragma unroll
                for (int l = 0; l < 3; ++l) sg[l] = mu * (H[f][l] + H[l][f]) + (f == l ? tr : T(0));
                T t0 = T(0);
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    T s = T(0);
#pragma unroll
                    for (int l = 0; l < 3; ++l) s += sg[l] * A[3 * m + l];
                    s *= om;
                    Tm[f][m] = s;
                    t0 -= s * gw[m];
                }
                tb[f] = t0;
            }
#pragma unroll
            for (int k = 0; k < NP1; ++k) {
                const T bz = sm.tab[2][0][e3][g2][k], dz = sm.tab[2][1][e3][g2][k];
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    W[k][f * 3 + 0] += tb[f] * bz + Tm[f][2] * dz;
                    W[k][f * 3 + 1] += Tm[f][0] * bz;
                    W[k][f * 3 + 2] += Tm[f][1] * bz;
                }
            }
          }
        }
        backward(e3);
    }
    // the last P node layers have all their contributions
    for (int n3 = ne3; n3 < nz; ++n3) backward(n3);
}

}  // namespace yeti
