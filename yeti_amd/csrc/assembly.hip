// assembly.hip -- FP64 stiffness / RHS assembly of 3D NURBS solid elements (YETI type U1) on
// MI355X (gfx950), behind the C ABI of include/yeti_amd.h.
//
// Reference path (paths relative to YETI's src/yeti_iga/):
//   sys_linmat_lindef_static         main/build_system_elemStorage.f90:24-315
//   sys_linmat_lindef_static_omp     main/build_system_elemStorage_OMP.f90:26-289
//   UELMAT_byCP / uelmat_byCP_omp    main/UELMAT.f90:22-225 / main/uelmat_solid_omp.f90:1-200
//   shap, dersbasisfuns              main/shap.f90:21-180, main/dersbasisfuns.f90:30-131
//   stiffmatrix_byCP, material_lib   main/stiffmatrix.F90:130-225, main/material_lib.f90:28-98
//   shapPress                        main/shapPress.f:18-265
//
// Element kernel (one workgroup per element):
//   1. stage the element's CP coordinates, weights, connectivity and the 1D B-spline values /
//      derivatives at its Gauss abscissae in LDS;
//   2. per Gauss point, reduce W, dW, sum x N, sum x (x) dN over the nodes -> J^-1, |det J| w;
//   3. write G_i[q][a] = d R_a / d x_i (q = Gauss point, a = node) to LDS;
//   4. S^{ij} = G_i^T diag(w|J|) G_j on v_mfma_f64_16x16x4f64: one wave owns whole 16x16 node
//      tiles (I <= J) and all nine (i, j) component tiles of them, so the isotropic
//      contraction K_ab[k][l] = lam S^{kl}_ab + mu S^{lk}_ab + mu d_kl tr S_ab is lane-local;
//   5. emit the element's upper-triangle COO slice in YETI's order (nodj-major, then nodi, then
//      the 3x3 block column-major; diagonal blocks: the 6 upper entries, diagonal halved).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace yeti {

std::string &last_error() {
    static thread_local std::string s;
    return s;
}

AllocCtx &alloc_ctx() {
    static thread_local AllocCtx c;
    return c;
}

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    last_error() = buf;
    return code;
}

// ------------------------------------------------------------------------------------------
// Gauss-Legendre rule (same nodes/weights as main/GaussLegendre.f90:27-160, computed by Newton
// iteration on P_n; agrees with the 30-digit table to ~1 ulp).
static void gauss_legendre(int n, double *x, double *w) {
    for (int i = 0; i < n; ++i) {
        double z = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            double dp = (n == 1) ? 1.0 : n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-17) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = pk;
        }
        double dp = (n == 1) ? 1.0 : n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
    for (int i = 0; i < n / 2; ++i) {
        double xs = 0.5 * (x[n - 1 - i] - x[i]);
        double ws = 0.5 * (w[i] + w[n - 1 - i]);
        x[i] = -xs;
        x[n - 1 - i] = xs;
        w[i] = ws;
        w[n - 1 - i] = ws;
    }
}

void gauss_rule_host(int n, double *x, double *w) { gauss_legendre(n, x, w); }


// ------------------------------------------------------------------------------------------
// Device: non-zero B-splines of degree P and first derivatives at u for 0-based span i
// (Cox-de Boor triangle + closed-form derivative; the contract of dersbasisfuns.f90:30-131).
template <int P>
__device__ inline void bspline_ders(int i, double u, const double *__restrict__ U, double *N,
                                    double *dN) {
    double left[P + 1], right[P + 1], cur[P + 1], low[P + 1];
    cur[0] = 1.0;
    low[0] = 1.0;
#pragma unroll
    for (int j = 1; j <= P; ++j) {
        left[j] = u - U[i + 1 - j];
        right[j] = U[i + j] - u;
        double saved = 0.0;
#pragma unroll
        for (int r = 0; r < j; ++r) {
            double tmp = cur[r] / (right[r + 1] + left[j - r]);
            cur[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        cur[j] = saved;
        if (j == P - 1) {
#pragma unroll
            for (int r = 0; r < P; ++r) low[r] = cur[r];
        }
    }
#pragma unroll
    for (int r = 0; r <= P; ++r) {
        N[r] = cur[r];
        int k = i - P + r;
        double d = 0.0;
        if (r >= 1) {
            double den = U[k + P] - U[k];
            if (den != 0.0) d += P * low[r - 1] / den;
        }
        if (r <= P - 1) {
            double den = U[k + P + 1] - U[k + 1];
            if (den != 0.0) d -= P * low[r] / den;
        }
        dN[r] = d;
    }
}

#include "elem_kernel.hpp"

// ------------------------------------------------------------------------------------------
// Right-hand side of loaded elements (UELMAT.f90:121-223): body loads 101/102 at the volume
// Gauss points, surface loads 10..99 at the face Gauss points. One 64-thread block per loaded
// element; each thread owns nodes a = lane, lane + 64, ...; block-wide sums go through LDS.
struct LoadTable {
    const int32_t *jdltype;        // [nb_load]
    const double *adlmag;          // [nb_load]
    const int32_t *add_off;        // [nb_load] offset into add_infos
    const double *add_infos;
    const int32_t *elem_loads;     // CSR: loads acting on each loaded element
    const int32_t *elem_loads_ptr; // [n_loaded + 1]
    const double *nodal_dist;      // (nb_n_dist, nb_cp)
    int32_t nb_n_dist;
    int32_t only_type = 0;         // load-gradient runs: 0 = every load, else only this JDLType
};

struct RhsPatch {
    const double *coords;
    const int32_t *ien;
    const double *weight;
    const int32_t *nijk;
    const double *U[3];
    const int32_t *elems;          // loaded elements of the patch (patch-local)
    int32_t n_loaded;
    int32_t slot0;                 // index of this patch's first loaded element in the RHS buffer
    int32_t p[3];
    int32_t ng;
    double rho;
    GaussRule g;
};

template <int MAXN>
__device__ void rhs_basis(const RhsPatch &rp, int e, const double *pt, int nn, const double *sW,
                          const double (*sX)[3], double *R, double (*dR)[3], double AJ[3][3], double *detJ) {
    // R[c] (and, if dR is given, dR[c][m] = dR/d(parent coordinate m)) for the nodes a = lane + 64 c
    // parametric point and 1D bases (general degree up to 8)
    const int lane = threadIdx.x;
    __shared__ double sb[3][10], sdb[3][10], shalf[3];
    if (lane < 3) {
        int d = lane;
        int span1 = rp.nijk[3 * (int64_t)e + d];
        const double *U = rp.U[d];
        double u0 = U[span1 - 1], u1 = U[span1];
        double u = ((u1 - u0) * pt[d] + (u1 + u0)) * 0.5;
        shalf[d] = 0.5 * (u1 - u0);
        int p = rp.p[d];
        int i = span1 - 1;
        double left[10], right[10], cur[10], low[10];
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
            int k = i - p + r;
            double dd = 0.0;
            if (r >= 1) {
                double den = U[k + p] - U[k];
                if (den != 0.0) dd += p * low[r - 1] / den;
            }
            if (r <= p - 1) {
                double den = U[k + p + 1] - U[k + 1];
                if (den != 0.0) dd -= p * low[r] / den;
            }
            sb[d][p - r] = cur[r];   // reversed local index
            sdb[d][p - r] = dd;
        }
    }
    __syncthreads();
    const int n0 = rp.p[0] + 1, n1 = rp.p[1] + 1;
    double s[16];
    for (int m = 0; m < 16; ++m) s[m] = 0.0;
    double Nl[MAXN / 64 + 1], dNl[MAXN / 64 + 1][3];
    int c = 0;
    for (int a = lane; a < nn; a += 64, ++c) {
        int i = a % n0, j = (a / n0) % n1, k = a / (n0 * n1);
        double w = sW[a];
        double Na = sb[0][i] * sb[1][j] * sb[2][k] * w;
        double d0 = sdb[0][i] * sb[1][j] * sb[2][k] * w;
        double d1 = sb[0][i] * sdb[1][j] * sb[2][k] * w;
        double d2 = sb[0][i] * sb[1][j] * sdb[2][k] * w;
        Nl[c] = Na;
        dNl[c][0] = d0; dNl[c][1] = d1; dNl[c][2] = d2;
        s[0] += Na; s[1] += d0; s[2] += d1; s[3] += d2;
        for (int r = 0; r < 3; ++r) {
            s[4 + r] += sX[a][r] * Na;
            s[7 + 3 * r] += sX[a][r] * d0;
            s[8 + 3 * r] += sX[a][r] * d1;
            s[9 + 3 * r] += sX[a][r] * d2;
        }
    }
    for (int m = 0; m < 16; ++m) {
        double v = s[m];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        s[m] = v;
    }
    double Wi = 1.0 / s[0];
    double dw[3] = {s[1] * Wi, s[2] * Wi, s[3] * Wi};
    double J[3][3];
    for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) J[r][cc] = (s[7 + 3 * r + cc] - s[4 + r] * dw[cc]) * Wi;
    for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) AJ[r][cc] = J[r][cc] * shalf[cc];
    *detJ = AJ[0][0] * (AJ[1][1] * AJ[2][2] - AJ[1][2] * AJ[2][1])
          - AJ[0][1] * (AJ[1][0] * AJ[2][2] - AJ[1][2] * AJ[2][0])
          + AJ[0][2] * (AJ[1][0] * AJ[2][1] - AJ[1][1] * AJ[2][0]);
    c = 0;
    for (int a = lane; a < nn; a += 64, ++c) {
        R[c] = Nl[c] * Wi;
        if (dR)
            for (int m = 0; m < 3; ++m) dR[c][m] = (dNl[c][m] - Nl[c] * dw[m]) * Wi * shalf[m];
    }
    __syncthreads();
}

static __device__ inline void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

constexpr int RHS_MAXN = 729;   // (8+1)^3

// Dynamic LDS of elem_rhs_kernel: the element's node coordinates, weights and CP ids (sized by
// the patch's node count instead of the largest degree, so that many elements share a CU).
static size_t rhs_smem_bytes(const RhsPatch &rp) {
    const size_t nn = (size_t)(rp.p[0] + 1) * (rp.p[1] + 1) * (rp.p[2] + 1);
    return nn * (4 * sizeof(double) + sizeof(int32_t));
}

__global__ void __launch_bounds__(64) elem_rhs_kernel(RhsPatch rp, LoadTable lt, double *rhs_buf) {
    const int le = blockIdx.x;
    if (le >= rp.n_loaded) return;
    const int e = rp.elems[le];
    const int lane = threadIdx.x;
    const int nn = (rp.p[0] + 1) * (rp.p[1] + 1) * (rp.p[2] + 1);
    // the element's nodes, sized at launch (rhs_smem_bytes): coordinates, weights, CP ids
    extern __shared__ double rhs_smem[];
    double (*sX)[3] = reinterpret_cast<double (*)[3]>(rhs_smem);
    double *sW = rhs_smem + 3 * nn;
    int32_t *sIen = reinterpret_cast<int32_t *>(rhs_smem + 4 * nn);
    for (int a = lane; a < nn; a += 64) {
        int cp = rp.ien[(int64_t)nn * e + a];
        sIen[a] = cp;
        for (int d = 0; d < 3; ++d) sX[a][d] = rp.coords[3 * (int64_t)(cp - 1) + d];
        sW[a] = rp.weight[(int64_t)nn * e + a];
    }
    __syncthreads();
    constexpr int NPL = RHS_MAXN / 64 + 1;
    double rhs[NPL][3];
    for (int c = 0; c < NPL; ++c) rhs[c][0] = rhs[c][1] = rhs[c][2] = 0.0;
    double R[NPL];
    const int ng = rp.ng;
    const int gle = rp.slot0 + le;
    const int l0 = lt.elem_loads_ptr[gle], l1 = lt.elem_loads_ptr[gle + 1];
    bool any_body = false;
    for (int li = l0; li < l1; ++li) {
        int t = lt.jdltype[lt.elem_loads[li]];
        any_body |= (t == 101 || t == 102);
    }
    if (any_body) {
        for (int gi = 0; gi < ng; ++gi)
            for (int gj = 0; gj < ng; ++gj)
                for (int gk = 0; gk < ng; ++gk) {
                    double pt[3] = {rp.g.x[gi], rp.g.x[gj], rp.g.x[gk]};
                    double AJ[3][3], detJ;
                    rhs_basis<RHS_MAXN>(rp, e, pt, nn, sW, sX, R, nullptr, AJ, &detJ);
                    double dvol = rp.g.w[gi] * rp.g.w[gj] * rp.g.w[gk] * fabs(detJ);
                    for (int li = l0; li < l1; ++li) {
                        int il = lt.elem_loads[li];
                        int t = lt.jdltype[il];
                        if (t != 101 && t != 102) continue;
                        const double *A = lt.add_infos + lt.add_off[il];
                        double vr[3], mag;
                        if (t == 101) {
                            double gp[3] = {0, 0, 0};
                            int c = 0;
                            for (int a = lane; a < nn; a += 64, ++c)
                                for (int d = 0; d < 3; ++d) gp[d] += R[c] * sX[a][d];
                            for (int d = 0; d < 3; ++d)
                                for (int o = 32; o > 0; o >>= 1) gp[d] += __shfl_xor(gp[d], o, 64);
                            double vd[3] = {A[3] - A[0], A[4] - A[1], A[5] - A[2]};
                            double nd = sqrt(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
                            for (int d = 0; d < 3; ++d) vd[d] /= nd;
                            double ag[3] = {gp[0] - A[0], gp[1] - A[1], gp[2] - A[2]};
                            double sc = ag[0] * vd[0] + ag[1] * vd[1] + ag[2] * vd[2];
                            for (int d = 0; d < 3; ++d) vr[d] = ag[d] - sc * vd[d];
                            mag = rp.rho * lt.adlmag[il] * lt.adlmag[il];
                        } else {
                            for (int d = 0; d < 3; ++d) vr[d] = A[d];
                            mag = lt.adlmag[il];
                        }
                        int c = 0;
                        for (int a = lane; a < nn; a += 64, ++c)
                            for (int d = 0; d < 3; ++d) rhs[c][d] += mag * vr[d] * R[c] * dvol;
                    }
                }
    }
    for (int li = l0; li < l1; ++li) {
        int il = lt.elem_loads[li];
        int t = lt.jdltype[il];
        if (!(t > 9 && t < 100)) continue;
        int face = t / 10, typ = t % 10;
        int field = typ == 4 ? (int)lt.add_infos[lt.add_off[il]] : 0;
        for (int fj = 0; fj < ng; ++fj)
            for (int fi = 0; fi < ng; ++fi) {
                double xi_i = rp.g.x[fi], xi_j = rp.g.x[fj];
                double pt[3];
                switch (face) {
                case 1: pt[0] = -1.0; pt[1] = xi_i; pt[2] = xi_j; break;
                case 2: pt[0] = 1.0; pt[1] = xi_i; pt[2] = xi_j; break;
                case 3: pt[0] = xi_i; pt[1] = -1.0; pt[2] = xi_j; break;
                case 4: pt[0] = xi_i; pt[1] = 1.0; pt[2] = xi_j; break;
                case 5: pt[0] = xi_j; pt[1] = xi_i; pt[2] = -1.0; break;
                default: pt[0] = xi_j; pt[1] = xi_i; pt[2] = 1.0; break;
                }
                double AJ[3][3], detJ;
                rhs_basis<RHS_MAXN>(rp, e, pt, nn, sW, sX, R, nullptr, AJ, &detJ);
                double c1[3] = {AJ[0][0], AJ[1][0], AJ[2][0]};
                double c2[3] = {AJ[0][1], AJ[1][1], AJ[2][1]};
                double c3[3] = {AJ[0][2], AJ[1][2], AJ[2][2]};
                double nrm[3];
                if (face <= 2) cross3(c2, c3, nrm);
                else if (face <= 4) cross3(c3, c1, nrm);
                else cross3(c1, c2, nrm);
                double dS = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
                double v[3] = {0.0, 0.0, 0.0};
                if (typ == 1) v[0] = 1.0;
                else if (typ == 2) v[1] = 1.0;
                else if (typ == 3) v[2] = 1.0;
                else if (typ == 0 || typ == 4) {
                    const double *A, *B;
                    switch (face) {
                    case 1: A = c2; B = c3; break;
                    case 2: A = c3; B = c2; break;
                    case 3: A = c3; B = c1; break;
                    case 4: A = c1; B = c3; break;
                    case 5: A = c1; B = c2; break;
                    default: A = c2; B = c1; break;
                    }
                    cross3(A, B, v);
                    for (int d = 0; d < 3; ++d) v[d] /= dS;
                }
                double dvol = rp.g.w[fi] * rp.g.w[fj] * dS;
                double fmag;
                if (typ == 4) {
                    double fm = 0.0;
                    int c = 0;
                    for (int a = lane; a < nn; a += 64, ++c)
                        fm += lt.nodal_dist[(int64_t)lt.nb_n_dist * (sIen[a] - 1) + (field - 1)] * R[c];
                    for (int o = 32; o > 0; o >>= 1) fm += __shfl_xor(fm, o, 64);
                    fmag = fm;
                } else {
                    fmag = lt.adlmag[il];
                }
                int c = 0;
                for (int a = lane; a < nn; a += 64, ++c)
                    for (int d = 0; d < 3; ++d) rhs[c][d] += R[c] * v[d] * dvol * fmag;
            }
    }
    int c = 0;
    for (int a = lane; a < nn; a += 64, ++c)
        for (int d = 0; d < 3; ++d) rhs_buf[(int64_t)gle * 3 * RHS_MAXN + 3 * a + d] = rhs[c][d];
}

// Shape derivative of the surface-load work, per loaded element and node: g_cj = 2 u^T dF/dX_cj
// for the face loads of elem_rhs_kernel (same faces, quadrature and load types). With the face
// tangents c_m = dx/dxi_m (columns of AJ), dc_m/dX_cj = e_j dR_c/dxi_m, so
//   fixed direction v (types 1-3), F_a = R_a v w dS f, dS = |cA x cB|:
//     d dS/dX_cj = dR_c/dxi_A (cB x n)_j + dR_c/dxi_B (n x cA)_j,  n = (cA x cB)/dS
//   normal load (types 0 and 4), F_a = R_a (A x B) w f:
//     u . d(A x B)/dX_cj = dR_c/dxi_A (B x u)_j + dR_c/dxi_B (u x A)_j
// with u = u(xi_q) = sum_a R_a u_a. Body loads (101 centrifugal, 102 constant) at the volume
// Gauss points as elem_rhs_kernel, derivative below.
__global__ void __launch_bounds__(64) elem_load_grad_kernel(RhsPatch rp, LoadTable lt, const double *__restrict__ u,
                                                            double *grad_buf) {
    const int le = blockIdx.x;
    if (le >= rp.n_loaded) return;
    const int e = rp.elems[le];
    const int lane = threadIdx.x;
    const int nn = (rp.p[0] + 1) * (rp.p[1] + 1) * (rp.p[2] + 1);
    __shared__ double sX[RHS_MAXN][3];
    __shared__ double sW[RHS_MAXN];
    __shared__ double sU[RHS_MAXN][3];
    __shared__ int32_t sIen[RHS_MAXN];
    for (int a = lane; a < nn; a += 64) {
        int cp = rp.ien[(int64_t)nn * e + a];
        sIen[a] = cp;
        for (int d = 0; d < 3; ++d) {
            sX[a][d] = rp.coords[3 * (int64_t)(cp - 1) + d];
            sU[a][d] = u[3 * (int64_t)(cp - 1) + d];
        }
        sW[a] = rp.weight[(int64_t)nn * e + a];
    }
    __syncthreads();
    constexpr int NPL = RHS_MAXN / 64 + 1;
    double g[NPL][3];
    for (int c = 0; c < NPL; ++c) g[c][0] = g[c][1] = g[c][2] = 0.0;
    double R[NPL], dR[NPL][3];
    const int ng = rp.ng;
    const int gle = rp.slot0 + le;
    const int l0 = lt.elem_loads_ptr[gle], l1 = lt.elem_loads_ptr[gle + 1];
    bool any_body = false;
    for (int li = l0; li < l1; ++li) {
        int t = lt.jdltype[lt.elem_loads[li]];
        any_body |= (t == 101 || t == 102);
    }
    // body loads, u . F = sum_q w |J| u(x_q) . f(x_q):
    //   d/dX_cj = w |J| [ (u . f) dR_c/dx_j + R_c (u . df/dx)_j ],
    // d|J|/dX_cj = |J| dR_c/dx_j; f = rho omega^2 (I - d d^T)(x - A) for 101 (so
    // u . df/dx = rho omega^2 (I - d d^T) u), a constant vector for 102.
    if (any_body) {
        for (int gi = 0; gi < ng; ++gi)
            for (int gj = 0; gj < ng; ++gj)
                for (int gk = 0; gk < ng; ++gk) {
                    double pt[3] = {rp.g.x[gi], rp.g.x[gj], rp.g.x[gk]};
                    double AJ[3][3], detJ;
                    rhs_basis<RHS_MAXN>(rp, e, pt, nn, sW, sX, R, dR, AJ, &detJ);
                    const double dvol = rp.g.w[gi] * rp.g.w[gj] * rp.g.w[gk] * fabs(detJ);
                    // rows of the inverse Jacobian: dxi_m/dx = (col_{m+1} x col_{m+2}) / detJ
                    double col[3][3], cof[3][3];
                    for (int m = 0; m < 3; ++m)
                        for (int d = 0; d < 3; ++d) col[m][d] = AJ[d][m];
                    for (int m = 0; m < 3; ++m) cross3(col[(m + 1) % 3], col[(m + 2) % 3], cof[m]);
                    double uq[3] = {0.0, 0.0, 0.0}, xq[3] = {0.0, 0.0, 0.0};
                    int c = 0;
                    for (int a = lane; a < nn; a += 64, ++c)
                        for (int d = 0; d < 3; ++d) {
                            uq[d] += R[c] * sU[a][d];
                            xq[d] += R[c] * sX[a][d];
                        }
                    for (int o = 32; o > 0; o >>= 1)
                        for (int d = 0; d < 3; ++d) {
                            uq[d] += __shfl_xor(uq[d], o, 64);
                            xq[d] += __shfl_xor(xq[d], o, 64);
                        }
                    double uf = 0.0, hv[3] = {0.0, 0.0, 0.0};
                    for (int li = l0; li < l1; ++li) {
                        const int il = lt.elem_loads[li];
                        const int t = lt.jdltype[il];
                        if (lt.only_type && t != lt.only_type) continue;
                        const double *A = lt.add_infos + lt.add_off[il];
                        if (t == 101) {
                            double vd[3] = {A[3] - A[0], A[4] - A[1], A[5] - A[2]};
                            const double nd = sqrt(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
                            for (int d = 0; d < 3; ++d) vd[d] /= nd;
                            const double ag[3] = {xq[0] - A[0], xq[1] - A[1], xq[2] - A[2]};
                            const double sc = ag[0] * vd[0] + ag[1] * vd[1] + ag[2] * vd[2];
                            const double ud = uq[0] * vd[0] + uq[1] * vd[1] + uq[2] * vd[2];
                            const double mag = rp.rho * lt.adlmag[il] * lt.adlmag[il];
                            for (int d = 0; d < 3; ++d) {
                                uf += mag * (ag[d] - sc * vd[d]) * uq[d];
                                hv[d] += mag * (uq[d] - ud * vd[d]);
                            }
                        } else if (t == 102) {
                            uf += lt.adlmag[il] * (A[0] * uq[0] + A[1] * uq[1] + A[2] * uq[2]);
                        }
                    }
                    const double wq = 2.0 * dvol;
                    c = 0;
                    for (int a = lane; a < nn; a += 64, ++c)
                        for (int d = 0; d < 3; ++d) {
                            const double dRdx =
                                (dR[c][0] * cof[0][d] + dR[c][1] * cof[1][d] + dR[c][2] * cof[2][d]) / detJ;
                            g[c][d] += wq * (uf * dRdx + R[c] * hv[d]);
                        }
                }
    }
    for (int li = l0; li < l1; ++li) {
        int il = lt.elem_loads[li];
        int t = lt.jdltype[il];
        if (!(t > 9 && t < 100) || (lt.only_type && t != lt.only_type)) continue;
        int face = t / 10, typ = t % 10;
        int field = typ == 4 ? (int)lt.add_infos[lt.add_off[il]] : 0;
        // columns spanning the face, in the order of elem_rhs_kernel's normal
        const int nA = face <= 2 ? 1 : face <= 4 ? 2 : 0, nB = face <= 2 ? 2 : face <= 4 ? 0 : 1;
        // (A, B) of the normal load per face
        int mA, mB;
        switch (face) {
        case 1: mA = 1; mB = 2; break;
        case 2: mA = 2; mB = 1; break;
        case 3: mA = 2; mB = 0; break;
        case 4: mA = 0; mB = 2; break;
        case 5: mA = 0; mB = 1; break;
        default: mA = 1; mB = 0; break;
        }
        for (int fj = 0; fj < ng; ++fj)
            for (int fi = 0; fi < ng; ++fi) {
                double xi_i = rp.g.x[fi], xi_j = rp.g.x[fj];
                double pt[3];
                switch (face) {
                case 1: pt[0] = -1.0; pt[1] = xi_i; pt[2] = xi_j; break;
                case 2: pt[0] = 1.0; pt[1] = xi_i; pt[2] = xi_j; break;
                case 3: pt[0] = xi_i; pt[1] = -1.0; pt[2] = xi_j; break;
                case 4: pt[0] = xi_i; pt[1] = 1.0; pt[2] = xi_j; break;
                case 5: pt[0] = xi_j; pt[1] = xi_i; pt[2] = -1.0; break;
                default: pt[0] = xi_j; pt[1] = xi_i; pt[2] = 1.0; break;
                }
                double AJ[3][3], detJ;
                rhs_basis<RHS_MAXN>(rp, e, pt, nn, sW, sX, R, dR, AJ, &detJ);
                double col[3][3];
                for (int m = 0; m < 3; ++m)
                    for (int d = 0; d < 3; ++d) col[m][d] = AJ[d][m];
                double nrm[3];
                cross3(col[nA], col[nB], nrm);
                const double dS = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
                // u and the load field at the face point
                double uq[3] = {0.0, 0.0, 0.0}, fm = 0.0;
                int c = 0;
                for (int a = lane; a < nn; a += 64, ++c) {
                    for (int d = 0; d < 3; ++d) uq[d] += R[c] * sU[a][d];
                    if (typ == 4) fm += lt.nodal_dist[(int64_t)lt.nb_n_dist * (sIen[a] - 1) + (field - 1)] * R[c];
                }
                for (int o = 32; o > 0; o >>= 1) {
                    for (int d = 0; d < 3; ++d) uq[d] += __shfl_xor(uq[d], o, 64);
                    fm += __shfl_xor(fm, o, 64);
                }
                const double fmag = typ == 4 ? fm : lt.adlmag[il];
                const double wq = 2.0 * rp.g.w[fi] * rp.g.w[fj] * fmag;
                double tA[3], tB[3];   // g_c = wq (dR_c/dxi_A tA + dR_c/dxi_B tB)
                int cA, cB;
                if (typ >= 1 && typ <= 3) {
                    const double un = uq[typ - 1];   // u . v
                    double n[3] = {nrm[0] / dS, nrm[1] / dS, nrm[2] / dS};
                    cross3(col[nB], n, tA);
                    cross3(n, col[nA], tB);
                    for (int d = 0; d < 3; ++d) {
                        tA[d] *= un;
                        tB[d] *= un;
                    }
                    cA = nA;
                    cB = nB;
                } else {
                    cross3(col[mB], uq, tA);
                    cross3(uq, col[mA], tB);
                    cA = mA;
                    cB = mB;
                }
                c = 0;
                for (int a = lane; a < nn; a += 64, ++c)
                    for (int d = 0; d < 3; ++d) g[c][d] += wq * (dR[c][cA] * tA[d] + dR[c][cB] * tB[d]);
            }
    }
    int c = 0;
    for (int a = lane; a < nn; a += 64, ++c)
        for (int d = 0; d < 3; ++d) grad_buf[(int64_t)gle * 3 * RHS_MAXN + 3 * a + d] = g[c][d];
}

// F = F_conc + sum over loaded elements (in element order) of their RHS (build_system_elemStorage.f90:295-303);
// F_conc may be null (then F starts from zero).
__global__ void rhs_gather_kernel(int32_t nb_cp, const double *__restrict__ F_conc,
                                  const int32_t *__restrict__ cp_ptr,
                                  const int32_t *__restrict__ cp_ent,
                                  const double *__restrict__ rhs_buf, double *__restrict__ F) {
    int cp = blockIdx.x * blockDim.x + threadIdx.x;
    if (cp >= nb_cp) return;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    if (F_conc) {
        f0 = F_conc[3 * (int64_t)cp];
        f1 = F_conc[3 * (int64_t)cp + 1];
        f2 = F_conc[3 * (int64_t)cp + 2];
    }
    for (int k = cp_ptr[cp]; k < cp_ptr[cp + 1]; ++k) {
        int ent = cp_ent[k];   // slot * RHS_MAXN + local node
        const double *r = rhs_buf + 3 * (int64_t)ent;
        f0 += r[0];
        f1 += r[1];
        f2 += r[2];
    }
    F[3 * (int64_t)cp] = f0;
    F[3 * (int64_t)cp + 1] = f1;
    F[3 * (int64_t)cp + 2] = f2;
}

// Zero a range of a COO buffer (slots of inactive elements in the SLOTTED layout).
__global__ void fill_zero_kernel(double *Kd, int32_t *Kr, int32_t *Kc, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        Kd[i] = 0.0;
        if (Kr) Kr[i] = 0;
        if (Kc) Kc[i] = 0;
    }
}

// ------------------------------------------------------------------------------------------
// Host side
struct PatchDev {
    int p[3];
    int ng;
    int nnode;
    int nelem;
    double lam, mu, rho;
    bool u1;
    DevBuf<double> U[3];
    DevBuf<int32_t> ien, nijk, elems;
    DevBuf<double> weight;
    DevBuf<int64_t> offs;
    int32_t n_active = 0;
    DevBuf<int32_t> loaded;
    int32_t n_loaded = 0;
    int32_t slot0 = 0;
    DevBuf<double> geo;     // per-Gauss-point geometry of one element batch (allocated on first run)
};

}  // namespace yeti

struct yeti_asm {
    // host copies of the per-patch device pointers (for the CSR builder)
    std::vector<const int32_t *> v_ien, v_elems;
    std::vector<const int64_t *> v_offs;
    std::vector<int32_t> v_nnode, v_nactive;
    int layout;
    int64_t nb_data;        // entries the layout uses
    int64_t nb_data_cap;    // nb_data requested by the caller
    int32_t nb_cp, nb_dof_tot;
    std::vector<yeti::PatchDev *> patches;
    yeti::DevBuf<double> coords;
    // zero ranges (slotted layout, inactive elements)
    std::vector<std::pair<int64_t, int64_t>> zero_ranges;
    // loads
    yeti::DevBuf<double> F_conc;
    yeti::DevBuf<int32_t> jdl, add_off, elem_loads, elem_loads_ptr, cp_ptr, cp_ent;
    yeti::DevBuf<double> adlmag, add_infos, nodal_dist, rhs_buf, grad_buf;
    int32_t nb_n_dist = 1;
    // optional per-launch timing of the element kernel (HIP events on the launch stream)
    bool timing = false;
    std::vector<hipEvent_t> ev;
    int32_t ev_used = 0;
    int32_t n_loaded_total = 0;
    // natural-frequency gradient (built on first use): CP -> element-node entries of the active
    // U1 elements in CSR form, per-entry and per-element partials, the per-mode mass
    yeti::DevBuf<int32_t> vib_ptr;
    yeti::DevBuf<int64_t> vib_ent;
    yeti::DevBuf<double> vib_eg, vib_em, vib_mass;
    yeti::DevBuf<double> vol_u;    // translation field of the volume gradient
    // gradLinElastWork_AN: offsets of the centrifugal loads' axis points in load_additionalInfos
    // as the reference reads them (the k-th load of type 101 at 6k), and the work buffers
    yeti::DevBuf<int32_t> add_off_ref;
    bool ref_infos_ok = true;
    yeti::DevBuf<double> gw_fields, gw_grads, gw_vec, gw_vec2, gw_sums;
    int64_t vib_nent = -1, vib_nslot = 0;
    ~yeti_asm() {
        for (auto *p : patches) delete p;
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

namespace yeti {

static int parse_elt_types(const yeti_system_inputs *in, std::vector<std::string> &out) {
    out.clear();
    std::string s(in->ELT_TYPE, in->ELT_TYPE + in->len_ELT_TYPE);
    size_t k = 0;
    while (k < s.size()) {
        if (s[k] == 'U') {
            size_t j = k + 1;
            while (j < s.size() && s[j] != 'U' && s[j] != ' ' && s[j] != '\0') ++j;
            out.push_back(s.substr(k, j - k));
            k = j;
        } else {
            ++k;
        }
    }
    if ((int)out.size() < in->nb_patch)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "ELT_TYPE lists %d types for %d patches",
                         (int)out.size(), in->nb_patch);
    return 0;
}

static int gauss_per_dir(int nbint) {
    int n = (int)std::pow((double)nbint, 1.0 / 3.0);
    if (n * n * n < nbint) n += 1;
    return n;
}

// the MFMA element kernel is built for equal degrees with p + 1 Gauss points per direction
static bool mfma_kernel_available(int p0, int p1, int p2, int ng) {
    if (p0 != p1 || p1 != p2) return false;
    return (p0 == 1 && ng == 2) || (p0 == 2 && ng == 3) || (p0 == 3 && ng == 4) || (p0 == 4 && ng == 5);
}

// every other case goes to elem_generic_kernel: up to 64 nodes, degrees up to 4, up to 9 points
static bool kernel_available(int p0, int p1, int p2, int ng) {
    if (mfma_kernel_available(p0, p1, p2, ng)) return true;
    return p0 >= 1 && p1 >= 1 && p2 >= 1 && p0 <= 4 && p1 <= 4 && p2 <= 4 && (p0 + 1) * (p1 + 1) * (p2 + 1) <= 64 &&
           ng >= 1 && ng <= MAX_NG;
}

static int device_cu_count() {
    static thread_local int cached_dev = -1, cached_n = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (dev != cached_dev) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached_dev = dev;
        cached_n = n;
    }
    return cached_n;
}

// Elements per launch pair (geometry pass + element kernel): bounds the geometry buffer
// (12 doubles per Gauss point: 400 MB at p = 3) while keeping every CU busy.
constexpr int32_t ELEM_BATCH = 65536;

// event pair k of the handle's timing record (created on first use)
static int timing_events(yeti_asm *h, int k, hipEvent_t *a, hipEvent_t *b) {
    while ((int)h->ev.size() < 2 * (k + 1)) {
        hipEvent_t e;
        YETI_HIP(hipEventCreate(&e));
        h->ev.push_back(e);
    }
    *a = h->ev[2 * k];
    *b = h->ev[2 * k + 1];
    return 0;
}

template <int P, int NG>
static int launch_elem(yeti_asm *h, PatchDev &pd, const double *coords, double *Kd, int32_t *Kr, int32_t *Kc,
                       hipStream_t st) {
    using C = ElemCfg<P, NG>;
    if (pd.n_active == 0) return 0;
    const int32_t batch = std::min(pd.n_active, ELEM_BATCH);
    const size_t geo_n = (size_t)batch * C::REC;
    if (pd.geo.n < geo_n) YETI_HIP(pd.geo.alloc(geo_n));
    PatchArgs pa;
    pa.coords = coords;
    pa.ien = pd.ien.p;
    pa.weight = pd.weight.p;
    pa.nijk = pd.nijk.p;
    pa.U0 = pd.U[0].p;
    pa.U1 = pd.U[1].p;
    pa.U2 = pd.U[2].p;
    pa.lam = pd.lam;
    pa.mu = pd.mu;
    gauss_legendre(NG, pa.g.x, pa.g.w);
    const int cus = device_cu_count();
    for (int32_t e0 = 0; e0 < pd.n_active; e0 += batch) {
        pa.elems = pd.elems.p + e0;
        pa.offs = pd.offs.p + e0;
        pa.n_active = std::min(batch, pd.n_active - e0);
        constexpr int GEO_EPB = 4 * geo_epw<C::NQ>();
        hipLaunchKernelGGL((elem_geo_kernel<P, NG>), dim3((unsigned)((pa.n_active + GEO_EPB - 1) / GEO_EPB)), dim3(256), 0,
                           st, pa, pd.geo.p);
        // persistent workgroups: one per CU (the kernel's LDS and registers allow one)
        const int grid = std::min(pa.n_active, cus);
        hipEvent_t ea = nullptr, eb = nullptr;
        if (h->timing) {
            int rc = timing_events(h, h->ev_used++, &ea, &eb);
            if (rc) return rc;
            YETI_HIP(hipEventRecord(ea, st));
        }
        if (Kr || C::VALUES_VIA_IDX)
            hipLaunchKernelGGL((elem_stiffness_kernel<P, NG, true>), dim3(grid), dim3(C::NTH), 0, st, pa,
                               (const double *)pd.geo.p, Kd, Kr, Kc);
        else
            hipLaunchKernelGGL((elem_stiffness_kernel<P, NG, false>), dim3(grid), dim3(C::NTH), 0, st, pa,
                               (const double *)pd.geo.p, Kd, Kr, Kc);
        YETI_HIP(hipGetLastError());
        if (h->timing) YETI_HIP(hipEventRecord(eb, st));
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// Element stiffness for the degree triples and Gauss orders the MFMA kernel is not built for
// (unequal degrees per direction, NBINT other than (p + 1)^3), up to GEN_MAXN nodes and degree
// GEN_MAXP per direction: the same
// K_ab[k][l] = lam S^kl_ab + mu S^lk_ab + mu delta_kl tr S_ab with S^kl_ab = sum_q w_q |det J_q|
// dR_a/dx_k dR_b/dx_l (UELMAT.f90:22-225 with stiffmatrix_byCP, stiffmatrix.F90:130-225), in the
// same COO layout. One 256-thread workgroup per element, Gauss points in chunks of GEN_QC:
// bases one thread per (point, direction), W / dW / dx/du of a point by 16 lanes and a shuffle
// sum, dR/dx of every (node, point) in LDS, then each thread adds the chunk to its node pairs
// (pair t, t + 256, ... in COO order: b (b + 1) / 2 + a, a <= b) in registers and finally writes
// their blocks.
constexpr int GEN_MAXN = 64, GEN_MAXP = 4, GEN_QC = 16, GEN_PPT = (GEN_MAXN * (GEN_MAXN + 1) / 2 + 255) / 256;

struct GenPatch {
    PatchArgs pa;
    int p[3];
    int ng;
    double rho;
};

// MODE 0: stiffness (|det J|); MODE 1: consistent mass rho sum_q w_q det J_q R_a R_b on the
// diagonal entries, det J signed (compute_CMASSMatrix.f:22-215, UMASSMAT.f:220-330)
// PPT: node pairs per thread (the launch picks the smallest of 1, 2, 3, 5, 9 that covers the
// element's pairs, so small elements keep few accumulators live)
template <int MODE, int PPT>
__global__ void __launch_bounds__(256) elem_generic_kernel(GenPatch gp, double *__restrict__ Kd,
                                                           int32_t *__restrict__ Kr, int32_t *__restrict__ Kc) {
    const PatchArgs &pa = gp.pa;
    const int n0 = gp.p[0] + 1, n1 = gp.p[1] + 1, n2 = gp.p[2] + 1, NN = n0 * n1 * n2;
    const int NG = gp.ng, NQ = NG * NG * NG, NPAIR = NN * (NN + 1) / 2;
    constexpr int G = 256 / GEN_QC;
    __shared__ double xw[GEN_MAXN][4];
    __shared__ int32_t dof[GEN_MAXN];
    __shared__ double Bq[GEN_QC][3][GEN_MAXP + 1], dBq[GEN_QC][3][GEN_MAXP + 1];
    __shared__ double gq[GEN_QC][14];        // du/dx (9), dv = w |det J|, 1/W, dW/du / W (3)
    __shared__ double Dx[GEN_MAXN][GEN_QC][MODE == 0 ? 3 : 1];
    __shared__ double hel[3];
    const int t = threadIdx.x, qq = t / G, part = t % G;
    for (int64_t gi = blockIdx.x; gi < pa.n_active; gi += gridDim.x) {
        const int e = pa.elems[gi];
        __syncthreads();
        for (int a = t; a < NN; a += 256) {
            const int32_t cp = pa.ien[(int64_t)NN * e + a] - 1;
            xw[a][0] = pa.coords[3 * (int64_t)cp];
            xw[a][1] = pa.coords[3 * (int64_t)cp + 1];
            xw[a][2] = pa.coords[3 * (int64_t)cp + 2];
            xw[a][3] = pa.weight[(int64_t)NN * e + a];
            dof[a] = 3 * cp;
        }
        double S[PPT][9];
#pragma unroll
        for (int k = 0; k < PPT; ++k)
#pragma unroll
            for (int m = 0; m < 9; ++m) S[k][m] = 0.0;
        for (int q0 = 0; q0 < NQ; q0 += GEN_QC) {
            const int nq = NQ - q0 < GEN_QC ? NQ - q0 : GEN_QC;
            __syncthreads();
            // 1D bases of direction `part` at point qq (the reference's 3D rule: i slowest,
            // k fastest, GaussLegendre.f90:175-186), local index reversed (shap.f90:94-119)
            if (part < 3 && qq < nq) {
                const int q = q0 + qq, d = part, pd = gp.p[d];
                const int gd = d == 0 ? q / (NG * NG) : d == 1 ? (q / NG) % NG : q % NG;
                const double *U = d == 0 ? pa.U0 : d == 1 ? pa.U1 : pa.U2;
                const int span1 = pa.nijk[3 * (int64_t)e + d], i0 = span1 - 1;
                const double u0 = U[span1 - 1], u1 = U[span1];
                const double u = ((u1 - u0) * pa.g.x[gd] + (u1 + u0)) * 0.5;
                double lf[GEN_MAXP + 1], rt[GEN_MAXP + 1], N[GEN_MAXP + 1], lo[GEN_MAXP + 1];
                N[0] = 1.0;
                lo[0] = 1.0;
                for (int j = 1; j <= pd; ++j) {
                    lf[j] = u - U[i0 + 1 - j];
                    rt[j] = U[i0 + j] - u;
                    double saved = 0.0;
                    for (int r = 0; r < j; ++r) {
                        const double tmp = N[r] / (rt[r + 1] + lf[j - r]);
                        N[r] = saved + rt[r + 1] * tmp;
                        saved = lf[j - r] * tmp;
                    }
                    N[j] = saved;
                    if (j == pd - 1)
                        for (int r = 0; r < pd; ++r) lo[r] = N[r];
                }
                for (int r = 0; r <= pd; ++r) {
                    const int k = i0 - pd + r;
                    double dd = 0.0;
                    if (r >= 1 && U[k + pd] != U[k]) dd += pd * lo[r - 1] / (U[k + pd] - U[k]);
                    if (r <= pd - 1 && U[k + pd + 1] != U[k + 1]) dd -= pd * lo[r] / (U[k + pd + 1] - U[k + 1]);
                    Bq[qq][d][pd - r] = N[r];
                    dBq[qq][d][pd - r] = dd;
                }
                if (qq == 0) hel[d] = 0.5 * (u1 - u0);
            }
            __syncthreads();
            // W, dW/du and dx/du of point qq: nodes part, part + G, ..., then a shuffle sum
            {
                double s[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) s[k] = 0.0;
                if (qq < nq) {
                    for (int a = part; a < NN; a += G) {
                        const int i = a % n0, j = (a / n0) % n1, k = a / (n0 * n1);
                        const double w = xw[a][3], x0 = xw[a][0], x1 = xw[a][1], x2 = xw[a][2];
                        const double B0 = Bq[qq][0][i], B1 = Bq[qq][1][j], B2 = Bq[qq][2][k];
                        const double Na = B0 * B1 * B2 * w;
                        const double d0 = dBq[qq][0][i] * B1 * B2 * w;
                        const double d1 = B0 * dBq[qq][1][j] * B2 * w;
                        const double d2 = B0 * B1 * dBq[qq][2][k] * w;
                        s[0] += Na;
                        s[1] += d0; s[2] += d1; s[3] += d2;
                        s[4] += x0 * Na; s[5] += x1 * Na; s[6] += x2 * Na;
                        s[7] += x0 * d0; s[8] += x0 * d1; s[9] += x0 * d2;
                        s[10] += x1 * d0; s[11] += x1 * d1; s[12] += x1 * d2;
                        s[13] += x2 * d0; s[14] += x2 * d1; s[15] += x2 * d2;
                    }
                }
#pragma unroll
                for (int o = 1; o < G; o <<= 1)
#pragma unroll
                    for (int k = 0; k < 16; ++k) s[k] += __shfl_xor(s[k], o, 64);
                if (part == 0) {
                    double *g = gq[qq];
                    if (qq < nq) {
                        const int q = q0 + qq;
                        const double Wi = 1.0 / s[0];
                        const double dw[3] = {s[1] * Wi, s[2] * Wi, s[3] * Wi};
                        double J[3][3];
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 3; ++c) J[r][c] = (s[7 + 3 * r + c] - s[4 + r] * dw[c]) * Wi;
                        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                                           J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                        const double di = 1.0 / det;
                        g[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * di;
                        g[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * di;
                        g[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * di;
                        g[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * di;
                        g[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * di;
                        g[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * di;
                        g[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * di;
                        g[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * di;
                        g[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * di;
                        const int g0 = q / (NG * NG), g1 = (q / NG) % NG, g2 = q % NG;
                        const double dj = det * hel[0] * hel[1] * hel[2];
                        g[9] = MODE == 0 ? (pa.g.w[g0] * pa.g.w[g1] * pa.g.w[g2]) * fabs(dj)
                                         : gp.rho * (pa.g.w[g0] * pa.g.w[g1] * pa.g.w[g2] * dj);
                        g[10] = Wi;
                        g[11] = dw[0];
                        g[12] = dw[1];
                        g[13] = dw[2];
                    } else {
                        for (int k = 0; k < 14; ++k) g[k] = 0.0;
                    }
                }
            }
            __syncthreads();
            // stiffness: dR/dx = ((dN w) / W - R dW / W) du/dx; mass: R, of every (node, point)
            for (int x = t; x < NN * GEN_QC; x += 256) {
                const int a = x / GEN_QC, c = x % GEN_QC;
                double r3[3] = {0.0, 0.0, 0.0};
                if (c < nq) {
                    const int i = a % n0, j = (a / n0) % n1, k = a / (n0 * n1);
                    const double w = xw[a][3], Wi = gq[c][10];
                    const double B0 = Bq[c][0][i], B1 = Bq[c][1][j], B2 = Bq[c][2][k];
                    const double R = B0 * B1 * B2 * w * Wi;
                    if (MODE == 0) {
                        const double du[3] = {dBq[c][0][i] * B1 * B2 * w * Wi - R * gq[c][11],
                                              B0 * dBq[c][1][j] * B2 * w * Wi - R * gq[c][12],
                                              B0 * B1 * dBq[c][2][k] * w * Wi - R * gq[c][13]};
                        for (int m = 0; m < 3; ++m)
                            r3[m] = du[0] * gq[c][m] + du[1] * gq[c][3 + m] + du[2] * gq[c][6 + m];
                    } else {
                        r3[0] = R;
                    }
                }
                for (int m = 0; m < (MODE == 0 ? 3 : 1); ++m) Dx[a][c][m] = r3[m];
            }
            __syncthreads();
            // pair sums S^kl_ab += dv dR_a/dx_k dR_b/dx_l
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const int pi = t + 256 * k;
                if (pi < NPAIR) {
                    int b = (int)((sqrt(8.0 * pi + 1.0) - 1.0) * 0.5);   // pi = b (b + 1) / 2 + a
                    while (b * (b + 1) / 2 > pi) --b;
                    while ((b + 1) * (b + 2) / 2 <= pi) ++b;
                    const int a = pi - b * (b + 1) / 2;
                    if (MODE == 1) {
                        for (int c = 0; c < nq; ++c) S[k][0] += gq[c][9] * Dx[b][c][0] * Dx[a][c][0];
                        continue;
                    }
                    for (int c = 0; c < nq; ++c) {
                        const double dv = gq[c][9];
                        const double ga0 = dv * Dx[a][c][0], ga1 = dv * Dx[a][c][MODE == 0 ? 1 : 0],
                                     ga2 = dv * Dx[a][c][MODE == 0 ? 2 : 0];
                        const double gb0 = Dx[b][c][0], gb1 = Dx[b][c][MODE == 0 ? 1 : 0],
                                     gb2 = Dx[b][c][MODE == 0 ? 2 : 0];
                        S[k][0] += ga0 * gb0; S[k][1] += ga0 * gb1; S[k][2] += ga0 * gb2;
                        S[k][3] += ga1 * gb0; S[k][4] += ga1 * gb1; S[k][5] += ga1 * gb2;
                        S[k][6] += ga2 * gb0; S[k][7] += ga2 * gb1; S[k][8] += ga2 * gb2;
                    }
                }
            }
        }
        // K_ab[k][l] = lam S^kl + mu S^lk + mu delta_kl tr S, written at the element's COO offset
        const int64_t base = pa.offs[gi];
        const double lam = pa.lam, mu = pa.mu;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int pi = t + 256 * k;
            if (pi >= NPAIR) continue;
            int b = (int)((sqrt(8.0 * pi + 1.0) - 1.0) * 0.5);
            while (b * (b + 1) / 2 > pi) --b;
            while ((b + 1) * (b + 2) / 2 <= pi) ++b;
            const int a = pi - b * (b + 1) / 2;
            const double tr = mu * (S[k][0] + S[k][4] + S[k][8]);
            double K[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    K[r][c] = MODE == 0 ? lam * S[k][3 * r + c] + mu * S[k][3 * c + r] + (r == c ? tr : 0.0)
                                        : (r == c ? S[k][0] : 0.0);
            const int64_t o = base + coo_start(b) + (a < b ? 9 * a : 9 * b);
            if (a < b) {
                for (int x = 0; x < 9; ++x) {
                    const int kk = x % 3, ll = x / 3;
                    Kd[o + x] = K[kk][ll];
                    if (Kr) {
                        Kr[o + x] = dof[a] + kk;
                        Kc[o + x] = dof[b] + ll;
                    }
                }
            } else {
                for (int x = 0; x < 6; ++x) {   // (0,0) (0,1) (1,1) (0,2) (1,2) (2,2)
                    const int kk = (0x210100 >> (4 * x)) & 15, ll = (0x222110 >> (4 * x)) & 15;
                    Kd[o + x] = kk == ll ? 0.5 * K[kk][ll] : K[kk][ll];
                    if (Kr) {
                        Kr[o + x] = dof[b] + kk;
                        Kc[o + x] = dof[b] + ll;
                    }
                }
            }
        }
    }
}

static int launch_generic(yeti_asm *h, PatchDev &pd, double *Kd, int32_t *Kr, int32_t *Kc, hipStream_t st,
                          bool mass = false) {
    GenPatch gp{};
    gp.pa.coords = h->coords.p;
    gp.pa.ien = pd.ien.p;
    gp.pa.weight = pd.weight.p;
    gp.pa.nijk = pd.nijk.p;
    gp.pa.U0 = pd.U[0].p;
    gp.pa.U1 = pd.U[1].p;
    gp.pa.U2 = pd.U[2].p;
    gp.pa.elems = pd.elems.p;
    gp.pa.offs = pd.offs.p;
    gp.pa.n_active = pd.n_active;
    gp.pa.lam = pd.lam;
    gp.pa.mu = pd.mu;
    gauss_legendre(pd.ng, gp.pa.g.x, gp.pa.g.w);
    for (int d = 0; d < 3; ++d) gp.p[d] = pd.p[d];
    gp.ng = pd.ng;
    gp.rho = pd.rho;
    const unsigned grid = (unsigned)std::min<int64_t>(pd.n_active, 1 << 20);
    const int nn = (pd.p[0] + 1) * (pd.p[1] + 1) * (pd.p[2] + 1), need = (nn * (nn + 1) / 2 + 255) / 256;
    if (need > GEN_PPT) return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "general element kernel: %d nodes", nn);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, gp, Kd, Kr, Kc); };
    if (mass) {
        if (need <= 1) go(elem_generic_kernel<1, 1>);
        else if (need <= 2) go(elem_generic_kernel<1, 2>);
        else if (need <= 3) go(elem_generic_kernel<1, 3>);
        else if (need <= 5) go(elem_generic_kernel<1, 5>);
        else go(elem_generic_kernel<1, GEN_PPT>);
    } else {
        if (need <= 1) go(elem_generic_kernel<0, 1>);
        else if (need <= 2) go(elem_generic_kernel<0, 2>);
        else if (need <= 3) go(elem_generic_kernel<0, 3>);
        else if (need <= 5) go(elem_generic_kernel<0, 5>);
        else go(elem_generic_kernel<0, GEN_PPT>);
    }
    YETI_HIP(hipGetLastError());
    return 0;
}

static int run_elements(yeti_asm *h, double *Kd, int32_t *Kr, int32_t *Kc, hipStream_t st) {
    h->ev_used = 0;
    for (auto &zr : h->zero_ranges) {
        int64_t n = zr.second;
        int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, st, Kd + zr.first,
                           Kr ? Kr + zr.first : nullptr, Kc ? Kc + zr.first : nullptr, n);
        YETI_HIP(hipGetLastError());
    }
    for (auto *pd : h->patches) {
        if (!pd->u1 || pd->n_active == 0) continue;
        int rc = 0;
        if (!mfma_kernel_available(pd->p[0], pd->p[1], pd->p[2], pd->ng)) {
            rc = launch_generic(h, *pd, Kd, Kr, Kc, st);
            if (rc) return rc;
            continue;
        }
        int key = pd->p[0] * 16 + pd->ng;
        switch (key) {
        case 1 * 16 + 2: rc = launch_elem<1, 2>(h, *pd, h->coords.p, Kd, Kr, Kc, st); break;
        case 2 * 16 + 3: rc = launch_elem<2, 3>(h, *pd, h->coords.p, Kd, Kr, Kc, st); break;
        case 3 * 16 + 4: rc = launch_elem<3, 4>(h, *pd, h->coords.p, Kd, Kr, Kc, st); break;
        case 4 * 16 + 5: rc = launch_elem<4, 5>(h, *pd, h->coords.p, Kd, Kr, Kc, st); break;
        default: return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "no kernel for p=%d ng=%d", pd->p[0], pd->ng);
        }
        if (rc) return rc;
    }
    return 0;
}

// ---------------------------------------------------------------------------------------
// Consistent mass matrix (compute_CMASSMatrix.f:22-215 with UMASSMAT_byCP, UMASSMAT.f:220-330,
// solid branch): M_ab[k][k] = sum_q rho w_q DetJac_q R_a(q) R_b(q) (DetJac signed, parent ->
// physical), off-diagonal k != l entries 0, emitted in the stiffness COO layout (a < b: 3 x 3
// block, kk fastest; a = b: the upper 6 with the diagonal halved) at the same element offsets.
// One 256-thread workgroup per element (grid-stride over the active elements), Gauss points in
// chunks of MQC:
//   1. the 1D bases of the chunk's points, one thread per (point, direction);
//   2. W, dW and the Jacobian of each point from G = 256 / MQC threads that split its nodes and
//      add their sums with lane shuffles;
//   3. R_a(q) = N_a w_a / W for the chunk in LDS (node-major, points contiguous);
//   4. the node-pair sums as 16 x 16 tiles (a tile, b tile) of (dv R)^T R on
//      v_mfma_f64_16x16x4f64, upper tiles spread over the four waves, accumulated over chunks;
// then the upper triangle of M goes to LDS (packed by column, b (b + 1) / 2 + a) and all threads
// write the element's COO range front to back, so every store of a wave is one contiguous run.
template <int P, int NG>
struct MassCfg {
    static constexpr int NP1 = P + 1, NN = NP1 * NP1 * NP1, NQ = NG * NG * NG;
    static constexpr int NNP = (NN + 15) / 16 * 16, NT = NNP / 16, NTILE = NT * (NT + 1) / 2;
    static constexpr int MAXT = (NTILE + 3) / 4;                       // tiles per wave
    // points per chunk: at most 32, so that at p = 3 the LDS (26 KB) and registers (compiled for
    // 4 waves per SIMD) allow 4 workgroups per CU (64-point chunks: 49 KB, 3 per CU, 16.6 ms
    // against 15.5 ms at 64^3)
    static constexpr int MQC = NQ <= 32 ? (NQ + 7) / 8 * 8 : 32;
    static constexpr int G = 256 / MQC;                                // threads per point
    static constexpr int RS = MQC + 2;                                 // R row stride (doubles)
    static constexpr int NPAIR = NN * (NN + 1) / 2;
    // LDS (doubles): xw, dvq, then R + bases, which the packed M reuses after the last chunk
    static constexpr int O_DV = 4 * NN, O_U = O_DV + MQC;
    static constexpr int O_B = O_U + NNP * RS, O_DB = O_B + MQC * 3 * NP1;
    static constexpr int N_U = NNP * RS + 2 * MQC * 3 * NP1;
    static constexpr int NDBL = O_U + (N_U > NPAIR ? N_U : NPAIR);
    static constexpr int64_t NCOO = coo_start(NN);
    static_assert(256 % MQC == 0 && MQC % 8 == 0 && G <= 64, "mass chunking");
};

template <int P, int NG, bool WRITE_IDX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(P == 3 ? 4 : 1)))
mass_kernel(PatchArgs pa, double rho, double *__restrict__ Md,
                                                   int32_t *__restrict__ Mr, int32_t *__restrict__ Mc) {
    using C = MassCfg<P, NG>;
    constexpr int NP1 = C::NP1, NN = C::NN, NQ = C::NQ, MQC = C::MQC, G = C::G, RS = C::RS;
    __shared__ __attribute__((aligned(16))) double sm[C::NDBL];
    __shared__ double hel[3], wiq[MQC];
    __shared__ int32_t dof[NN];
    double(*xw)[4] = reinterpret_cast<double(*)[4]>(sm);
    double *dvq = sm + C::O_DV;
    double *Rq = sm + C::O_U;                                              // [NNP][RS]
    double(*Bq)[3][NP1] = reinterpret_cast<double(*)[3][NP1]>(sm + C::O_B);
    double(*dBq)[3][NP1] = reinterpret_cast<double(*)[3][NP1]>(sm + C::O_DB);
    double *mp = sm + C::O_U;                                              // packed upper M
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int nl = lane & 15, kg = lane >> 4;
    const int qq = t / G, part = t % G;
    // this wave's tiles u = wv + 4 s, numbered column by column: u = tb (tb + 1) / 2 + ta
    int ta[C::MAXT], tb[C::MAXT];
#pragma unroll
    for (int s = 0; s < C::MAXT; ++s) {
        const int u = wv + 4 * s;
        int j = 0;
        while ((j + 1) * (j + 2) / 2 <= u) ++j;
        tb[s] = j;
        ta[s] = u - j * (j + 1) / 2;
    }
    for (int64_t gi = blockIdx.x; gi < pa.n_active; gi += gridDim.x) {
        const int e = pa.elems[gi];
        __syncthreads();   // the previous element is done with the LDS arrays
        for (int a = t; a < NN; a += 256) {
            const int32_t cp = pa.ien[(int64_t)NN * e + a] - 1;
            xw[a][0] = pa.coords[3 * (int64_t)cp];
            xw[a][1] = pa.coords[3 * (int64_t)cp + 1];
            xw[a][2] = pa.coords[3 * (int64_t)cp + 2];
            xw[a][3] = pa.weight[(int64_t)NN * e + a];
            dof[a] = 3 * cp;
        }
        f64x4 acc[C::MAXT];
#pragma unroll
        for (int s = 0; s < C::MAXT; ++s) acc[s] = f64x4{0.0, 0.0, 0.0, 0.0};
        for (int q0 = 0; q0 < NQ; q0 += MQC) {
            const int nq = NQ - q0 < MQC ? NQ - q0 : MQC;
            __syncthreads();
            // 1. bases: direction `part` of point qq (shap.f90:75-178); the reference's 3D rule
            // runs i (xi) slowest and k (zeta) fastest (GaussLegendre.f90:175-186)
            if (part < 3 && qq < nq) {
                const int q = q0 + qq, d = part;
                const int gd = d == 0 ? q / (NG * NG) : d == 1 ? (q / NG) % NG : q % NG;
                const double *U = d == 0 ? pa.U0 : d == 1 ? pa.U1 : pa.U2;
                const int span1 = pa.nijk[3 * (int64_t)e + d];
                const double u0 = U[span1 - 1], u1 = U[span1];
                const double u = ((u1 - u0) * pa.g.x[gd] + (u1 + u0)) * 0.5;
                double N[NP1], dN[NP1];
                bspline_ders<P>(span1 - 1, u, U, N, dN);
#pragma unroll
                for (int i = 0; i < NP1; ++i) {
                    Bq[qq][d][i] = N[P - i];
                    dBq[qq][d][i] = dN[P - i];
                }
                if (qq == 0) hel[d] = 0.5 * (u1 - u0);
            }
            __syncthreads();
            // 2. W, dW/dxi and dX/dxi of point qq: nodes part, part + G, ... then a shuffle sum
            {
                double s[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) s[k] = 0.0;
                if (qq < nq) {
#pragma unroll 1
                    for (int a = part; a < NN; a += G) {
                        const int i = a % NP1, j = (a / NP1) % NP1, k = a / (NP1 * NP1);
                        const double w = xw[a][3], x0 = xw[a][0], x1 = xw[a][1], x2 = xw[a][2];
                        const double B0 = Bq[qq][0][i], B1 = Bq[qq][1][j], B2 = Bq[qq][2][k];
                        const double b12 = B1 * B2 * w;
                        const double Na = B0 * b12;
                        const double d0 = dBq[qq][0][i] * b12;
                        const double d1 = B0 * dBq[qq][1][j] * B2 * w;
                        const double d2 = B0 * B1 * dBq[qq][2][k] * w;
                        s[0] += Na;
                        s[1] += d0; s[2] += d1; s[3] += d2;
                        s[4] += x0 * Na; s[5] += x1 * Na; s[6] += x2 * Na;
                        s[7] += x0 * d0; s[8] += x0 * d1; s[9] += x0 * d2;
                        s[10] += x1 * d0; s[11] += x1 * d1; s[12] += x1 * d2;
                        s[13] += x2 * d0; s[14] += x2 * d1; s[15] += x2 * d2;
                    }
                }
#pragma unroll
                for (int o = 1; o < G; o <<= 1)   // the G threads of a point are adjacent lanes
#pragma unroll
                    for (int k = 0; k < 16; ++k) s[k] += __shfl_xor(s[k], o, 64);
                if (part == 0) {
                    double dv = 0.0, Wi = 0.0;
                    if (qq < nq) {
                        const int q = q0 + qq;
                        Wi = 1.0 / s[0];
                        const double dw[3] = {s[1] * Wi, s[2] * Wi, s[3] * Wi};
                        double J[3][3];
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int c = 0; c < 3; ++c) J[r][c] = (s[7 + 3 * r + c] - s[4 + r] * dw[c]) * Wi;
                        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                                           J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                        const int g0 = q / (NG * NG), g1 = (q / NG) % NG, g2 = q % NG;
                        dv = rho * (pa.g.w[g0] * pa.g.w[g1] * pa.g.w[g2]) * (det * hel[0] * hel[1] * hel[2]);
                    }
                    dvq[qq] = dv;   // 0 at the padding points of the last chunk
                    wiq[qq] = Wi;
                }
            }
            __syncthreads();
            // 3. R_a(q) = N_a w_a / W; zero rows past NN and columns past nq
            for (int x = t; x < C::NNP * MQC; x += 256) {
                const int a = x / MQC, c = x % MQC;
                double r = 0.0;
                if (a < NN && c < nq) {
                    const int i = a % NP1, j = (a / NP1) % NP1, k = a / (NP1 * NP1);
                    r = Bq[c][0][i] * Bq[c][1][j] * Bq[c][2][k] * xw[a][3] * wiq[c];
                }
                Rq[a * RS + c] = r;
            }
            __syncthreads();
            // 4. tile sums; K step 2v uses point 8v + 2 kg, step 2v + 1 point 8v + 2 kg + 1 (A and
            // B with the same permutation of the K sum)
#pragma unroll 1
            for (int v = 0; v < MQC / 8; ++v) {
                const int qo = 8 * v + 2 * kg;
                const f64x2 dv2 = *reinterpret_cast<const f64x2 *>(dvq + qo);
#pragma unroll
                for (int s = 0; s < C::MAXT; ++s) {
                    if (wv + 4 * s < C::NTILE) {
                        f64x2 A = *reinterpret_cast<const f64x2 *>(Rq + (16 * ta[s] + nl) * RS + qo);
                        const f64x2 B = *reinterpret_cast<const f64x2 *>(Rq + (16 * tb[s] + nl) * RS + qo);
                        A.x *= dv2.x;
                        A.y *= dv2.y;
                        acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(A.x, B.x, acc[s], 0, 0, 0);
                        acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(A.y, B.y, acc[s], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();   // every wave is done with R before M overwrites it
        // accumulator lane layout: register r of lane l holds row (l >> 4) + 4 r, column l & 15
#pragma unroll
        for (int s = 0; s < C::MAXT; ++s) {
            if (wv + 4 * s < C::NTILE) {
                const int b = 16 * tb[s] + nl;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int a = 16 * ta[s] + kg + 4 * r;
                    if (a <= b && b < NN) mp[b * (b + 1) / 2 + a] = acc[s][r];
                }
            }
        }
        __syncthreads();
        // emission: entry x of the element's range lies in column b with start(b) <= x < start(b+1)
        const int64_t base = pa.offs[gi];
        int b = 0;
        for (int x = t; x < (int)C::NCOO; x += 256) {
            while (b + 1 < NN && (int)coo_start(b + 1) <= x) ++b;
            const unsigned r = (unsigned)(x - (int)coo_start(b));
            const int mb = b * (b + 1) / 2;
            double v;
            int32_t row, col;
            if (r < 9u * (unsigned)b) {
                const unsigned a = r / 9u, y = r - 9u * a, kk = y % 3u, ll = y / 3u;
                v = kk == ll ? mp[mb + (int)a] : 0.0;
                row = dof[a] + (int)kk;
                col = dof[b] + (int)ll;
            } else {
                const unsigned y = r - 9u * (unsigned)b;   // (0,0) (0,1) (1,1) (0,2) (1,2) (2,2)
                const int kk = (0x210100 >> (4 * y)) & 15, ll = (0x222110 >> (4 * y)) & 15;
                v = kk == ll ? 0.5 * mp[mb + b] : 0.0;
                row = dof[b] + kk;
                col = dof[b] + ll;
            }
            // streaming output, read back by no kernel of this launch
            __builtin_nontemporal_store(v, Md + base + x);
            if (WRITE_IDX) {
                __builtin_nontemporal_store(row, Mr + base + x);
                __builtin_nontemporal_store(col, Mc + base + x);
            }
        }
    }
}

template <int P, int NG>
static int launch_mass(yeti_asm *h, PatchDev &pd, double *Md, int32_t *Mr, int32_t *Mc, hipStream_t st) {
    PatchArgs pa{};
    pa.coords = h->coords.p;
    pa.ien = pd.ien.p;
    pa.weight = pd.weight.p;
    pa.nijk = pd.nijk.p;
    pa.U0 = pd.U[0].p;
    pa.U1 = pd.U[1].p;
    pa.U2 = pd.U[2].p;
    pa.elems = pd.elems.p;
    pa.offs = pd.offs.p;
    pa.n_active = pd.n_active;
    gauss_legendre(NG, pa.g.x, pa.g.w);
    const unsigned grid = (unsigned)std::min<int64_t>(pd.n_active, 1 << 20);
    if (Mr)
        hipLaunchKernelGGL((mass_kernel<P, NG, true>), dim3(grid), dim3(256), 0, st, pa, pd.rho, Md, Mr, Mc);
    else
        hipLaunchKernelGGL((mass_kernel<P, NG, false>), dim3(grid), dim3(256), 0, st, pa, pd.rho, Md, Mr, Mc);
    YETI_HIP(hipGetLastError());
    return 0;
}

static int run_mass(yeti_asm *h, double *Md, int32_t *Mr, int32_t *Mc, hipStream_t st) {
    for (auto &zr : h->zero_ranges) {
        int64_t n = zr.second;
        int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, st, Md + zr.first,
                           Mr ? Mr + zr.first : nullptr, Mc ? Mc + zr.first : nullptr, n);
        YETI_HIP(hipGetLastError());
    }
    for (auto *pd : h->patches) {
        if (!pd->u1 || pd->n_active == 0) continue;
        int rc = 0;
        if (!mfma_kernel_available(pd->p[0], pd->p[1], pd->p[2], pd->ng)) {   // other degree triples / orders
            rc = launch_generic(h, *pd, Md, Mr, Mc, st, true);
            if (rc) return rc;
            continue;
        }
        switch (pd->p[0] * 16 + pd->ng) {
        case 1 * 16 + 2: rc = launch_mass<1, 2>(h, *pd, Md, Mr, Mc, st); break;
        case 2 * 16 + 3: rc = launch_mass<2, 3>(h, *pd, Md, Mr, Mc, st); break;
        case 3 * 16 + 4: rc = launch_mass<3, 4>(h, *pd, Md, Mr, Mc, st); break;
        case 4 * 16 + 5: rc = launch_mass<4, 5>(h, *pd, Md, Mr, Mc, st); break;
        default: return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "mass: no kernel for p=%d ng=%d", pd->p[0], pd->ng);
        }
        if (rc) return rc;
    }
    return 0;
}

// B-spline values and first derivatives at u for a run-time degree p <= PMAX, span i (0-based):
// the triangular recurrence with the degree p - 1 values kept for the derivatives (the contract
// of dersbasisfuns.f90:30-131, first derivative only)
template <int PMAX>
__device__ inline void bspline_ders_rt(int i, int p, double u, const double *__restrict__ U, double *N, double *dN) {
    double lf[PMAX + 1], rt[PMAX + 1], lo[PMAX + 1];
    N[0] = 1.0;
    lo[0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        lf[j] = u - U[i + 1 - j];
        rt[j] = U[i + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            const double tmp = N[r] / (rt[r + 1] + lf[j - r]);
            N[r] = saved + rt[r + 1] * tmp;
            saved = lf[j - r] * tmp;
        }
        N[j] = saved;
        if (j == p - 1)
            for (int r = 0; r < p; ++r) lo[r] = N[r];
    }
    for (int r = 0; r <= p; ++r) {
        const int k = i - p + r;
        double dd = 0.0;
        if (r >= 1 && U[k + p] != U[k]) dd += p * lo[r - 1] / (U[k + p] - U[k]);
        if (r <= p - 1 && U[k + p + 1] != U[k + 1]) dd -= p * lo[r] / (U[k + p + 1] - U[k + 1]);
        dN[r] = dd;
    }
}

// ---------------------------------------------------------------------------------------
// Natural-frequency shape gradient (gradvibration_fullAN.f:22-189 with gradUELMASSMAT1,
// gradUELMASSMAT_solid.f:22-332), solid patches, up to VIB_MAXF modes per launch. For mode f with
// eigenvalue lambda_f and eigenvector u, element e adds to its node a
//   g_a[m] = sum_q dv_q [ (eps:sig - lambda_f rho |u|^2) dR_a/dx_m - 2 sum_ik sig_ik du_i/dx_m dR_a/dx_k ],
// the derivative of u^T (K - lambda_f M) u with respect to control point a (the reference writes
// the same derivative in the covariant basis as its Jacobian, strain and metric terms), and the
// mass rho |u|^2 dv; the CP gradient is the sum over elements divided by the total mass per mode.
// One 256-thread workgroup per active element, Gauss points in chunks of VQC:
//   1. bases, one thread per (point, direction);
//   2. W, dW and dx/du of a point by 256 / VQC threads and a shuffle sum, then du/dx and dv;
//   3. R and dR/dx of every (node, point) of the chunk in LDS;
//   4. per (point, mode): u, du/dx, sig, the tensor dv (W' I - 2 (du/dx)^T sig) and the mass;
//   5. per (node, mode): g += tensor . dR/dx, accumulated over the chunks in the registers of the
//      (node, mode) pair's thread.
constexpr int VIB_MAXF = 4, VQC = 16;

struct VibModes {
    const double *vect;            // (nb_frq, 3, nb_cp) column-major
    int32_t nf_total, f0, nf;      // modes f0 .. f0 + nf - 1 of nf_total
    double vals[VIB_MAXF];
    double scale;                  // the CP sums are multiplied by scale ...
    int32_t normalize;             // ... and divided by each mode's mass when set
    int32_t unit_density;          // density 1 in place of the patches' (volume)
};

// RT = false: degree P in every direction and NG points (compile-time); RT = true: the degrees and
// Gauss order of `dg` at run time, P and GEN_MAXN bounding the arrays.
struct VibDegrees {
    int p[3];
    int ng;
};

// waves per SIMD the kernel is compiled for: at p = 3 three (its LDS allows three workgroups per
// CU once the node sums are in registers), at p = 2 four, elsewhere what the registers give
template <int P, bool RT>
constexpr int vib_wpe() { return RT ? 1 : P == 3 ? 3 : P == 2 ? 4 : 1; }

template <int P, int NG, bool RT = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(vib_wpe<P, RT>())))
vib_grad_kernel(PatchArgs pa, double rho, VibModes vm, int64_t ebase,
                                                       int64_t mbase, double *__restrict__ eg,
                                                       double *__restrict__ em, VibDegrees dg = VibDegrees{}) {
    constexpr int NP1 = P + 1, NN = RT ? GEN_MAXN : NP1 * NP1 * NP1, G = 256 / VQC;
    const int p0 = RT ? dg.p[0] : P, p1 = RT ? dg.p[1] : P, p2 = RT ? dg.p[2] : P;
    const int n0 = p0 + 1, n1 = p1 + 1, nn = RT ? n0 * n1 * (p2 + 1) : NN;
    const int ng = RT ? dg.ng : NG, NQ = ng * ng * ng;
    __shared__ double xw[NN][4];
    __shared__ double uf[VIB_MAXF][NN][3];
    __shared__ double Bq[VQC][3][NP1], dBq[VQC][3][NP1];
    __shared__ double gq[VQC][14];     // du/dx (9, [c][k] = du_c/dx_k), dv, 1/W, dW/du / W (3)
    __shared__ double Rx[NN][VQC][4];  // R, dR/dx
    __shared__ double Tq[VQC][VIB_MAXF][9];
    __shared__ double mpart[VQC][VIB_MAXF];
    __shared__ double hel[3];
    const int t = threadIdx.x;
    const int nf = vm.nf;
    const int qq = t / G, part = t % G;
    for (int64_t gi = blockIdx.x; gi < pa.n_active; gi += gridDim.x) {
        const int e = pa.elems[gi];
        __syncthreads();   // the previous element is done with the LDS arrays
        for (int a = t; a < nn; a += 256) {
            const int64_t cp = (int64_t)pa.ien[(int64_t)nn * e + a] - 1;
            xw[a][0] = pa.coords[3 * cp];
            xw[a][1] = pa.coords[3 * cp + 1];
            xw[a][2] = pa.coords[3 * cp + 2];
            xw[a][3] = pa.weight[(int64_t)nn * e + a];
            for (int f = 0; f < nf; ++f)
                for (int c = 0; c < 3; ++c)
                    uf[f][a][c] = vm.vect[(vm.f0 + f) + (int64_t)vm.nf_total * (c + 3 * cp)];
        }
        // this thread's node sums: (node, mode) x = t + 256 s, a = x mod nn, f = x / nn
        constexpr int GSL = (NN * VIB_MAXF + 255) / 256;
        double gr[GSL][3];
#pragma unroll
        for (int s = 0; s < GSL; ++s) gr[s][0] = gr[s][1] = gr[s][2] = 0.0;
        double macc = 0.0;   // this thread's (point, mode) share of the element mass
        for (int q0 = 0; q0 < NQ; q0 += VQC) {
            const int nq = NQ - q0 < VQC ? NQ - q0 : VQC;
            __syncthreads();
            // 1. bases of direction `part` at point qq (shap.f90:75-178 for the parametric form)
            if (part < 3 && qq < nq) {
                const int q = q0 + qq, d = part;
                const int gd = d == 0 ? q / (ng * ng) : d == 1 ? (q / ng) % ng : q % ng;
                const double *U = d == 0 ? pa.U0 : d == 1 ? pa.U1 : pa.U2;
                const int span1 = pa.nijk[3 * (int64_t)e + d];
                const double u0 = U[span1 - 1], u1 = U[span1];
                const double u = ((u1 - u0) * pa.g.x[gd] + (u1 + u0)) * 0.5;
                double N[NP1], dN[NP1];
                if constexpr (RT) {
                    const int pd = d == 0 ? p0 : d == 1 ? p1 : p2;
                    bspline_ders_rt<P>(span1 - 1, pd, u, U, N, dN);
                    for (int i = 0; i <= pd; ++i) {
                        Bq[qq][d][i] = N[pd - i];
                        dBq[qq][d][i] = dN[pd - i];
                    }
                } else {
                    bspline_ders<P>(span1 - 1, u, U, N, dN);
#pragma unroll
                    for (int i = 0; i < NP1; ++i) {
                        Bq[qq][d][i] = N[P - i];
                        dBq[qq][d][i] = dN[P - i];
                    }
                }
                if (qq == 0) hel[d] = 0.5 * (u1 - u0);
            }
            __syncthreads();
            // 2. W, dW/du and dx/du of point qq: nodes part, part + G, ... then a shuffle sum
            {
                double s[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) s[k] = 0.0;
                if (qq < nq) {
#pragma unroll 1
                    for (int a = part; a < nn; a += G) {
                        const int i = a % n0, j = (a / n0) % n1, k = a / (n0 * n1);
                        const double w = xw[a][3], x0 = xw[a][0], x1 = xw[a][1], x2 = xw[a][2];
                        const double B0 = Bq[qq][0][i], B1 = Bq[qq][1][j], B2 = Bq[qq][2][k];
                        const double b12 = B1 * B2 * w;
                        const double Na = B0 * b12;
                        const double d0 = dBq[qq][0][i] * b12;
                        const double d1 = B0 * dBq[qq][1][j] * B2 * w;
                        const double d2 = B0 * B1 * dBq[qq][2][k] * w;
                        s[0] += Na;
                        s[1] += d0; s[2] += d1; s[3] += d2;
                        s[4] += x0 * Na; s[5] += x1 * Na; s[6] += x2 * Na;
                        s[7] += x0 * d0; s[8] += x0 * d1; s[9] += x0 * d2;
                        s[10] += x1 * d0; s[11] += x1 * d1; s[12] += x1 * d2;
                        s[13] += x2 * d0; s[14] += x2 * d1; s[15] += x2 * d2;
                    }
                }
#pragma unroll
                for (int o = 1; o < G; o <<= 1)   // the G threads of a point are adjacent lanes
#pragma unroll
                    for (int k = 0; k < 16; ++k) s[k] += __shfl_xor(s[k], o, 64);
                if (part == 0) {
                    double *gp = gq[qq];
                    if (qq < nq) {
                        const int q = q0 + qq;
                        const double Wi = 1.0 / s[0];
                        const double dw[3] = {s[1] * Wi, s[2] * Wi, s[3] * Wi};
                        double J[3][3];   // dx_r / du_c
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int c = 0; c < 3; ++c) J[r][c] = (s[7 + 3 * r + c] - s[4 + r] * dw[c]) * Wi;
                        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                                           J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                        const double di = 1.0 / det;
                        gp[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * di;
                        gp[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * di;
                        gp[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * di;
                        gp[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * di;
                        gp[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * di;
                        gp[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * di;
                        gp[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * di;
                        gp[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * di;
                        gp[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * di;
                        const int g0 = q / (ng * ng), g1 = (q / ng) % ng, g2 = q % ng;
                        gp[9] = (pa.g.w[g0] * pa.g.w[g1] * pa.g.w[g2]) * (det * hel[0] * hel[1] * hel[2]);
                        gp[10] = Wi;
                        gp[11] = dw[0];
                        gp[12] = dw[1];
                        gp[13] = dw[2];
                    } else {
                        for (int k = 0; k < 14; ++k) gp[k] = 0.0;
                    }
                }
            }
            __syncthreads();
            // 3. R = N w / W and dR/dx = (dN w / W - R dW / W) du/dx
            for (int x = t; x < nn * VQC; x += 256) {
                const int a = x / VQC, c = x % VQC;
                double r4[4] = {0.0, 0.0, 0.0, 0.0};
                if (c < nq) {
                    const int i = a % n0, j = (a / n0) % n1, k = a / (n0 * n1);
                    const double w = xw[a][3], Wi = gq[c][10];
                    const double B0 = Bq[c][0][i], B1 = Bq[c][1][j], B2 = Bq[c][2][k];
                    const double R = B0 * B1 * B2 * w * Wi;
                    const double du[3] = {dBq[c][0][i] * B1 * B2 * w * Wi - R * gq[c][11],
                                          B0 * dBq[c][1][j] * B2 * w * Wi - R * gq[c][12],
                                          B0 * B1 * dBq[c][2][k] * w * Wi - R * gq[c][13]};
                    r4[0] = R;
                    for (int kk = 0; kk < 3; ++kk)
                        r4[1 + kk] = du[0] * gq[c][kk] + du[1] * gq[c][3 + kk] + du[2] * gq[c][6 + kk];
                }
                for (int kk = 0; kk < 4; ++kk) Rx[a][c][kk] = r4[kk];
            }
            __syncthreads();
            // 4. per (point, mode): u, du/dx, stress, the tensor and the mass; the nodes of a
            // (point, mode) pair are split over 4 adjacent lanes and their sums added by shuffles
            {
                const int pr = t >> 2, part4 = t & 3, c = pr % VQC, f = pr / VQC;
                const bool act = f < nf && c < nq;
                double u[3] = {0.0, 0.0, 0.0}, D[3][3] = {};   // D[i][k] = du_i/dx_k
                if (act) {
#pragma unroll 1
                    for (int a = part4; a < nn; a += 4) {
                        const double R = Rx[a][c][0], g0 = Rx[a][c][1], g1 = Rx[a][c][2], g2 = Rx[a][c][3];
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            const double ua = uf[f][a][i];
                            u[i] += R * ua;
                            D[i][0] += ua * g0;
                            D[i][1] += ua * g1;
                            D[i][2] += ua * g2;
                        }
                    }
                }
#pragma unroll
                for (int o = 1; o < 4; o <<= 1) {
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        u[i] += __shfl_xor(u[i], o, 64);
#pragma unroll
                        for (int k = 0; k < 3; ++k) D[i][k] += __shfl_xor(D[i][k], o, 64);
                    }
                }
                if (part4 == 0 && f < nf) {
                    double T[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                    if (c < nq) {
                        const double lam = pa.lam, mu = pa.mu, dv = gq[c][9];
                        const double tr = D[0][0] + D[1][1] + D[2][2];
                        double sig[3][3], W = 0.0;
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                const double eik = 0.5 * (D[i][k] + D[k][i]);
                                sig[i][k] = 2.0 * mu * eik + (i == k ? lam * tr : 0.0);
                                W += eik * sig[i][k];
                            }
                        const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                        const double Wp = W - vm.vals[f] * rho * uu;
#pragma unroll
                        for (int m = 0; m < 3; ++m)
#pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                double v = m == k ? Wp : 0.0;
                                v -= 2.0 * (D[0][m] * sig[0][k] + D[1][m] * sig[1][k] + D[2][m] * sig[2][k]);
                                T[3 * m + k] = dv * v;
                            }
                        macc += rho * uu * dv;
                    }
                    for (int k = 0; k < 9; ++k) Tq[c][f][k] = T[k];
                }
            }
            __syncthreads();
            // 5. per (node, mode): g[m] += sum_q T[m][k] dR/dx_k
#pragma unroll
            for (int s = 0; s < GSL; ++s) {
                const int x = t + 256 * s;
                if (x >= nn * nf) continue;
                const int a = x % nn, f = x / nn;
                double g[3] = {0.0, 0.0, 0.0};
                for (int c = 0; c < nq; ++c) {
                    const double g0 = Rx[a][c][1], g1 = Rx[a][c][2], g2 = Rx[a][c][3];
#pragma unroll
                    for (int m = 0; m < 3; ++m)
                        g[m] += Tq[c][f][3 * m] * g0 + Tq[c][f][3 * m + 1] * g1 + Tq[c][f][3 * m + 2] * g2;
                }
#pragma unroll
                for (int m = 0; m < 3; ++m) gr[s][m] += g[m];
            }
        }
        if ((t & 3) == 0) mpart[(t >> 2) % VQC][(t >> 2) / VQC] = (t >> 2) / VQC < nf ? macc : 0.0;
        __syncthreads();
        const int64_t eb = ebase + gi * nn;
#pragma unroll
        for (int s = 0; s < GSL; ++s) {
            const int x = t + 256 * s;
            if (x >= nn * VIB_MAXF) continue;
            const int a = x % nn, f = x / nn;
#pragma unroll
            for (int m = 0; m < 3; ++m) eg[(eb + a) * (VIB_MAXF * 3) + 3 * f + m] = f < nf ? gr[s][m] : 0.0;
        }
        if (t < VIB_MAXF) {
            double ms = 0.0;
            for (int c = 0; c < VQC; ++c) ms += mpart[c][t];
            em[(mbase + gi) * VIB_MAXF + t] = ms;
        }
    }
}

// Total mass of mode f over every element slot, in a fixed order (one block per mode).
__global__ void __launch_bounds__(256) vib_mass_kernel(const double *em, int64_t nslot, double *mass) {
    __shared__ double sh[256];
    const int f = blockIdx.x;
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < nslot; i += 256) v += em[i * VIB_MAXF + f];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) mass[f] = sh[0];
}

// gradV(f0 + f, m, cp) = sum over the element-node entries of cp, in entry order, / mass[f]
__global__ void __launch_bounds__(256) vib_gather_kernel(int32_t nb_cp, const int32_t *ptr, const int64_t *ent,
                                                         const double *eg, const double *mass, VibModes vm,
                                                         double *gradV) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < 3 * (int64_t)nb_cp;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t cp = x / 3;
        const int m = (int)(x % 3);
        for (int f = 0; f < vm.nf; ++f) {
            double s = 0.0;
            for (int32_t j = ptr[cp]; j < ptr[cp + 1]; ++j) s += eg[ent[j] * (VIB_MAXF * 3) + 3 * f + m];
            const double v = s * vm.scale;
            gradV[(vm.f0 + f) + (int64_t)vm.nf_total * (m + 3 * cp)] = vm.normalize ? v / mass[f] : v;
        }
    }
}

template <int P, int NG, bool RT = false>
static int launch_vib(yeti_asm *h, PatchDev &pd, const VibModes &vm, int64_t ebase, int64_t mbase, hipStream_t st) {
    PatchArgs pa{};
    pa.coords = h->coords.p;
    pa.ien = pd.ien.p;
    pa.weight = pd.weight.p;
    pa.nijk = pd.nijk.p;
    pa.U0 = pd.U[0].p;
    pa.U1 = pd.U[1].p;
    pa.U2 = pd.U[2].p;
    pa.elems = pd.elems.p;
    pa.n_active = pd.n_active;
    pa.lam = pd.lam;
    pa.mu = pd.mu;
    gauss_legendre(RT ? pd.ng : NG, pa.g.x, pa.g.w);
    VibDegrees dg{{pd.p[0], pd.p[1], pd.p[2]}, pd.ng};
    const unsigned grid = (unsigned)std::min<int64_t>(pd.n_active, 1 << 20);
    const double rho = vm.unit_density ? 1.0 : pd.rho;
    hipLaunchKernelGGL((vib_grad_kernel<P, NG, RT>), dim3(grid), dim3(256), 0, st, pa, rho, vm, ebase, mbase,
                       h->vib_eg.p, h->vib_em.p, dg);
    YETI_HIP(hipGetLastError());
    return 0;
}

// CP -> element-node entries of the active U1 elements (entry = patch base + gi nnode + a), in
// patch, element and node order; buffers for the per-entry and per-element partials.
static int vib_setup(yeti_asm *h) {
    if (h->vib_nent >= 0) return 0;
    std::vector<int32_t> cnt(h->nb_cp + 1, 0);
    std::vector<std::vector<int32_t>> ien_h(h->patches.size()), el_h(h->patches.size());
    int64_t nent = 0, nslot = 0;
    for (size_t k = 0; k < h->patches.size(); ++k) {
        PatchDev &pd = *h->patches[k];
        if (!pd.u1 || pd.n_active == 0) continue;
        ien_h[k].resize((size_t)pd.nnode * pd.nelem);
        el_h[k].resize(pd.n_active);
        YETI_HIP(hipMemcpy(ien_h[k].data(), pd.ien.p, ien_h[k].size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        YETI_HIP(hipMemcpy(el_h[k].data(), pd.elems.p, el_h[k].size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int32_t gi = 0; gi < pd.n_active; ++gi)
            for (int a = 0; a < pd.nnode; ++a) {
                const int32_t id = ien_h[k][(size_t)pd.nnode * el_h[k][gi] + a];   // 1-based
                if (id < 1 || id > h->nb_cp)
                    return set_error(YETI_ERR_INVALID_ARGUMENT, "vibration gradient: IEN entry %d out of range", id);
                ++cnt[id];
            }
        nent += (int64_t)pd.n_active * pd.nnode;
        nslot += pd.n_active;
    }
    if (nent > INT32_MAX) return set_error(YETI_ERR_SIZE_OVERFLOW, "vibration gradient: %lld element nodes", (long long)nent);
    std::vector<int32_t> ptr(h->nb_cp + 1, 0);
    for (int32_t c = 0; c < h->nb_cp; ++c) ptr[c + 1] = ptr[c] + cnt[c + 1];
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    std::vector<int64_t> ent(std::max<int64_t>(nent, 1));
    int64_t base = 0;
    for (size_t k = 0; k < h->patches.size(); ++k) {
        PatchDev &pd = *h->patches[k];
        if (!pd.u1 || pd.n_active == 0) continue;
        for (int32_t gi = 0; gi < pd.n_active; ++gi)
            for (int a = 0; a < pd.nnode; ++a) {
                const int32_t cp = ien_h[k][(size_t)pd.nnode * el_h[k][gi] + a] - 1;
                ent[pos[cp]++] = base + (int64_t)gi * pd.nnode + a;
            }
        base += (int64_t)pd.n_active * pd.nnode;
    }
    YETI_HIP(h->vib_ptr.upload(ptr.data(), ptr.size()));
    YETI_HIP(h->vib_ent.upload(ent.data(), ent.size()));
    YETI_HIP(h->vib_eg.alloc((size_t)std::max<int64_t>(nent, 1) * VIB_MAXF * 3));
    YETI_HIP(h->vib_em.alloc((size_t)std::max<int64_t>(nslot, 1) * VIB_MAXF));
    YETI_HIP(h->vib_mass.alloc(VIB_MAXF));
    h->vib_nent = nent;
    h->vib_nslot = nslot;
    return 0;
}

static int run_vib_grad(yeti_asm *h, const double *vect, const double *vals, int32_t nf_total, double *gradV,
                        hipStream_t st, double scale = 1.0, bool normalize = true, bool unit_density = false) {
    for (auto *pd : h->patches) {
        if (!pd->u1 || pd->n_active == 0) continue;
        const bool fixed = mfma_kernel_available(pd->p[0], pd->p[1], pd->p[2], pd->ng);
        const bool general = pd->p[0] >= 1 && pd->p[1] >= 1 && pd->p[2] >= 1 && pd->p[0] <= GEN_MAXP &&
                             pd->p[1] <= GEN_MAXP && pd->p[2] <= GEN_MAXP && pd->nnode <= GEN_MAXN && pd->ng >= 1 &&
                             pd->ng <= MAX_NG;
        if (!fixed && !general)
            return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "vibration gradient: no kernel for degrees (%d,%d,%d), ng=%d",
                             pd->p[0], pd->p[1], pd->p[2], pd->ng);
    }
    int rc = vib_setup(h);
    if (rc) return rc;
    for (int32_t f0 = 0; f0 < nf_total; f0 += VIB_MAXF) {
        VibModes vm{};
        vm.vect = vect;
        vm.nf_total = nf_total;
        vm.f0 = f0;
        vm.nf = std::min<int32_t>(VIB_MAXF, nf_total - f0);
        for (int f = 0; f < vm.nf; ++f) vm.vals[f] = vals[f0 + f];
        vm.scale = scale;
        vm.normalize = normalize ? 1 : 0;
        vm.unit_density = unit_density ? 1 : 0;
        int64_t ebase = 0, mbase = 0;
        for (auto *pd : h->patches) {
            if (!pd->u1 || pd->n_active == 0) continue;
            if (!mfma_kernel_available(pd->p[0], pd->p[1], pd->p[2], pd->ng))
                rc = launch_vib<GEN_MAXP, MAX_NG, true>(h, *pd, vm, ebase, mbase, st);
            else switch (pd->p[0] * 16 + pd->ng) {
            case 1 * 16 + 2: rc = launch_vib<1, 2>(h, *pd, vm, ebase, mbase, st); break;
            case 2 * 16 + 3: rc = launch_vib<2, 3>(h, *pd, vm, ebase, mbase, st); break;
            case 3 * 16 + 4: rc = launch_vib<3, 4>(h, *pd, vm, ebase, mbase, st); break;
            default: rc = launch_vib<4, 5>(h, *pd, vm, ebase, mbase, st); break;
            }
            if (rc) return rc;
            ebase += (int64_t)pd->n_active * pd->nnode;
            mbase += pd->n_active;
        }
        if (h->vib_nslot > 0) {
            hipLaunchKernelGGL(vib_mass_kernel, dim3(vm.nf), dim3(256), 0, st, (const double *)h->vib_em.p, h->vib_nslot,
                               h->vib_mass.p);
        } else {   // no active element: zero mass, as the reference's empty sum
            hipLaunchKernelGGL(fill_zero_kernel, dim3(1), dim3(256), 0, st, h->vib_mass.p, (int32_t *)nullptr,
                               (int32_t *)nullptr, (int64_t)VIB_MAXF);
        }
        YETI_HIP(hipGetLastError());
        const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((3 * (int64_t)h->nb_cp + 255) / 256, 65536));
        hipLaunchKernelGGL(vib_gather_kernel, dim3(blocks), dim3(256), 0, st, h->nb_cp, (const int32_t *)h->vib_ptr.p,
                           (const int64_t *)h->vib_ent.p, (const double *)h->vib_eg.p, (const double *)h->vib_mass.p,
                           vm, gradV);
        YETI_HIP(hipGetLastError());
    }
    return 0;
}

// -d(u^T K u)/dX (3 per CP, COORDS3D order) over the active solid elements: the gradient kernel
// with eigenvalue 0 (no mass term), the sums scaled by -1 and not normalised. u is dof-interleaved
// (3 cp + c), which is the (1, 3, nb_cp) column-major layout of one mode.
static int run_strain_energy_grad(yeti_asm *h, const double *u, double *g, hipStream_t st) {
    const double zero = 0.0;
    return run_vib_grad(h, u, &zero, 1, g, st, -1.0, false);
}

// u = e_x at every control point (a rigid translation)
__global__ void translation_kernel(double *u, int64_t n3) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) u[i] = i % 3 == 0 ? 1.0 : 0.0;
}

// Volume of the active solid elements and its derivative dV/dX (3 per CP): the gradient kernel
// for the translation u = e_x with density 1 and eigenvalue -1. Then du/dx = 0, |u| = 1, the
// per-point tensor is w|J| I, so the CP sums are sum_q w|J| dR_a/dx_m = dV/dX_am, and the
// "mass" u^T M u is sum_q w|J| = V. vol (device, 1 value) and g (device, 3 nb_cp).
static int run_volume_grad(yeti_asm *h, double *vol, double *g, hipStream_t st) {
    const int64_t n3 = 3 * (int64_t)h->nb_cp;
    if (!h->vol_u.p) YETI_HIP(h->vol_u.alloc((size_t)std::max<int64_t>(n3, 1)));
    hipLaunchKernelGGL(translation_kernel, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, h->vol_u.p, n3);
    YETI_HIP(hipGetLastError());
    const double eig = -1.0;
    int rc = run_vib_grad(h, h->vol_u.p, &eig, 1, g, st, 1.0, false, true);
    if (rc) return rc;
    YETI_HIP(hipMemcpyAsync(vol, h->vib_mass.p, sizeof(double), hipMemcpyDeviceToDevice, st));
    return 0;
}

// u^T F shape derivative of the face and body loads: g = 2 u^T dF/dX (3 per CP), gathered like F.
// reference_terms: only the terms gradLinElastWork_AN computes (centrifugal loads, their axis
// points read as it reads them; its face-load term is a stub, optim/gradUELMAT_solid.f90:463-477).
static int run_load_grad(yeti_asm *h, const double *u, double *g, hipStream_t st, bool reference_terms = false) {
    if (h->n_loaded_total > 0 && !h->grad_buf.p) YETI_HIP(h->grad_buf.alloc((size_t)h->n_loaded_total * 3 * RHS_MAXN));
    LoadTable lt;
    lt.jdltype = h->jdl.p;
    lt.adlmag = h->adlmag.p;
    lt.add_off = reference_terms ? h->add_off_ref.p : h->add_off.p;
    lt.only_type = reference_terms ? 101 : 0;
    lt.add_infos = h->add_infos.p;
    lt.elem_loads = h->elem_loads.p;
    lt.elem_loads_ptr = h->elem_loads_ptr.p;
    lt.nodal_dist = h->nodal_dist.p;
    lt.nb_n_dist = h->nb_n_dist;
    for (auto *pd : h->patches) {
        if (pd->n_loaded == 0) continue;
        RhsPatch rp;
        rp.coords = h->coords.p;
        rp.ien = pd->ien.p;
        rp.weight = pd->weight.p;
        rp.nijk = pd->nijk.p;
        for (int d = 0; d < 3; ++d) {
            rp.U[d] = pd->U[d].p;
            rp.p[d] = pd->p[d];
        }
        rp.elems = pd->loaded.p;
        rp.n_loaded = pd->n_loaded;
        rp.slot0 = pd->slot0;
        rp.ng = pd->ng;
        rp.rho = pd->rho;
        gauss_legendre(pd->ng, rp.g.x, rp.g.w);
        hipLaunchKernelGGL(elem_load_grad_kernel, dim3(pd->n_loaded), dim3(64), 0, st, rp, lt, u, h->grad_buf.p);
        YETI_HIP(hipGetLastError());
    }
    int blocks = (h->nb_cp + 255) / 256;
    hipLaunchKernelGGL(rhs_gather_kernel, dim3(blocks), dim3(256), 0, st, h->nb_cp, (const double *)nullptr,
                       h->cp_ptr.p, h->cp_ent.p, (const double *)h->grad_buf.p, g);
    YETI_HIP(hipGetLastError());
    return 0;
}

// ---- gradLinElastWork_AN: internal and external work gradients for given SOL and ADJ --------
// Fields f of an (nf, 3, nb_cp) column-major array sit at f + nf (3 cp + d).

// sums[f] = sum over (d, cp) of V[f, d, cp]^2, one block per field, fixed order (deterministic)
__global__ void __launch_bounds__(256) field_sumsq_kernel(const double *V, int32_t nf, int64_t n3, double *sums) {
    const int f = blockIdx.x;
    double s = 0.0;
    for (int64_t j = threadIdx.x; j < n3; j += blockDim.x) {
        const double v = V[f + (int64_t)nf * j];
        s += v * v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[f] = red[0];
}

// V[2k] = u + s_k a_k, V[2k + 1] = u - s_k a_k (fields of an (2 nadj, 3, nb_cp) array)
__global__ void polar_fields_kernel(const double *u, const double *adj, int32_t nadj, int64_t n3, const double *scale,
                                    double *V) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n3 * nadj) return;
    const int64_t j = x / nadj;
    const int k = (int)(x % nadj);
    const double a = scale[k] * adj[k + (int64_t)nadj * j];
    V[2 * k + (int64_t)(2 * nadj) * j] = u[j] + a;
    V[2 * k + 1 + (int64_t)(2 * nadj) * j] = u[j] - a;
}

// out[k, j] = (G[2k, j] - G[2k + 1, j]) / (4 s_k): the polarisation of the quadratic form
__global__ void polar_combine_kernel(const double *G, int32_t nadj, int64_t n3, const double *scale, double *out) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n3 * nadj) return;
    const int64_t j = x / nadj;
    const int k = (int)(x % nadj);
    out[k + (int64_t)nadj * j] =
        (G[2 * k + (int64_t)(2 * nadj) * j] - G[2 * k + 1 + (int64_t)(2 * nadj) * j]) / (4.0 * scale[k]);
}

// a[j] = ADJ[k, j] and back: out[k, j] = c g[j]
__global__ void field_extract_kernel(const double *adj, int32_t nadj, int32_t k, int64_t n3, double *a) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n3) a[j] = adj[k + (int64_t)nadj * j];
}
__global__ void field_insert_kernel(const double *g, double c, int32_t nadj, int32_t k, int64_t n3, double *out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n3) out[k + (int64_t)nadj * j] = c * g[j];
}

// gradWint[k] = -ADJ_k^T (dK/dX) SOL and gradWext[k] = ADJ_k^T (dF/dX) over the active solid
// elements (gradtotalwork_fullAN.f90:17-300). The internal term is bilinear and symmetric, and
// the gradient kernel gives the quadratic form q(v) = -v^T (dK/dX) v, so
//   -a^T (dK/dX) u = (q(u + s a) - q(u - s a)) / (4 s),
// with s = |u| / |a| so that both fields have the norm of u (no cancellation beyond rounding);
// the 2 nadj fields run through the kernel together (its field-independent work is shared).
static int run_grad_lin_elast_work(yeti_asm *h, const double *sol, const double *adj, int32_t nadj, bool wint,
                                   bool wext, double *gint, double *gext, hipStream_t st) {
    const int64_t n3 = 3 * (int64_t)h->nb_cp;
    const unsigned blk = (unsigned)((n3 * nadj + 255) / 256), blk1 = (unsigned)((n3 + 255) / 256);
    if (h->gw_sums.n < (size_t)(2 * nadj + 1)) YETI_HIP(h->gw_sums.alloc(2 * (size_t)nadj + 1));
    if (wint) {
        if (h->gw_fields.n < (size_t)(2 * nadj * n3)) {
            YETI_HIP(h->gw_fields.alloc((size_t)(2 * nadj * n3)));
            YETI_HIP(h->gw_grads.alloc((size_t)(2 * nadj * n3)));
        }
        double *sums = h->gw_sums.p, *scale = h->gw_sums.p + nadj + 1;
        hipLaunchKernelGGL(field_sumsq_kernel, dim3(1), dim3(256), 0, st, sol, 1, n3, sums);
        hipLaunchKernelGGL(field_sumsq_kernel, dim3(nadj), dim3(256), 0, st, adj, nadj, n3, sums + 1);
        YETI_HIP(hipGetLastError());
        std::vector<double> hs(nadj + 1), sc(nadj);
        YETI_HIP(hipMemcpyAsync(hs.data(), sums, (nadj + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
        YETI_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < nadj; ++k)
            sc[k] = (hs[0] > 0.0 && hs[k + 1] > 0.0) ? std::sqrt(hs[0] / hs[k + 1]) : 1.0;
        YETI_HIP(hipMemcpyAsync(scale, sc.data(), nadj * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(polar_fields_kernel, dim3(blk), dim3(256), 0, st, sol, adj, nadj, n3, (const double *)scale,
                           h->gw_fields.p);
        YETI_HIP(hipGetLastError());
        std::vector<double> zeros(2 * nadj, 0.0);
        int rc = run_vib_grad(h, h->gw_fields.p, zeros.data(), 2 * nadj, h->gw_grads.p, st, -1.0, false);
        if (rc) return rc;
        hipLaunchKernelGGL(polar_combine_kernel, dim3(blk), dim3(256), 0, st, (const double *)h->gw_grads.p, nadj, n3,
                           (const double *)scale, gint);
        YETI_HIP(hipGetLastError());
        // the host copy of the scales must outlive the copy above
        YETI_HIP(hipStreamSynchronize(st));
    } else {
        hipLaunchKernelGGL(fill_zero_kernel, dim3(blk), dim3(256), 0, st, gint, (int32_t *)nullptr, (int32_t *)nullptr,
                           n3 * nadj);
        YETI_HIP(hipGetLastError());
    }
    if (wext) {
        if (!h->ref_infos_ok)
            return set_error(YETI_ERR_INVALID_ARGUMENT,
                             "load_additionalInfos holds fewer than 6 values per centrifugal load");
        if (h->gw_vec.n < (size_t)n3) {
            YETI_HIP(h->gw_vec.alloc(n3));
            YETI_HIP(h->gw_vec2.alloc(n3));
        }
        for (int k = 0; k < nadj; ++k) {
            hipLaunchKernelGGL(field_extract_kernel, dim3(blk1), dim3(256), 0, st, adj, nadj, k, n3, h->gw_vec.p);
            YETI_HIP(hipGetLastError());
            int rc = run_load_grad(h, h->gw_vec.p, h->gw_vec2.p, st, true);
            if (rc) return rc;
            // run_load_grad gives 2 a^T dF/dX
            hipLaunchKernelGGL(field_insert_kernel, dim3(blk1), dim3(256), 0, st, (const double *)h->gw_vec2.p, 0.5, nadj,
                               k, n3, gext);
            YETI_HIP(hipGetLastError());
        }
    } else {
        hipLaunchKernelGGL(fill_zero_kernel, dim3(blk), dim3(256), 0, st, gext, (int32_t *)nullptr, (int32_t *)nullptr,
                           n3 * nadj);
        YETI_HIP(hipGetLastError());
    }
    return 0;
}

static int run_rhs(yeti_asm *h, double *F, hipStream_t st) {
    LoadTable lt;
    lt.jdltype = h->jdl.p;
    lt.adlmag = h->adlmag.p;
    lt.add_off = h->add_off.p;
    lt.add_infos = h->add_infos.p;
    lt.elem_loads = h->elem_loads.p;
    lt.elem_loads_ptr = h->elem_loads_ptr.p;
    lt.nodal_dist = h->nodal_dist.p;
    lt.nb_n_dist = h->nb_n_dist;
    for (auto *pd : h->patches) {
        if (pd->n_loaded == 0) continue;
        RhsPatch rp;
        rp.coords = h->coords.p;
        rp.ien = pd->ien.p;
        rp.weight = pd->weight.p;
        rp.nijk = pd->nijk.p;
        for (int d = 0; d < 3; ++d) {
            rp.U[d] = pd->U[d].p;
            rp.p[d] = pd->p[d];
        }
        rp.elems = pd->loaded.p;
        rp.n_loaded = pd->n_loaded;
        rp.slot0 = pd->slot0;
        rp.ng = pd->ng;
        rp.rho = pd->rho;
        gauss_legendre(pd->ng, rp.g.x, rp.g.w);
        hipLaunchKernelGGL(elem_rhs_kernel, dim3(pd->n_loaded), dim3(64), rhs_smem_bytes(rp), st, rp, lt, h->rhs_buf.p);
        YETI_HIP(hipGetLastError());
    }
    int blocks = (h->nb_cp + 255) / 256;
    hipLaunchKernelGGL(rhs_gather_kernel, dim3(blocks), dim3(256), 0, st, h->nb_cp, h->F_conc.p,
                       h->cp_ptr.p, h->cp_ent.p, h->rhs_buf.p, F);
    YETI_HIP(hipGetLastError());
    return 0;
}

static int create(const yeti_system_inputs *in, int layout, yeti_asm **out) {
    *out = nullptr;
    if (!in) return set_error(YETI_ERR_INVALID_ARGUMENT, "null inputs");
    if (layout != YETI_LAYOUT_PACKED && layout != YETI_LAYOUT_SLOTTED)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "unknown layout %d", layout);
    // every id the set-up and the kernels gather through is checked here, before any device call
    if (int rc0 = validate_system_inputs(in, false)) return rc0;
    if ((int64_t)in->nb_cp * 3 > INT32_MAX)
        return set_error(YETI_ERR_SIZE_OVERFLOW, "dof ids exceed int32");
    std::vector<std::string> types;
    int rc = parse_elt_types(in, types);
    if (rc) return rc;

    auto *h = new yeti_asm();
    h->layout = layout;
    h->nb_cp = in->nb_cp;
    h->nb_dof_tot = in->nb_dof_tot;
    h->nb_data_cap = in->nb_data;
    h->nb_n_dist = std::max(1, in->nb_n_dist);
    auto fail = [&](int code) {
        delete h;
        return code;
    };
    if (hipError_t e = h->coords.upload(in->COORDS3D, 3 * (size_t)in->nb_cp); e != hipSuccess)
        return fail(set_error(YETI_ERR_DEVICE, "upload coords: %s", hipGetErrorString(e)));

    // load bookkeeping: which loads act on which (global, 1-based) element
    std::vector<int64_t> load_off(in->nb_load + 1, 0);
    for (int il = 0; il < in->nb_load; ++il) load_off[il + 1] = load_off[il] + in->load_target_nbelem[il];
    std::vector<std::vector<int32_t>> loads_of_elem(in->nb_elem);
    for (int il = 0; il < in->nb_load; ++il) {
        int t = in->JDLType[il];
        bool elem_load = (t > 9 && t < 100) || t == 101 || t == 102;
        if (!elem_load) continue;
        for (int64_t k = load_off[il]; k < load_off[il + 1]; ++k) {
            int64_t el = in->indDLoad[k];
            if (el >= 1 && el <= in->nb_elem) {
                auto &v = loads_of_elem[el - 1];
                if (v.empty() || v.back() != il) v.push_back(il);
            }
        }
    }

    int64_t kv = 0, woff = 0, ieoff = 0, eloff = 0, count = 0;
    std::vector<int32_t> elem_loads, elem_loads_ptr(1, 0);
    std::vector<std::pair<int32_t, int64_t>> cp_contrib;   // (cp0, entry)
    int32_t slot = 0;
    for (int ip = 0; ip < in->nb_patch; ++ip) {
        auto *pd = new PatchDev();
        h->patches.push_back(pd);
        for (int d = 0; d < 3; ++d) pd->p[d] = in->Jpqr[3 * ip + d];
        int nkv[3] = {in->Nkv[3 * ip], in->Nkv[3 * ip + 1], in->Nkv[3 * ip + 2]};
        pd->nnode = in->NNODE[ip];
        pd->nelem = in->nb_elem_patch[ip];
        int nn_geo = (pd->p[0] + 1) * (pd->p[1] + 1) * (pd->p[2] + 1);
        const std::string &ty = types[ip];
        pd->u1 = ty == "U1";
        bool u0 = ty == "U0";
        pd->ng = gauss_per_dir(in->NBINT[ip]);
        double E = in->MATERIAL_PROPERTIES[(int64_t)in->n_mat_props_max * ip];
        double nu = in->MATERIAL_PROPERTIES[(int64_t)in->n_mat_props_max * ip + 1];
        pd->lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        pd->mu = E / (2.0 * (1.0 + nu));
        pd->rho = in->RHO[ip];
        const int64_t nde = 9LL * pd->nnode * (pd->nnode - 1) / 2 + 6LL * pd->nnode;
        // active elements and their slots
        std::vector<int32_t> act;
        std::vector<int64_t> offs;
        int64_t c = count;
        for (int e = 0; e < pd->nelem; ++e) {
            bool a = in->activeElement[eloff + e] == 1;
            if (layout == YETI_LAYOUT_SLOTTED) {
                if (a) {
                    act.push_back(e);
                    offs.push_back(count + (int64_t)e * nde);
                } else {
                    h->zero_ranges.push_back({count + (int64_t)e * nde, nde});
                }
            } else if (a) {
                act.push_back(e);
                offs.push_back(c);
                c += nde;
            }
        }
        int64_t next = layout == YETI_LAYOUT_SLOTTED ? count + (int64_t)pd->nelem * nde : c;
        if (!act.empty() && !pd->u1 && !u0)
            return fail(set_error(YETI_ERR_UNSUPPORTED_ELEMENT, "patch %d: element type %s is not supported",
                                  ip + 1, ty.c_str()));
        if (!act.empty()) {
            if (pd->nnode != nn_geo)
                return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "patch %d: NNODE=%d but degrees give %d",
                                      ip + 1, pd->nnode, nn_geo));
            if (!kernel_available(pd->p[0], pd->p[1], pd->p[2], pd->ng))
                return fail(set_error(YETI_ERR_UNSUPPORTED_DEGREE,
                                      "patch %d: degrees (%d,%d,%d) with %d Gauss points per direction",
                                      ip + 1, pd->p[0], pd->p[1], pd->p[2], pd->ng));
        }
        // void elements (U0) keep their slots with zero values and real dof ids, as the
        // reference emits its zero AMATRX (build_system_elemStorage.f90:185-188,260-293)
        if (u0) pd->lam = pd->mu = 0.0;
        pd->u1 = pd->u1 || u0;
        pd->n_active = (int32_t)act.size();
        hipError_t e1 = hipSuccess;
        for (int d = 0; d < 3 && e1 == hipSuccess; ++d) {
            e1 = pd->U[d].upload(in->Ukv + kv, nkv[d]);
            kv += nkv[d];
        }
        if (e1 == hipSuccess) e1 = pd->ien.upload(in->IEN + ieoff, (size_t)pd->nnode * pd->nelem);
        if (e1 == hipSuccess) e1 = pd->weight.upload(in->weight + woff, (size_t)nn_geo * pd->nelem);
        if (e1 == hipSuccess) e1 = pd->nijk.upload(in->Nijk + 3 * eloff, 3 * (size_t)pd->nelem);
        if (e1 == hipSuccess) e1 = pd->elems.upload(act.data(), act.size());
        if (e1 == hipSuccess) e1 = pd->offs.upload(offs.data(), offs.size());
        // loaded elements of this patch
        std::vector<int32_t> loaded;
        pd->slot0 = slot;
        if (pd->u1 && !u0) {
            for (int e = 0; e < pd->nelem; ++e) {
                if (in->activeElement[eloff + e] != 1) continue;
                auto &ls = loads_of_elem[eloff + e];
                if (ls.empty()) continue;
                loaded.push_back(e);
                for (int il : ls) elem_loads.push_back(il);
                elem_loads_ptr.push_back((int32_t)elem_loads.size());
                for (int a = 0; a < pd->nnode; ++a) {
                    int32_t cp = in->IEN[ieoff + (int64_t)pd->nnode * e + a] - 1;
                    cp_contrib.push_back({cp, (int64_t)slot * RHS_MAXN + a});
                }
                ++slot;
            }
        }
        pd->n_loaded = (int32_t)loaded.size();
        if (e1 == hipSuccess) e1 = pd->loaded.upload(loaded.data(), loaded.size());
        if (e1 != hipSuccess) return fail(set_error(YETI_ERR_DEVICE, "upload patch %d: %s", ip + 1, hipGetErrorString(e1)));
        if (pd->u1 && pd->nnode > RHS_MAXN)
            return fail(set_error(YETI_ERR_UNSUPPORTED_DEGREE, "patch %d: too many nodes", ip + 1));
        woff += (int64_t)nn_geo * pd->nelem;
        ieoff += (int64_t)pd->nnode * pd->nelem;
        eloff += pd->nelem;
        count = next;
    }
    h->nb_data = count;
    h->n_loaded_total = slot;
    for (auto *pd : h->patches) {
        h->v_ien.push_back(pd->ien.p);
        h->v_elems.push_back(pd->elems.p);
        h->v_offs.push_back(pd->offs.p);
        h->v_nnode.push_back(pd->nnode);
        h->v_nactive.push_back(pd->u1 ? pd->n_active : 0);
    }
    // concentrated loads: assignment in load order (build_system_elemStorage.f90:137-152)
    std::vector<double> Fc((size_t)in->nb_dof_tot, 0.0);
    for (int il = 0; il < in->nb_load; ++il) {
        int t = in->JDLType[il];
        if (t / 10 < 1)
            for (int64_t k = load_off[il]; k < load_off[il + 1]; ++k) {
                int64_t dof = (int64_t)(in->indDLoad[k] - 1) * 3 + t - 1;
                if (dof < 0 || dof >= in->nb_dof_tot)
                    return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "concentrated load %d targets dof %lld",
                                          il + 1, (long long)dof));
                Fc[dof] = in->ADLMAG[il];
            }
    }
    // CP -> RHS entries, element order kept by a stable sort on the CP id
    std::stable_sort(cp_contrib.begin(), cp_contrib.end(),
                     [](const std::pair<int32_t, int64_t> &a, const std::pair<int32_t, int64_t> &b) {
                         return a.first < b.first;
                     });
    std::vector<int32_t> cp_ptr(in->nb_cp + 1, 0), cp_ent(cp_contrib.size());
    for (size_t k = 0; k < cp_contrib.size(); ++k) {
        cp_ptr[cp_contrib[k].first + 1]++;
        cp_ent[k] = (int32_t)cp_contrib[k].second;
    }
    for (int i = 0; i < in->nb_cp; ++i) cp_ptr[i + 1] += cp_ptr[i];
    std::vector<int32_t> jdl(in->JDLType, in->JDLType + in->nb_load);
    std::vector<int32_t> add_off(in->nb_load + 1, 0);
    for (int il = 0; il < in->nb_load; ++il) add_off[il + 1] = add_off[il] + in->nb_load_additionalInfos[il];
    // gradLinElastWork_AN reads the k-th centrifugal load's axis points A, B at values 6k+1..6k+6
    // of load_additionalInfos, whatever the other loads hold (its loadcount advances for type 101
    // only, optim/gradUELMAT_solid.f90:378-398)
    std::vector<int32_t> add_off_ref(in->nb_load + 1, 0);
    int64_t n101 = 0;
    for (int il = 0; il < in->nb_load; ++il)
        if (in->JDLType[il] == 101) add_off_ref[il] = (int32_t)(6 * n101++);
    h->ref_infos_ok = 6 * n101 <= std::max<int64_t>(0, in->len_load_additionalInfos);
    if (in->nb_load > 0 && add_off[in->nb_load] > std::max<int64_t>(0, in->len_load_additionalInfos))
        return fail(set_error(YETI_ERR_INVALID_ARGUMENT, "load_additionalInfos holds %lld values, the loads name %d",
                              (long long)in->len_load_additionalInfos, add_off[in->nb_load]));
    // a nodal-distribution load (type 4) reads row `field` of nodal_dist at every CP of its elements
    for (int il = 0; il < in->nb_load; ++il) {
        const int t = in->JDLType[il];
        if (t < 10 || t > 99 || t % 10 != 4) continue;
        const int field = in->nb_load_additionalInfos[il] > 0 ? (int)in->load_additionalInfos[add_off[il]] : 0;
        if (field < 1 || field > std::max(1, in->nb_n_dist) || in->nb_cp_n_dist < in->nb_cp)
            return fail(set_error(YETI_ERR_INVALID_ARGUMENT,
                                  "load %d: distribution %d not in nodal_dist (%d x %d, %d control points)", il + 1,
                                  field, in->nb_n_dist, in->nb_cp_n_dist, in->nb_cp));
    }
    hipError_t e2 = h->F_conc.upload(Fc.data(), Fc.size());
    if (e2 == hipSuccess) e2 = h->jdl.upload(jdl.data(), jdl.size());
    if (e2 == hipSuccess) e2 = h->adlmag.upload(in->ADLMAG, in->nb_load);
    if (e2 == hipSuccess) e2 = h->add_off.upload(add_off.data(), add_off.size());
    if (e2 == hipSuccess) e2 = h->add_off_ref.upload(add_off_ref.data(), add_off_ref.size());
    if (e2 == hipSuccess) e2 = h->add_infos.upload(in->load_additionalInfos, std::max<int64_t>(1, in->len_load_additionalInfos));
    if (e2 == hipSuccess) e2 = h->elem_loads.upload(elem_loads.data(), elem_loads.size());
    if (e2 == hipSuccess) e2 = h->elem_loads_ptr.upload(elem_loads_ptr.data(), elem_loads_ptr.size());
    if (e2 == hipSuccess) e2 = h->cp_ptr.upload(cp_ptr.data(), cp_ptr.size());
    if (e2 == hipSuccess) e2 = h->cp_ent.upload(cp_ent.data(), cp_ent.size());
    if (e2 == hipSuccess)
        e2 = h->nodal_dist.upload(in->nodal_dist, (size_t)std::max(1, in->nb_n_dist) * std::max(1, in->nb_cp_n_dist));
    if (e2 == hipSuccess && slot > 0) e2 = h->rhs_buf.alloc((size_t)slot * 3 * RHS_MAXN);
    if (e2 != hipSuccess) return fail(set_error(YETI_ERR_DEVICE, "upload loads: %s", hipGetErrorString(e2)));
    *out = h;
    return 0;
}

struct AsmView {
    int32_t nb_cp;
    int32_t npatch;
    const int32_t *const *ien;
    const int32_t *const *elems;
    const int64_t *const *offs;
    const int32_t *nnode;
    const int32_t *n_active;
};

int asm_view(const yeti_asm *h, AsmView *v) {
    if (!h) return set_error(YETI_ERR_INVALID_ARGUMENT, "null assembly handle");
    v->nb_cp = h->nb_cp;
    v->npatch = (int32_t)h->patches.size();
    v->ien = h->v_ien.data();
    v->elems = h->v_elems.data();
    v->offs = h->v_offs.data();
    v->nnode = h->v_nnode.data();
    v->n_active = h->v_nactive.data();
    return 0;
}

}  // namespace yeti

// ------------------------------------------------------------------------------------------
// C ABI
extern "C" {

const char *yeti_last_error(void) { return yeti::last_error().c_str(); }

int32_t yeti_version(void) { return 0x000100; }

int32_t yeti_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int32_t yeti_asm_create(const yeti_system_inputs *in, int32_t layout, yeti_asm **out) {
    if (!out) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null handle pointer");
    yeti::last_error().clear();
    return yeti::create(in, layout, out);
}

int32_t yeti_asm_create_in(const yeti_system_inputs *in, int32_t layout, void *arena, size_t arena_bytes,
                           void *stream, yeti_asm **out, size_t *used) {
    if (!out) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null handle pointer");
    *out = nullptr;
    yeti::last_error().clear();
    yeti::Arena a;
    a.base = (char *)arena;
    a.size = arena ? arena_bytes : 0;
    yeti::AllocScope scope(&a, (hipStream_t)stream);
    yeti_asm *h = nullptr;
    int rc = yeti::create(in, layout, &h);
    if (used) *used = a.used;
    if (rc) return rc;
    if (!arena) {
        delete h;   // counting only
        return 0;
    }
    *out = h;
    return 0;
}

int64_t yeti_asm_nb_data(const yeti_asm *h) { return h ? h->nb_data : -1; }

int32_t yeti_asm_run_coo(yeti_asm *h, double *Kd, int32_t *Kr, int32_t *Kc, double *F, void *stream) {
    if (!h || !Kd || !Kr || !Kc || !F) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = yeti::run_elements(h, Kd, Kr, Kc, st);
    if (rc) return rc;
    return yeti::run_rhs(h, F, st);
}

int32_t yeti_asm_run_coo_matrix(yeti_asm *h, double *Kd, int32_t *Kr, int32_t *Kc, void *stream) {
    if (!h || !Kd || !Kr || !Kc) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return yeti::run_elements(h, Kd, Kr, Kc, (hipStream_t)stream);
}

int32_t yeti_asm_run_coo_values(yeti_asm *h, double *Kd, void *stream) {
    if (!h || !Kd) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return yeti::run_elements(h, Kd, nullptr, nullptr, (hipStream_t)stream);
}

int32_t yeti_asm_run_mass_coo(yeti_asm *h, double *Md, int32_t *Mr, int32_t *Mc, void *stream) {
    if (!h || !Md || (!Mr) != (!Mc)) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return yeti::run_mass(h, Md, Mr, Mc, (hipStream_t)stream);
}

int32_t yeti_asm_run_vib_grad(yeti_asm *h, const double *vect, const double *vals, int32_t nb_frq, double *gradV,
                              void *stream) {
    if (!h || !vect || !vals || !gradV || nb_frq < 1)
        return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument or nb_frq < 1");
    return yeti::run_vib_grad(h, vect, vals, nb_frq, gradV, (hipStream_t)stream);
}

int32_t yeti_asm_strain_energy_grad(yeti_asm *h, const double *u, double *g, void *stream) {
    if (!h || !u || !g) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return yeti::run_strain_energy_grad(h, u, g, (hipStream_t)stream);
}

int32_t yeti_asm_volume_grad(yeti_asm *h, double *vol, double *g, void *stream) {
    if (!h || !vol || !g) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return yeti::run_volume_grad(h, vol, g, (hipStream_t)stream);
}

int32_t yeti_grad_lin_elast_work(yeti_asm *h, const double *SOL, const double *ADJ, int32_t nadj, int32_t computeWint,
                                 int32_t computeWext, double *gradWint, double *gradWext, void *stream) {
    if (!h || !SOL || !ADJ || !gradWint || !gradWext || nadj < 1)
        return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument or nadj < 1");
    return yeti::run_grad_lin_elast_work(h, SOL, ADJ, nadj, computeWint != 0, computeWext != 0, gradWint, gradWext,
                                         (hipStream_t)stream);
}

int32_t yeti_asm_run_rhs(yeti_asm *h, double *F, void *stream) {
    if (!h || !F) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return yeti::run_rhs(h, F, (hipStream_t)stream);
}

int32_t yeti_asm_load_shape_grad(yeti_asm *h, const double *u, double *g, void *stream) {
    if (!h || !u || !g) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    return yeti::run_load_grad(h, u, g, (hipStream_t)stream);
}

int32_t yeti_asm_update_coords(yeti_asm *h, const double *coords, void *stream) {
    if (!h || !coords) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    YETI_HIP(hipMemcpyAsync(h->coords.p, coords, 3 * sizeof(double) * (size_t)h->nb_cp, hipMemcpyDeviceToDevice,
                            (hipStream_t)stream));
    return 0;
}

int32_t yeti_asm_set_timing(yeti_asm *h, int32_t on) {
    if (!h) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null handle");
    h->timing = on != 0;
    return 0;
}

int32_t yeti_asm_kernel_times(yeti_asm *h, float *ms, int32_t cap, int32_t *n) {
    if (!h || !n) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    *n = h->ev_used;
    for (int k = 0; k < h->ev_used && k < cap; ++k) {
        YETI_HIP(hipEventSynchronize(h->ev[2 * k + 1]));
        YETI_HIP(hipEventElapsedTime(&ms[k], h->ev[2 * k], h->ev[2 * k + 1]));
    }
    return 0;
}

void yeti_asm_destroy(yeti_asm *h) { delete h; }

int32_t yeti_sys_linmat_lindef_static(const yeti_system_inputs *in, int32_t layout, double *Kdata,
                                      int32_t *Krow, int32_t *Kcol, double *F) {
    if (!in || !Kdata || !Krow || !Kcol || !F) return yeti::set_error(YETI_ERR_INVALID_ARGUMENT, "null argument");
    yeti::last_error().clear();
    yeti_asm *h = nullptr;
    int rc = yeti::create(in, layout, &h);
    if (rc) return rc;
    int64_t nd = h->nb_data;
    if (nd > in->nb_data) {
        yeti_asm_destroy(h);
        return yeti::set_error(YETI_ERR_SIZE_OVERFLOW, "nb_data=%lld but the layout needs %lld",
                               (long long)in->nb_data, (long long)nd);
    }
    yeti::DevBuf<double> dK, dF;
    yeti::DevBuf<int32_t> dR, dC;
    hipError_t e = dK.alloc(std::max<int64_t>(nd, 1));
    if (e == hipSuccess) e = dR.alloc(std::max<int64_t>(nd, 1));
    if (e == hipSuccess) e = dC.alloc(std::max<int64_t>(nd, 1));
    if (e == hipSuccess) e = dF.alloc(std::max(in->nb_dof_tot, 1));
    if (e != hipSuccess) {
        yeti_asm_destroy(h);
        return yeti::set_error(YETI_ERR_DEVICE, "device allocation: %s", hipGetErrorString(e));
    }
    rc = yeti_asm_run_coo(h, dK.p, dR.p, dC.p, dF.p, nullptr);
    yeti_asm_destroy(h);
    if (rc) return rc;
    YETI_HIP(hipMemcpy(Kdata, dK.p, nd * sizeof(double), hipMemcpyDeviceToHost));
    YETI_HIP(hipMemcpy(Krow, dR.p, nd * sizeof(int32_t), hipMemcpyDeviceToHost));
    YETI_HIP(hipMemcpy(Kcol, dC.p, nd * sizeof(int32_t), hipMemcpyDeviceToHost));
    YETI_HIP(hipMemcpy(F, dF.p, in->nb_dof_tot * sizeof(double), hipMemcpyDeviceToHost));
    // entries beyond what the layout uses stay zero, as the reference's zero-initialised arrays
    if (in->nb_data > nd) {
        std::memset(Kdata + nd, 0, (in->nb_data - nd) * sizeof(double));
        std::memset(Krow + nd, 0, (in->nb_data - nd) * sizeof(int32_t));
        std::memset(Kcol + nd, 0, (in->nb_data - nd) * sizeof(int32_t));
    }
    return 0;
}

}  // extern "C"
