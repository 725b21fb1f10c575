// post.hip -- the steps after the solve: the full displacement from the free-dof solution,
// reconstruction.f:22-135 (U = scatter of the free values, prescribed values of the boundary
// conditions in their order, sol = U reshaped to (nb_cp, MCRD) column-major), the
// displacement at parametric points, evaluate_disp.f:20-189, and the VTU output of solid
// patches, postproc_vtu.f:22-297 (fields at each element's vertex grid on the GPU,
// compute_svars_solid.f:24-188; the text written on the host, writepatch.f).
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <tuple>

#include "common.hpp"

namespace yeti {

__global__ void __launch_bounds__(256) recon_fill_kernel(int64_t n, double *U) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        U[i] = 0.0;
}

// U[ind_dof_free[d] - 1] = u_inv[d] (reconstruction.f:60-62); the free indices are distinct
__global__ void __launch_bounds__(256) recon_free_kernel(int32_t nfree, const double *u_inv, const int32_t *ind,
                                                         int64_t ndof, double *U) {
    for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < nfree; d += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = (int64_t)ind[d] - 1;
        if (g >= 0 && g < ndof) U[g] = u_inv[d];
    }
}

// one boundary condition: U[(target - 1) mcrd + dir - 1] = value (reconstruction.f:64-74)
__global__ void __launch_bounds__(256) recon_bc_kernel(int32_t n, const int32_t *targets, int32_t mcrd, int32_t dir,
                                                       double value, int64_t ndof, double *U) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = ((int64_t)targets[k] - 1) * mcrd + dir - 1;
        if (g >= 0 && g < ndof) U[g] = value;
    }
}

// sol(cp, i) = U(cp mcrd + i), sol column-major (reconstruction.f:131-135)
__global__ void __launch_bounds__(256) recon_sol_kernel(int32_t nb_cp, int32_t mcrd, const double *U, double *sol) {
    const int64_t n = (int64_t)nb_cp * mcrd;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / nb_cp, cp = t % nb_cp;
        sol[t] = U[cp * mcrd + i];
    }
}

static unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256)); }

constexpr int EVAL_MAXP = 8;

// BINARYSEARCH of igaparmetrization.f90:291-319 over uk[0..n-1] (1-based result, intervals
// [left, right); a value on no interval ends with the last midpoint, as there)
__device__ inline int eval_search(double val, const double *uk, int n) {
    int low = 1, high = n, middle = 1;
    bool found = false;
    while (low != high && !found) {
        middle = (low + high) / 2;
        if (val >= uk[middle - 1] && val < uk[middle]) found = true;
        else if (val < uk[middle - 1]) high = middle;
        else low = middle + 1;
    }
    return middle;
}

// B-spline values N_{span-p..span} at u (the triangular recurrence of dersbasisfuns.f90:30-131,
// values only); span1 is the 1-based span index
__device__ inline void eval_basis(int span1, int p, double u, const double *U, double *N) {
    const int i = span1 - 1;
    double left[EVAL_MAXP + 1], right[EVAL_MAXP + 1];
    N[0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[i + 1 - j];
        right[j] = U[i + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            const double tmp = N[r] / (right[r + 1] + left[j - r]);
            N[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        N[j] = saved;
    }
}

struct EvalPatch {
    int p[3], nkv[3], nint[3];
    const double *U[3];
    const int32_t *nijk, *ien, *table;   // table: interval triple -> element (0-based)
    const double *weight;
    int nnode;
};

// one thread per point (evaluate_disp.f:83-97 with evalnurbs, evaluatenurbsfcts.f:94-201)
__global__ void __launch_bounds__(256) evaldisp_kernel(EvalPatch ep, int32_t mcrd, const double *sol, int64_t npts,
                                                       const double *xi, double *disp) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npts; t += (int64_t)gridDim.x * blockDim.x) {
        int iv[3];
        for (int d = 0; d < 3; ++d) {
            const int idx = eval_search(xi[3 * t + d], ep.U[d] + ep.p[d], ep.nkv[d] - 2 * ep.p[d]);   // span - p
            iv[d] = min(max(idx - 1, 0), ep.nint[d] - 1);
        }
        const int e = ep.table[iv[0] + ep.nint[0] * (iv[1] + ep.nint[1] * iv[2])];
        double B[3][EVAL_MAXP + 1];
        for (int d = 0; d < 3; ++d) eval_basis(ep.nijk[3 * e + d], ep.p[d], xi[3 * t + d], ep.U[d], B[d]);
        double sum = 0.0;
        int a = 0;
        for (int k = 0; k <= ep.p[2]; ++k)
            for (int j = 0; j <= ep.p[1]; ++j)
                for (int i = 0; i <= ep.p[0]; ++i, ++a)
                    sum += B[0][ep.p[0] - i] * B[1][ep.p[1] - j] * B[2][ep.p[2] - k] * ep.weight[(int64_t)ep.nnode * e + a];
        const double inv = 1.0 / sum;
        double u[3] = {0.0, 0.0, 0.0};
        a = 0;
        for (int k = 0; k <= ep.p[2]; ++k)
            for (int j = 0; j <= ep.p[1]; ++j)
                for (int i = 0; i <= ep.p[0]; ++i, ++a) {
                    const double r = B[0][ep.p[0] - i] * B[1][ep.p[1] - j] * B[2][ep.p[2] - k] *
                                     ep.weight[(int64_t)ep.nnode * e + a] * inv;
                    const int64_t cp = (int64_t)ep.ien[(int64_t)ep.nnode * e + a] - 1;
                    for (int c = 0; c < mcrd && c < 3; ++c) u[c] += r * sol[cp * mcrd + c];
                }
        for (int c = 0; c < 3; ++c) disp[3 * t + c] = u[c];
    }
}

// B-spline values and first derivatives at u, runtime degree p <= EVAL_MAXP: the triangular
// recurrence of eval_basis, keeping the degree p - 1 values for the derivatives
// (dersbasisfuns.f90:30-131, first derivative only)
__device__ inline void eval_ders(int span1, int p, double u, const double *U, double *N, double *dN) {
    const int i = span1 - 1;
    double left[EVAL_MAXP + 1], right[EVAL_MAXP + 1], low[EVAL_MAXP + 1];
    N[0] = 1.0;
    low[0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[i + 1 - j];
        right[j] = U[i + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            const double tmp = N[r] / (right[r + 1] + left[j - r]);
            N[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        N[j] = saved;
        if (j == p - 1)
            for (int r = 0; r < p; ++r) low[r] = N[r];
    }
    // dN_k/du = p N_{k,p-1} / (U[k+p] - U[k]) - p N_{k+1,p-1} / (U[k+p+1] - U[k+1]), k = i - p + r
    for (int r = 0; r <= p; ++r) {
        const int k = i - p + r;
        double d = 0.0;
        if (r >= 1) {
            const double den = U[k + p] - U[k];
            if (den != 0.0) d += p * low[r - 1] / den;
        }
        if (r <= p - 1) {
            const double den = U[k + p + 1] - U[k + 1];
            if (den != 0.0) d -= p * low[r] / den;
        }
        dN[r] = d;
    }
}

struct SvarsPatch {
    int p[3];
    const double *U[3];
    const int32_t *nijk, *ien;
    const double *weight;
    int nnode, nelem;
};

// Fields at the vertex grid of every element of a solid patch (compute_svars_solid.f:24-188):
// one workgroup per element with its control points, displacements and weights in LDS, one
// thread per vertex (parent coordinate 2/(nb-1) i - 1, xi fastest). W, the point, the
// displacement and their parametric derivatives come from one pass over the nodes (the quotient
// rule of shap.f90:122-129 applied to the sums), then dx/dxi is inverted, du/dx = du/dxi dxi/dx,
// the strain has engineering shears (12, 13, 23) and the stress is the isotropic D times it.
// Output per vertex: point(3), [u(3)], [stress(6)], [strain(6)], element after element.
__global__ void __launch_bounds__(128) svars_kernel(SvarsPatch sp, const double *coords, const double *sol, int nx,
                                                    int ny, int nz, int fdisp, int fstress, int fstrain, int nsv,
                                                    double lam, double mu, double *svars) {
    extern __shared__ double svsh[];
    const int nn = sp.nnode, nv = nx * ny * nz;
    double *xs = svsh, *us = svsh + 3 * nn, *ws = svsh + 6 * nn;
    const int p0 = sp.p[0], p1 = sp.p[1], p2 = sp.p[2];
    for (int e = blockIdx.x; e < sp.nelem; e += gridDim.x) {
        __syncthreads();
        for (int a = threadIdx.x; a < nn; a += blockDim.x) {
            const int64_t cp = (int64_t)sp.ien[(int64_t)nn * e + a] - 1;
            for (int c = 0; c < 3; ++c) {
                xs[3 * a + c] = coords[3 * cp + c];
                us[3 * a + c] = sol[3 * cp + c];
            }
            ws[a] = sp.weight[(int64_t)nn * e + a];
        }
        __syncthreads();
        for (int v = threadIdx.x; v < nv; v += blockDim.x) {
            const int iv[3] = {v % nx, (v / nx) % ny, v / (nx * ny)};
            const int nb[3] = {nx, ny, nz};
            double B[3][EVAL_MAXP + 1], dB[3][EVAL_MAXP + 1];
            for (int d = 0; d < 3; ++d) {
                const double pt = 2.0 / (double)(nb[d] - 1) * (double)iv[d] - 1.0;
                const int span1 = sp.nijk[3 * (int64_t)e + d];
                const double u0 = sp.U[d][span1 - 1], u1 = sp.U[d][span1];
                const double xi = ((u1 - u0) * pt + (u1 + u0)) * 0.5;
                eval_ders(span1, sp.p[d], xi, sp.U[d], B[d], dB[d]);
            }
            double W = 0.0, dW[3] = {0.0, 0.0, 0.0}, SX[3] = {0.0, 0.0, 0.0}, SU[3] = {0.0, 0.0, 0.0};
            double SXd[3][3] = {}, SUd[3][3] = {};
            int a = 0;
            for (int k = 0; k <= p2; ++k)
                for (int j = 0; j <= p1; ++j)
                    for (int i = 0; i <= p0; ++i, ++a) {
                        const double b0 = B[0][p0 - i], b1 = B[1][p1 - j], b2 = B[2][p2 - k], w = ws[a];
                        const double r = b0 * b1 * b2 * w;
                        const double g[3] = {dB[0][p0 - i] * b1 * b2 * w, b0 * dB[1][p1 - j] * b2 * w,
                                             b0 * b1 * dB[2][p2 - k] * w};
                        W += r;
                        for (int c = 0; c < 3; ++c) {
                            const double x = xs[3 * a + c], u = us[3 * a + c];
                            dW[c] += g[c];
                            SX[c] += r * x;
                            SU[c] += r * u;
                            for (int b = 0; b < 3; ++b) {
                                SXd[c][b] += g[b] * x;
                                SUd[c][b] += g[b] * u;
                            }
                        }
                    }
            const double Wi = 1.0 / W;
            double J[3][3], Gu[3][3];   // dx_c/dxi_b, du_c/dxi_b
            for (int c = 0; c < 3; ++c)
                for (int b = 0; b < 3; ++b) {
                    J[c][b] = (SXd[c][b] - SX[c] * dW[b] * Wi) * Wi;
                    Gu[c][b] = (SUd[c][b] - SU[c] * dW[b] * Wi) * Wi;
                }
            const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            const double di = 1.0 / det;
            double Ji[3][3];   // dxi_b/dx_m
            Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * di;
            Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * di;
            Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * di;
            Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * di;
            Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * di;
            Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * di;
            Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * di;
            Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * di;
            Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * di;
            double du[3][3];   // du_c/dx_m
            for (int c = 0; c < 3; ++c)
                for (int m = 0; m < 3; ++m) du[c][m] = Gu[c][0] * Ji[0][m] + Gu[c][1] * Ji[1][m] + Gu[c][2] * Ji[2][m];
            const double eps[6] = {du[0][0], du[1][1], du[2][2], du[0][1] + du[1][0], du[0][2] + du[2][0],
                                   du[2][1] + du[1][2]};
            const double tr = eps[0] + eps[1] + eps[2];
            double *o = svars + ((int64_t)e * nv + v) * nsv;
            for (int c = 0; c < 3; ++c) *o++ = SX[c] * Wi;
            if (fdisp)
                for (int c = 0; c < 3; ++c) *o++ = SU[c] * Wi;
            if (fstress) {
                for (int c = 0; c < 3; ++c) *o++ = 2.0 * mu * eps[c] + lam * tr;
                for (int c = 3; c < 6; ++c) *o++ = mu * eps[c];
            }
            if (fstrain)
                for (int c = 0; c < 6; ++c) *o++ = eps[c];
        }
    }
}

// Vertices per direction and per element, cells per element, values per vertex
// (postproc_vtu.f:93-96, 136-151)
struct VtuCounts {
    int nb[3];
    int64_t nv, ncell;
    int nsv;
};
static VtuCounts vtu_counts(const int32_t *nb_ref, const int32_t *flags) {
    VtuCounts c{};
    c.nv = 1;
    c.ncell = 1;
    for (int d = 0; d < 3; ++d) {
        const int s = 1 << std::max(nb_ref[d] - 1, 0);
        c.nb[d] = s + 1;
        c.nv *= s + 1;
        c.ncell *= s;
    }
    c.nsv = 3 + (flags[0] ? 3 : 0) + (flags[1] ? 6 : 0) + (flags[2] ? 6 : 0);
    return c;
}

// Buffered text output of a VTU piece: numbers in the shortest form that reads back exactly
struct TextOut {
    FILE *f;
    std::string buf;
    explicit TextOut(FILE *f_) : f(f_) { buf.reserve(1 << 20); }
    void flush() {
        if (!buf.empty()) fwrite(buf.data(), 1, buf.size(), f);
        buf.clear();
    }
    void line(const char *s) {
        buf += s;
        buf += '\n';
        if (buf.size() > (1u << 20)) flush();
    }
    void num(double v) {
        char t[32];
        auto r = std::to_chars(t, t + sizeof(t), v);
        buf += ' ';
        buf.append(t, r.ptr);
    }
    void num(int64_t v) {
        char t[24];
        auto r = std::to_chars(t, t + sizeof(t), v);
        buf += ' ';
        buf.append(t, r.ptr);
    }
    void end_line() {
        buf += '\n';
        if (buf.size() > (1u << 20)) flush();
    }
};

}  // namespace yeti

using namespace yeti;

extern "C" int32_t yeti_reconstruction(const double *u_inv, const int32_t *ind_dof_free, int32_t nb_dof_free,
                                       const int32_t *bc_target, const int32_t *bc_target_nbelem_host,
                                       const double *bc_values_host, int32_t nb_bc, int32_t mcrd, int32_t nb_cp,
                                       double *sol, double *U, void *stream) {
    if (mcrd < 1 || nb_cp < 0 || nb_dof_free < 0 || nb_bc < 0 || !sol || !U)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "reconstruction: bad sizes or null output");
    if ((nb_dof_free > 0 && (!u_inv || !ind_dof_free)) || (nb_bc > 0 && (!bc_target_nbelem_host || !bc_values_host)))
        return set_error(YETI_ERR_INVALID_ARGUMENT, "reconstruction: null input");
    hipStream_t st = (hipStream_t)stream;
    const int64_t ndof = (int64_t)nb_cp * mcrd;
    if (ndof == 0) return 0;
    hipLaunchKernelGGL(recon_fill_kernel, dim3(grid_for(ndof)), dim3(256), 0, st, ndof, U);
    if (nb_dof_free > 0)
        hipLaunchKernelGGL(recon_free_kernel, dim3(grid_for(nb_dof_free)), dim3(256), 0, st, nb_dof_free, u_inv,
                           ind_dof_free, ndof, U);
    int64_t count = 0;
    for (int32_t b = 0; b < nb_bc; ++b) {
        const int32_t n = bc_target_nbelem_host[b];
        if (n < 0) return set_error(YETI_ERR_INVALID_ARGUMENT, "reconstruction: negative target count");
        if (bc_values_host[2 * b] <= 3.0 && n > 0) {
            if (!bc_target) return set_error(YETI_ERR_INVALID_ARGUMENT, "reconstruction: null bc_target");
            hipLaunchKernelGGL(recon_bc_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, bc_target + count, mcrd,
                               (int32_t)bc_values_host[2 * b], bc_values_host[2 * b + 1], ndof, U);
        }
        count += n;
    }
    hipLaunchKernelGGL(recon_sol_kernel, dim3(grid_for(ndof)), dim3(256), 0, st, nb_cp, mcrd, (const double *)U, sol);
    YETI_HIP(hipGetLastError());
    return 0;
}

extern "C" int32_t yeti_evaldisp(const yeti_nurbs_patch *patch, int32_t nb_cp, int32_t mcrd, const double *sol,
                                 int64_t npts, const double *xi, double *disp) {
    if (!patch || !sol || (npts > 0 && (!xi || !disp)) || mcrd < 1 || nb_cp < 1 || npts < 0)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "evaldisp: bad arguments");
    const yeti_nurbs_patch &P = *patch;
    if (P.nelem < 1 || !P.nijk || !P.ien || !P.weight)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "evaldisp: empty patch");
    int nn = 1;
    EvalPatch ep{};
    for (int d = 0; d < 3; ++d) {
        if (P.p[d] < 0 || P.p[d] > EVAL_MAXP || !P.U[d] || P.nkv[d] < 2 * P.p[d] + 2)
            return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "evaldisp: degree %d / %d knots in direction %d", P.p[d],
                             P.nkv[d], d + 1);
        ep.p[d] = P.p[d];
        ep.nkv[d] = P.nkv[d];
        ep.nint[d] = P.nkv[d] - 2 * P.p[d] - 1;
        nn *= P.p[d] + 1;
    }
    if (nn != P.nnode) return set_error(YETI_ERR_INVALID_ARGUMENT, "evaldisp: nnode %d, degrees give %d", P.nnode, nn);
    // interval triple -> element: the unique element with those spans, else element 1
    // (igaparmetrization.f90:357-368)
    const int64_t ntab = (int64_t)ep.nint[0] * ep.nint[1] * ep.nint[2];
    std::vector<int32_t> table(ntab, -1), count(ntab, 0);
    for (int32_t e = 0; e < P.nelem; ++e) {
        int iv[3];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {
            iv[d] = P.nijk[3 * (int64_t)e + d] - P.p[d] - 1;
            ok &= iv[d] >= 0 && iv[d] < ep.nint[d];
        }
        if (!ok) continue;
        const int64_t k = iv[0] + (int64_t)ep.nint[0] * (iv[1] + (int64_t)ep.nint[1] * iv[2]);
        if (count[k]++ == 0) table[k] = e;
    }
    for (int64_t k = 0; k < ntab; ++k)
        if (count[k] != 1) table[k] = 0;
    for (int32_t e = 0; e < P.nelem; ++e)
        for (int a = 0; a < P.nnode; ++a) {
            const int32_t cp = P.ien[(int64_t)P.nnode * e + a];
            if (cp < 1 || cp > nb_cp) return set_error(YETI_ERR_INVALID_ARGUMENT, "evaldisp: IEN entry out of range");
        }
    if (npts == 0) return 0;
    DevBuf<double> dU[3], dw, dsol, dxi, ddisp;
    DevBuf<int32_t> dnijk, dien, dtab;
    hipError_t er = hipSuccess;
    for (int d = 0; d < 3 && er == hipSuccess; ++d) er = dU[d].upload(P.U[d], P.nkv[d]);
    if (er == hipSuccess) er = dnijk.upload(P.nijk, 3 * (size_t)P.nelem);
    if (er == hipSuccess) er = dien.upload(P.ien, (size_t)P.nnode * P.nelem);
    if (er == hipSuccess) er = dw.upload(P.weight, (size_t)P.nnode * P.nelem);
    if (er == hipSuccess) er = dtab.upload(table.data(), table.size());
    if (er == hipSuccess) er = dsol.upload(sol, (size_t)mcrd * nb_cp);
    if (er == hipSuccess) er = dxi.upload(xi, 3 * (size_t)npts);
    if (er == hipSuccess) er = ddisp.alloc(3 * (size_t)npts);
    if (er != hipSuccess) return set_error(YETI_ERR_DEVICE, "evaldisp: device memory: %s", hipGetErrorString(er));
    for (int d = 0; d < 3; ++d) ep.U[d] = dU[d].p;
    ep.nijk = dnijk.p;
    ep.ien = dien.p;
    ep.weight = dw.p;
    ep.table = dtab.p;
    ep.nnode = P.nnode;
    hipLaunchKernelGGL(evaldisp_kernel, dim3(grid_for(npts)), dim3(256), 0, 0, ep, mcrd, (const double *)dsol.p, npts,
                       (const double *)dxi.p, ddisp.p);
    YETI_HIP(hipGetLastError());
    YETI_HIP(hipMemcpy(disp, ddisp.p, 3 * (size_t)npts * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int32_t yeti_svars_solid(const yeti_nurbs_patch *patch, int32_t nb_cp, const double *coords3d,
                                    const double *sol, const int32_t *nb_ref, const int32_t *flags, double E, double nu,
                                    double *svars) {
    if (!patch || !coords3d || !sol || !nb_ref || !flags || !svars || nb_cp < 1)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "svars: bad arguments");
    const yeti_nurbs_patch &P = *patch;
    if (P.nelem < 1 || !P.nijk || !P.ien || !P.weight)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "svars: empty patch");
    int nn = 1;
    SvarsPatch sp{};
    for (int d = 0; d < 3; ++d) {
        if (P.p[d] < 1 || P.p[d] > EVAL_MAXP || !P.U[d] || P.nkv[d] < 2 * P.p[d] + 2)
            return set_error(YETI_ERR_UNSUPPORTED_DEGREE, "svars: degree %d / %d knots in direction %d", P.p[d],
                             P.nkv[d], d + 1);
        if (nb_ref[d] < 1 || nb_ref[d] > 8)
            return set_error(YETI_ERR_INVALID_ARGUMENT, "svars: refinement %d in direction %d (1..8)", nb_ref[d], d + 1);
        sp.p[d] = P.p[d];
        nn *= P.p[d] + 1;
    }
    if (nn != P.nnode) return set_error(YETI_ERR_INVALID_ARGUMENT, "svars: nnode %d, degrees give %d", P.nnode, nn);
    for (int32_t e = 0; e < P.nelem; ++e) {
        for (int d = 0; d < 3; ++d) {   // the spans the basis recurrence reads around
            const int32_t s1 = P.nijk[3 * (int64_t)e + d];
            if (s1 < P.p[d] + 1 || s1 > P.nkv[d] - P.p[d] - 1)
                return set_error(YETI_ERR_INVALID_ARGUMENT, "svars: span %d of element %d out of range", s1, e + 1);
        }
        for (int a = 0; a < P.nnode; ++a) {
            const int32_t cp = P.ien[(int64_t)P.nnode * e + a];
            if (cp < 1 || cp > nb_cp) return set_error(YETI_ERR_INVALID_ARGUMENT, "svars: IEN entry out of range");
        }
    }
    const VtuCounts vc = vtu_counts(nb_ref, flags);
    const int64_t nout = (int64_t)P.nelem * vc.nv * vc.nsv;
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));   // material_lib.f90:52-87
    DevBuf<double> dU[3], dw, dx, dsol, dout;
    DevBuf<int32_t> dnijk, dien;
    hipError_t er = hipSuccess;
    for (int d = 0; d < 3 && er == hipSuccess; ++d) er = dU[d].upload(P.U[d], P.nkv[d]);
    if (er == hipSuccess) er = dnijk.upload(P.nijk, 3 * (size_t)P.nelem);
    if (er == hipSuccess) er = dien.upload(P.ien, (size_t)P.nnode * P.nelem);
    if (er == hipSuccess) er = dw.upload(P.weight, (size_t)P.nnode * P.nelem);
    if (er == hipSuccess) er = dx.upload(coords3d, 3 * (size_t)nb_cp);
    if (er == hipSuccess) er = dsol.upload(sol, 3 * (size_t)nb_cp);
    if (er == hipSuccess) er = dout.alloc((size_t)nout);
    if (er != hipSuccess) return set_error(YETI_ERR_DEVICE, "svars: device memory: %s", hipGetErrorString(er));
    for (int d = 0; d < 3; ++d) sp.U[d] = dU[d].p;
    sp.nijk = dnijk.p;
    sp.ien = dien.p;
    sp.weight = dw.p;
    sp.nnode = P.nnode;
    sp.nelem = P.nelem;
    const unsigned grid = (unsigned)std::min<int64_t>(P.nelem, 1 << 20);
    hipLaunchKernelGGL(svars_kernel, dim3(grid), dim3(128), 7 * (size_t)P.nnode * sizeof(double), 0, sp,
                       (const double *)dx.p, (const double *)dsol.p, vc.nb[0], vc.nb[1], vc.nb[2], flags[0] ? 1 : 0,
                       flags[1] ? 1 : 0, flags[2] ? 1 : 0, vc.nsv, lam, mu, dout.p);
    YETI_HIP(hipGetLastError());
    YETI_HIP(hipMemcpy(svars, dout.p, (size_t)nout * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int32_t yeti_vtu_write_piece(const char *path, const double *svars, int32_t nb_elem, const int32_t *nb_ref,
                                        const int32_t *flags) {
    if (!path || !svars || !nb_ref || !flags || nb_elem < 0)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "vtu: bad arguments");
    for (int d = 0; d < 3; ++d)
        if (nb_ref[d] < 1 || nb_ref[d] > 8)
            return set_error(YETI_ERR_INVALID_ARGUMENT, "vtu: refinement %d in direction %d (1..8)", nb_ref[d], d + 1);
    FILE *f = fopen(path, "ab");
    if (!f) return set_error(YETI_ERR_INVALID_ARGUMENT, "vtu: cannot open %s", path);
    const VtuCounts vc = vtu_counts(nb_ref, flags);
    const int nx = vc.nb[0], ny = vc.nb[1], nz = vc.nb[2], nsv = vc.nsv;
    const int64_t nv = vc.nv, ncells = (int64_t)nb_elem * vc.ncell;
    TextOut o(f);
    char hdr[160];
    snprintf(hdr, sizeof(hdr), " <Piece  NumberOfPoints=\"%lld\"  NumberOfCells=\"%lld\">",
             (long long)(nb_elem * nv), (long long)ncells);
    o.line(hdr);
    // one 3- or 6-vector per vertex from column c0 of svars (writepatch.f:56-66, 136-176, 212-254)
    auto vectors = [&](int c0, int nc) {
        for (int64_t e = 0; e < nb_elem; ++e)
            for (int64_t v = 0; v < nv; ++v) {
                const double *s = svars + (e * nv + v) * nsv + c0;
                for (int c = 0; c < nc; ++c) o.num(s[c]);
                o.end_line();
            }
    };
    o.line(" <Points>");
    o.line(" <DataArray  type=\"Float64\" NumberOfComponents=\"3\"  format=\"ascii\" >");
    vectors(0, 3);
    o.line(" </DataArray>");
    o.line(" </Points>");
    o.line(" <Cells>");
    o.line(" <DataArray  type=\"Int32\"  Name=\"connectivity\"  format=\"ascii\">");
    for (int64_t e = 0; e < nb_elem; ++e)   // writepatch.f:84-98
        for (int k = 0; k < nz - 1; ++k)
            for (int j = 0; j < ny - 1; ++j)
                for (int i = 0; i < nx - 1; ++i) {
                    const int64_t c = e * nv + (int64_t)k * nx * ny + (int64_t)j * nx + i, t = c + (int64_t)nx * ny;
                    const int64_t q[8] = {c, c + 1, c + nx + 1, c + nx, t, t + 1, t + nx + 1, t + nx};
                    for (int m = 0; m < 8; ++m) o.num(q[m]);
                    o.end_line();
                }
    o.line(" </DataArray>");
    o.line(" <DataArray  type=\"Int32\"  Name=\"offsets\"  format=\"ascii\"> ");
    for (int64_t c = 1; c <= ncells; ++c) {
        o.num(8 * c);
        o.end_line();
    }
    o.line(" </DataArray>");
    o.line(" <DataArray  type=\"UInt8\"  Name=\"types\"  format=\"ascii\">");
    for (int64_t c = 0; c < ncells; ++c) o.line(" 12");
    o.line(" </DataArray>");
    o.line(" </Cells> ");
    o.line(" <PointData>");
    int col = 3;
    if (flags[0]) {
        o.line(" <DataArray  type=\"Float64\" Name=\"disp\" NumberOfComponents=\"3\" format=\"ascii\">");
        vectors(col, 3);
        o.line(" </DataArray>");
        col += 3;
    }
    if (flags[1]) {
        o.line(" <DataArray  type=\"Float64\" Name=\"stress\" NumberOfComponents=\"6\" format=\"ascii\">");
        vectors(col, 6);
        o.line(" </DataArray>");
        // von Mises, sqrt(((s1-s2)^2 + (s1-s3)^2 + (s2-s3)^2 + 6 (s4^2 + s5^2 + s6^2)) / 2)
        // (writepatch.f:180-201)
        o.line(" <DataArray  type=\"Float64\" Name=\"VM\" NumberOfComponents=\"1\" format=\"ascii\">");
        for (int64_t e = 0; e < nb_elem; ++e)
            for (int64_t v = 0; v < nv; ++v) {
                const double *s = svars + (e * nv + v) * nsv + col;
                const double vm = (s[0] - s[1]) * (s[0] - s[1]) + (s[0] - s[2]) * (s[0] - s[2]) +
                                  (s[1] - s[2]) * (s[1] - s[2]) + 6.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
                o.num(std::sqrt(0.5 * vm));
                o.end_line();
            }
        o.line(" </DataArray>");
        col += 6;
    }
    if (flags[2]) {
        o.line(" <DataArray  type=\"Float64\" Name=\"strain\" NumberOfComponents=\"6\" format=\"ascii\">");
        vectors(col, 6);
        o.line(" </DataArray>");
    }
    o.line(" </PointData>");
    o.line(" </Piece>");
    o.flush();
    const bool bad = ferror(f) != 0;
    if (fclose(f) != 0 || bad) return set_error(YETI_ERR_INVALID_ARGUMENT, "vtu: write to %s failed", path);
    return 0;
}
