/*
 * yeti_amd.h — C ABI of libyeti_hip.so, the MI355X (gfx950) implementation of YETI's
 * 3D NURBS solid (element type U1) linear-elasticity hot path.
 *
 * Plain pointers and sizes only. Array layouts are exactly the ones YETI's f2py routines
 * receive (Fortran column-major, 1-based control-point / span / element ids as given).
 * Every function returns 0 on success or one of the YETI_ERR_* codes below; the message of
 * the last error on the calling thread is available from yeti_last_error().
 *
 * Reference interfaces replaced (paths relative to the YETI source tree):
 *   yeti_sys_linmat_lindef_static   <- stiffmtrx_elemstorage.sys_linmat_lindef_static
 *                                      src/yeti_iga/main/build_system_elemStorage.f90:24-315
 *   (layout YETI_LAYOUT_SLOTTED)    <- stiffmtrx_elemstorage_omp.sys_linmat_lindef_static_omp
 *                                      src/yeti_iga/main/build_system_elemStorage_OMP.f90:26-289
 *   yeti_asm_*                      <- the same routines, split into set-up (geometry resident in
 *                                      HBM) and run (device-resident outputs)
 *   yeti_csr_*                      <- scipy COO->CSC + (K + K^T) done by the callers,
 *                                      src/yeti_iga/overlay/iga_model.py:391-395
 *   yeti_pcg_*                      <- scipy spsolve / the matrix-free CG iga_mf_cg_3d,
 *                                      src/yeti_iga/iga_wq_mf/fsrc/mf_iga_solver.f90:642-958
 *   yeti_mf_*                       <- mf_iga_get_ku_3d, src/yeti_iga/iga_wq_mf/fsrc/mf_iga_solver.f90:292-555
 *   yeti_grad_lin_elast_work        <- optim.gradadjointwork.gradlinelastwork_an,
 *                                      src/yeti_iga/optim/gradtotalwork_fullAN.f90:17-300
 */
#ifndef YETI_AMD_H
#define YETI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------------------- */
#define YETI_OK                      0
#define YETI_ERR_INVALID_ARGUMENT    1  /* inconsistent sizes, null pointers, bad layout id  */
#define YETI_ERR_UNSUPPORTED_ELEMENT 2  /* element type other than U1 / U0                    */
#define YETI_ERR_UNSUPPORTED_DEGREE  3  /* degree / Gauss order with no compiled kernel       */
#define YETI_ERR_SIZE_OVERFLOW       4  /* nb_data too small, or a dof id beyond int32        */
#define YETI_ERR_DEVICE              5  /* HIP runtime error (includes out of memory)         */
#define YETI_ERR_NOT_CONVERGED       6  /* iterative solve hit its iteration cap              */
#define YETI_ERR_BUDGET              7  /* request rejected by the HBM admission budget       */

/* ---- COO layouts ---------------------------------------------------------------------- */
/* PACKED: inactive elements take no COO slots (build_system_elemStorage.f90:156,276).
 * SLOTTED: element e of a patch owns slots [count + e*nb_data_elem, ...), inactive slots stay 0
 *          (build_system_elemStorage_OMP.f90:229). */
#define YETI_LAYOUT_PACKED  0
#define YETI_LAYOUT_SLOTTED 1

/* All inputs of sys_linmat_lindef_static, in the reference's argument order
 * (build_system_elemStorage.f90:24-33; marshalled by IGA_parametrization.py:1705-1716).
 * Fortran shapes are noted as (rows, cols); storage is column-major. */
typedef struct yeti_system_inputs {
    const int32_t *activeElement;        /* [nb_elem] 1 = assemble this element              */
    int64_t        nb_data;              /* length of Kdata/Krow/Kcol                        */
    const double  *COORDS3D;             /* (3, nb_cp)                                       */
    const int32_t *IEN;                  /* concatenated per-patch (nnode, nb_elem_patch)    */
    int64_t        len_IEN;
    const int32_t *nb_elem_patch;        /* [nb_patch]                                       */
    const int32_t *Nkv;                  /* (3, nb_patch) knot-vector lengths                */
    const double  *Ukv;                  /* concatenated knot vectors                        */
    int64_t        len_Ukv;
    const int32_t *Nijk;                 /* (3, nb_elem) 1-based knot-span ids               */
    const double  *weight;               /* concatenated per-element CP weights              */
    int64_t        len_weight;
    const int32_t *Jpqr;                 /* (3, nb_patch) degrees                            */
    const char    *ELT_TYPE;             /* "U1U1..." (not NUL-terminated necessarily)       */
    int64_t        len_ELT_TYPE;
    const double  *PROPS;
    int64_t        len_PROPS;
    const int32_t *JPROPS;               /* [nb_patch]                                       */
    const double  *MATERIAL_PROPERTIES;  /* (n_mat_props_max, nb_patch): E, nu, ...          */
    const int32_t *n_mat_props;          /* [nb_patch]                                       */
    const double  *RHO;                  /* [nb_patch]                                       */
    const char    *TENSOR;               /* "/THREED/THREED..."                              */
    int64_t        len_TENSOR;
    const int32_t *indDLoad;             /* concatenated load targets (elements or CPs)     */
    int64_t        len_indDLoad;
    const int32_t *JDLType;              /* [nb_load]                                        */
    const double  *ADLMAG;               /* [nb_load]                                        */
    const int32_t *load_target_nbelem;   /* [nb_load]                                        */
    const double  *load_additionalInfos;
    int64_t        len_load_additionalInfos;
    const int32_t *nb_load_additionalInfos; /* [nb_load]                                     */
    const double  *bc_values;            /* (2, nb_bc)                                       */
    const int32_t *bc_target;
    int64_t        len_bc_target;
    const int32_t *bc_target_nbelem;     /* [nb_bc]                                          */
    const int32_t *ind_dof_free;         /* [nb_dof_tot] 1-based, zero padded                */
    int32_t        nb_dof_free;
    int32_t        MCRD;                 /* must be 3                                        */
    const int32_t *NBINT;                /* [nb_patch] Gauss points per element              */
    const int32_t *NNODE;                /* [nb_patch] CPs per element                       */
    const double  *nodal_dist;           /* (nb_n_dist, nb_cp_n_dist)                        */
    int32_t        nb_bc;
    int32_t        nb_load;
    int32_t        nb_patch;
    int32_t        nb_elem;
    int32_t        nb_cp;
    int32_t        nb_dof_tot;
    int32_t        nb_n_dist;
    int32_t        nb_cp_n_dist;
    int32_t        n_mat_props_max;
} yeti_system_inputs;

/* Last error message of the calling thread ("" when none). */
const char *yeti_last_error(void);

/* Library/ABI version, e.g. 0x000100 for 0.1.0. */
int32_t yeti_version(void);

/* Number of HIP devices visible (0 when none); never fails. */
int32_t yeti_device_count(void);

/* ---- drop-in: host in, host out --------------------------------------------------------
 * Same contract as sys_linmat_lindef_static (build_system_elemStorage.f90:24-315): fills
 * Kdata/Krow/Kcol (nb_data entries, upper-triangle element blocks with halved diagonals,
 * 0-based dof ids) and F (nb_dof_tot). All pointers are host pointers. layout selects the
 * serial (PACKED) or OpenMP (SLOTTED, build_system_elemStorage_OMP.f90:229) slot rule. */
int32_t yeti_sys_linmat_lindef_static(const yeti_system_inputs *in, int32_t layout,
                                      double *Kdata, int32_t *Krow, int32_t *Kcol, double *F);

/* ---- device-resident assembly ----------------------------------------------------------
 * Set-up copies the geometry/load description to the current HIP device once; run writes the
 * COO triplets and F into caller-owned device buffers on `stream` (a hipStream_t, 0 = null
 * stream). The handle may be run any number of times; it is not thread-safe. */
typedef struct yeti_asm yeti_asm;

int32_t yeti_asm_create(const yeti_system_inputs *in, int32_t layout, yeti_asm **out);
/* Number of COO entries the layout needs (>= what Kdata must hold). */
int64_t yeti_asm_nb_data(const yeti_asm *h);
int32_t yeti_asm_run_coo(yeti_asm *h, double *Kdata_dev, int32_t *Krow_dev, int32_t *Kcol_dev,
                         double *F_dev, void *stream);
/* Matrix only: Kdata + Krow + Kcol, no right-hand side. */
int32_t yeti_asm_run_coo_matrix(yeti_asm *h, double *Kdata_dev, int32_t *Krow_dev, int32_t *Kcol_dev,
                                void *stream);
/* Values only (indices are a pure function of IEN and can be emitted once). */
int32_t yeti_asm_run_coo_values(yeti_asm *h, double *Kdata_dev, void *stream);
/* Consistent mass matrix of the resident model (f2py massmtrx.build_cmassmatrix,
 * src/yeti_iga/main/compute_CMASSMatrix.f:22-215 with UMASSMAT_byCP, UMASSMAT.f:220-330, solid
 * elements): M_ab[k][k] = sum_q RHO w_q DetJac_q R_a R_b, k != l entries 0, emitted in the same
 * COO layout and at the same element offsets as the stiffness (nb_data entries). Mrow / Mcol may
 * both be null (values only). Device pointers. */
int32_t yeti_asm_run_mass_coo(yeti_asm *h, double *Mdata, int32_t *Mrow, int32_t *Mcol, void *stream);
/* Natural-frequency shape gradient: gradVibration_AN, optim/gradvibration_fullAN.f:22-189 with
 * gradUELMASSMAT1 (optim/gradUELMASSMAT_solid.f:22-332), f2py vibration.gradvibration_an (caller
 * OPT_modelling.py:823-854). vect (device) holds nb_frq eigenvectors as (nb_frq, 3, nb_cp)
 * column-major, vals (host) their eigenvalues; gradV (device, same layout) receives
 * d lambda_f / d P over the active U1 elements of the model, divided per mode by the mass
 * u_f^T M u_f of those elements. */
int32_t yeti_asm_run_vib_grad(yeti_asm *h, const double *vect, const double *vals, int32_t nb_frq, double *gradV,
                              void *stream);
/* Strain-energy shape derivative on the resident model: g = -d(u^T K u)/dX (3 per control
 * point, COORDS3D order) over its active solid elements, for any degree triple the stiffness
 * accepts; the internal-work term of gradLinElastWork_AN (optim/gradtotalwork_fullAN.f90:17-300,
 * gradUELMAT1adj) for SOL = ADJ = u, in Cartesian form. u and g are device vectors of 3 nb_cp
 * values (dof-interleaved). */
int32_t yeti_asm_strain_energy_grad(yeti_asm *h, const double *u, double *g, void *stream);
/* Volume of the active solid elements and its shape derivative dV/dX (3 per control point,
 * COORDS3D order), for any degree triple the stiffness accepts: f2py volume.computevolume /
 * computegradvolume (optim/volume.f90:19-210, optim/gradvolume.f90:16-355; callers
 * OPT_modelling.py:508-530, all patches active). vol is a device pointer to one double, g a
 * device vector of 3 nb_cp values. */
int32_t yeti_asm_volume_grad(yeti_asm *h, double *vol, double *g, void *stream);
/* Work gradients of gradLinElastWork_AN (f2py optim.gradadjointwork.gradlinelastwork_an,
 * src/yeti_iga/optim/gradtotalwork_fullAN.f90:17-300 with gradUELMAT1adj,
 * optim/gradUELMAT_solid.f90:18-479; caller OPT_modelling.py:447-481) over the active solid
 * elements of the resident model, for a state SOL (3, nb_cp) and nadj adjoint fields ADJ
 * (nadj, 3, nb_cp), both column-major device arrays:
 *   gradWint (nadj, 3, nb_cp) = -ADJ_k^T (dK/dX) SOL   (computeWint != 0, else zeros)
 *   gradWext (nadj, 3, nb_cp) =  ADJ_k^T (dF/dX)       (computeWext != 0, else zeros)
 * with the external term restricted to what the reference computes: the centrifugal body loads
 * (type 101), whose axis points the k-th such load reads at values 6k+1..6k+6 of
 * load_additionalInfos as the reference does; its surface-load term is a stub (zero) there and
 * here. activeDir plays no part (the reference's solid routine does not read it). Synchronises
 * `stream` once (the polarisation scales). */
int32_t yeti_grad_lin_elast_work(yeti_asm *h, const double *SOL, const double *ADJ, int32_t nadj, int32_t computeWint,
                                 int32_t computeWext, double *gradWint, double *gradWext, void *stream);
/* Right-hand side only (concentrated, surface and body loads). */
int32_t yeti_asm_run_rhs(yeti_asm *h, double *F_dev, void *stream);
/* Replace the control-point coordinates (device, 3 nb_cp values, COORDS3D order) used by later
 * runs; the design-update step of the reference's shape optimisation
 * (OPT_modelling.py:268-312 re-marshals COORDS3D before each assembly). */
int32_t yeti_asm_update_coords(yeti_asm *h, const double *coords_dev, void *stream);
/* g[3 I + j] = 2 u^T dF/dX_Ij: shape derivative of the load work for a displacement u (device,
 * 3 nb_cp). Face loads 10..99 (YETI's own gradient leaves this term as a stub,
 * optim/gradUELMAT_solid.f90:463-477) and body loads 101 / 102 (the centrifugal term of
 * gradUELMAT1adj, optim/gradUELMAT_solid.f90:369-455, plus the constant volume force). */
int32_t yeti_asm_load_shape_grad(yeti_asm *h, const double *u_dev, double *g_dev, void *stream);
/* Measurement: with timing on, each element-kernel launch of the next runs is bracketed by HIP
 * events on its stream; yeti_asm_kernel_times waits for them and returns the durations (ms) of
 * the last run's launches (up to cap) and their count. */
int32_t yeti_asm_set_timing(yeti_asm *h, int32_t on);
int32_t yeti_asm_kernel_times(yeti_asm *h, float *ms, int32_t cap, int32_t *n);
void    yeti_asm_destroy(yeti_asm *h);

/* ---- CSR of K_side + K_side^T (device) --------------------------------------------------
 * Replaces the callers' scipy step `coo_matrix((Kdata,(Krow,Kcol))).tocsc(); K + K.T`
 * (overlay/iga_model.py:391-395). The pattern is built once from the handle's connectivity;
 * rows are dofs 3 I + k, each row lists the dof columns of every CP sharing an element with
 * CP I, ascending (the structural pattern; exact zeros are kept). rowptr is int64. */
typedef struct yeti_csr yeti_csr;

int32_t yeti_csr_create(const yeti_asm *h, void *stream, yeti_csr **out);
int64_t yeti_csr_nnz(const yeti_csr *c);
int32_t yeti_csr_nrows(const yeti_csr *c);
const int64_t *yeti_csr_rowptr(const yeti_csr *c);   /* device, [nrows + 1] */
const int32_t *yeti_csr_colind(const yeti_csr *c);   /* device, [nnz]       */
/* Copy the pattern to host buffers (rowptr[nrows + 1], colind[nnz]), e.g. to build a scipy matrix. */
int32_t yeti_csr_pattern_to_host(const yeti_csr *c, int64_t *rowptr_host, int32_t *colind_host);
/* vals[nnz] = CSR values of K_side + K_side^T from the COO values the handle wrote. */
int32_t yeti_csr_values(const yeti_csr *c, const double *Kdata_dev, double *vals_dev, void *stream);
/* y = A x, walking the nodal 3x3 blocks (one wave per control point, each block column read once). */
int32_t yeti_csr_spmv(const yeti_csr *c, const double *vals_dev, const double *x_dev, double *y_dev,
                      void *stream);
/* y = |A| |x| (entrywise absolute values): eps ||(|A||x|)|| bounds the FP64 rounding of A x, the
 * floor below which no residual of this matrix can be resolved. */
int32_t yeti_csr_spmv_abs(const yeti_csr *c, const double *vals_dev, const double *x_dev, double *y_dev,
                          void *stream);
void    yeti_csr_destroy(yeti_csr *c);
/* scipy's canonical K_side + K_side^T: the structural CSR with every entry whose sum is exactly 0
 * removed (scipy's sparse add drops them; overlay/iga_model.py:391-395 returns that matrix).
 * Outputs (device): rowptr_out[3 nb_cp + 1], colind_out / vals_out with room for yeti_csr_nnz
 * entries; *nnz_out (host) receives the kept count. Order within rows is kept (ascending). */
int32_t yeti_csr_drop_zeros(const yeti_csr *c, const double *vals_dev, int64_t *rowptr_out, int32_t *colind_out,
                            double *vals_out, int64_t *nnz_out, void *stream);

/* ---- Jacobi-preconditioned CG, FP64, on a device CSR --------------------------------------
 * Solves A[free][:, free] x[free] = b[free] (free_mask[i] = 1 for free dofs; masked entries of x
 * are 0 on return), stopping at ||r||_2 <= rtol ||b||_2; the residual is read back every
 * check_every iterations. Replaces the callers' spsolve on the sliced matrix
 * (benchs/volume_force/test_volume_force_3d.py:67-79) and the CG loop of iga_mf_cg_3d
 * (iga_wq_mf/fsrc/mf_iga_solver.f90:642-958). Returns YETI_ERR_NOT_CONVERGED at maxit. */
int32_t yeti_pcg_csr(int64_t nrows, const int64_t *rowptr, const int32_t *colind, const double *vals,
                     const double *b, const uint8_t *free_mask, double *x, double rtol, int32_t maxit,
                     int32_t check_every, int32_t *iters, double *relres, void *stream);

/* Jacobi PCG on a device CSR handle (its nodal-block product; same iteration as yeti_pcg_csr).
 * norm 2: stop at ||r||_2 <= rtol ||b||_2; norm 0: stop at max|r| <= rtol max|b| over the free
 * dofs, the rule of iga_mf_cg_3d (maxval(abs(r)) / maxval(abs(b)),
 * iga_wq_mf/fsrc/mf_iga_solver.f90:779,935). relres is reported in the chosen norm. */
int32_t yeti_csr_pcg(const yeti_csr *c, const double *vals, const double *b, const uint8_t *free_mask, double *x,
                     double rtol, int32_t maxit, int32_t check_every, int32_t norm, int32_t *iters, double *relres,
                     void *stream);

/* Same solve as yeti_pcg_csr, allocating nothing: no counterpart in the reference. For callers
 * that cannot allocate device memory (tight solve loops, harnesses that forbid hipMalloc). The
 * work vectors live in a file-scope device array of the library (up to 1,048,576 rows, about
 * 42 MB per GPU, held until the process ends); larger systems return YETI_ERR_SIZE_OVERFLOW.
 * Every call on one GPU shares that array: calls that overlap on one GPU (other threads or
 * streams) race and must be serialised by the caller. */
int32_t yeti_pcg_csr_noalloc(int64_t nrows, const int64_t *rowptr, const int32_t *colind, const double *vals,
                             const double *b, const uint8_t *free_mask, double *x, double rtol, int32_t maxit,
                             int32_t check_every, int32_t *iters, double *relres, void *stream);

/* ---- matrix-free operator (sum-factorised, FP64 and FP32) --------------------------------
 * y = (K_side + K_side^T) x without forming K, for patches whose elements fill the tensor grid
 * of their knot spans. Generalises YETI's scalar matrix-free apply mf_iga_get_ku_3d
 * (iga_wq_mf/fsrc/mf_iga_solver.f90:292-555) to 3-component NURBS elasticity. precisions: bit 0
 * prepares FP64, bit 1 FP32 (FP64 geometry is always kept for the diagonal blocks). x, y are
 * device arrays of 3 nb_cp values (dof 3 I + k); the apply is deterministic (no atomics). */
typedef struct yeti_mf yeti_mf;

int32_t yeti_mf_create(const yeti_system_inputs *in, int32_t precisions, yeti_mf **out);
int32_t yeti_mf_nb_cp(const yeti_mf *m);
int32_t yeti_mf_apply_f64(const yeti_mf *m, const double *x, double *y, void *stream);
int32_t yeti_mf_apply_f32(const yeti_mf *m, const float *x, float *y, void *stream);
/* 3x3 diagonal block of every CP, blocks[9 I + 3 k + l] (FP64). */
int32_t yeti_mf_diag_blocks_f64(const yeti_mf *m, double *blocks, void *stream);
/* g[3 I + j] = -u^T (dK/dX_Ij) u, the shape derivative of the stiffness part of the compliance
 * (FP64, deterministic). Twice YETI's gradLinElastWork_AN internal term for the adjoint 1/2 u
 * (optim/gradtotalwork_fullAN.f90:17-300, OPT_modelling.py:447-481). */
int32_t yeti_mf_strain_energy_grad_f64(const yeti_mf *m, const double *u_dev, double *g_dev, void *stream);
/* Replace the control-point coordinates (device, 3 nb_cp) and recompute the stored geometry. */
int32_t yeti_mf_update_coords(yeti_mf *m, const double *coords_dev, void *stream);
void    yeti_mf_destroy(yeti_mf *m);

/* FP64 PCG with the FP64 matrix-free apply and a 3x3 nodal block-Jacobi preconditioner on the
 * free dofs (the CG loop of iga_mf_cg_3d, mf_iga_solver.f90:642-958). x = 0 on fixed dofs.
 * norm 2: stop at ||r||_2 <= rtol ||b||_2; norm 0: stop at max|r| <= rtol max|b| over the free
 * dofs, iga_mf_cg_3d's own rule (maxval(abs(r)) / maxval(abs(b)), mf_iga_solver.f90:779,935);
 * relres is reported in that norm. */
int32_t yeti_mf_pcg_f64(const yeti_mf *m, const double *b, const uint8_t *free_mask, double *x, double rtol,
                        int32_t maxit, int32_t check_every, int32_t norm, int32_t *iters, double *relres,
                        void *stream);
/* FP64 iterative refinement with FP32 inner PCG (FP32 apply + FP32 block-Jacobi, FP64 dots):
 * stops when ||b - K x||_2 <= rtol ||b||_2 over the free dofs. */
int32_t yeti_mf_solve_ir(const yeti_mf *m, const double *b, const uint8_t *free_mask, double *x, double rtol,
                         double inner_rtol, int32_t max_outer, int32_t max_inner, int32_t check_every,
                         int32_t *outer_iters, int32_t *inner_iters, double *relres, void *stream);

/* ---- set-up inside caller-owned device memory ------------------------------------------
 * Same as yeti_mf_create / yeti_asm_create, but every device buffer of the handle is carved from
 * [arena, arena + arena_bytes) and the set-up's copies and kernels run on `stream` (only that
 * stream is synchronised, no hipMalloc / hipFree / device-wide synchronisation). With
 * arena == NULL nothing touches the device: *used returns the bytes the set-up needs and *out
 * stays NULL. The arena must outlive the handle; destroying the handle frees nothing in it.
 * Used by the serving loop to admit requests against an HBM budget allocated once. */
int32_t yeti_mf_create_in(const yeti_system_inputs *in, int32_t precisions, void *arena, size_t arena_bytes,
                          void *stream, yeti_mf **out, size_t *used);
int32_t yeti_asm_create_in(const yeti_system_inputs *in, int32_t layout, void *arena, size_t arena_bytes,
                           void *stream, yeti_asm **out, size_t *used);

/* ---- steppable, device-controlled solve (serving) ---------------------------------------
 * Solves K[free][:, free] x = b[free] with the matrix-free operator m by 3x3 block-Jacobi PCG.
 * precision 32: vectors, operator and preconditioner in FP32 (FP64 dot products), with the
 * solution and the true residual kept in FP64 and the residual replaced by mask (b - K x) from
 * the FP64 apply whenever the iterated residual has fallen by `delta` since the last replacement
 * (mixed-precision refinement that keeps the Krylov space); precision 64: the same in FP64.
 * Stops at ||b - K x||_2 <= rtol ||b||_2 on the FP64 residual. A replacement is also forced every
 * 256 iterations; when four replacements in a row do not halve the true residual the session
 * stops with `floor` set (the FP64 rounding floor of an ill-conditioned system lies above rtol);
 * `failed` is set at max_iters or max_updates. Control state lives on the device:
 * yeti_ir_step enqueues `chunk` iterations, every kernel predicated on device flags, so a step is
 * a fixed launch sequence; with use_graph it is captured once into a hipGraph (on a non-null
 * stream) and replayed. The iterates do not depend on how steps of different sessions
 * interleave. All device memory is taken at creation (arena semantics as above; arena == NULL
 * counts). No reference counterpart: the reference solves one system per call
 * (iga_wq_mf/fsrc/mf_iga_solver.f90:642-958; scipy spsolve in its callers). */
typedef struct yeti_ir yeti_ir;
typedef struct yeti_fdm yeti_fdm;   /* fast diagonalisation data of a session (opaque) */
typedef struct yeti_ir_status {
    int32_t active, update_pending, stagnant, done, failed, iters, floor, updates;
    double rz, pq, rr, rz_new, bb, rr_true, rr_max, tol2, rr_true_best, last_update_iter;
    double reserved[6];
} yeti_ir_status;

/* precond 0: 3x3 nodal block-Jacobi; 1: fast diagonalisation (single-patch handles whose fixed
 * dofs are whole patch faces per component), the preconditioner family of iga_mf_cg_3d
 * (iga_wq_mf/fsrc/tensor_algebra.f90:992-1161): per component the separable operator
 * sum_d alpha_d K_d (x) M (x) M built from 1D B-spline Gram matrices and the patch-mean
 * Jacobian, inverted exactly by 1D generalized eigendecompositions. */
int32_t yeti_ir_create(const yeti_mf *m, const uint8_t *free_mask_dev, int32_t precision, int32_t precond, double rtol,
                       double delta, int32_t chunk, int32_t max_iters, int32_t max_updates, void *arena,
                       size_t arena_bytes, void *stream, yeti_ir **out, size_t *used);
/* Make the session's FP64 products (every product of a precision-64 session, and the residual
 * replacement of either precision) go through an assembled device CSR of K + K^T (values vals,
 * which must stay valid) instead of the matrix-free operator: the solve is then the FP64 solve of
 * the assembled matrix. c == NULL restores the matrix-free operator. Call before the first step
 * of a graph-stepped session. */
int32_t yeti_ir_set_csr(yeti_ir *s, const yeti_csr *c, const double *vals);
/* x = 0 and the first direction set up for right-hand side b (device buffers kept by the session). */
int32_t yeti_ir_begin(yeti_ir *s, const double *b_dev, double *x_dev, void *stream);
int32_t yeti_ir_step(yeti_ir *s, int32_t use_graph, void *stream);
/* Enqueue a copy of the status into host memory (pinned for an asynchronous copy); read it once
 * the stream has reached this point. Relative residual = sqrt(rr_true / bb). */
int32_t yeti_ir_status_async(const yeti_ir *s, yeti_ir_status *host_status, void *stream);
void    yeti_ir_destroy(yeti_ir *s);

/* ---- batches of sessions (serving) ----------------------------------------------------------
 * A batch advances many sessions of one precision (32 or 64) with the same launches: each launch
 * of a step walks device task lists naming a member and a piece of its work (a brick of elements,
 * a block of control points, a chunk of a vector, a tile of a preconditioner product), so one
 * launch moves every active member forward. Members must be single-patch sessions with the fast
 * diagonalisation preconditioner, created and begun (yeti_ir_begin) by the caller; their status
 * flags and solutions stay in the sessions. A member's reductions are split into chunks that
 * depend only on its own size, so its iterates are the same whatever else shares the batch.
 * yeti_batch_step enqueues `chunk` iterations of every member (captured once into a hipGraph
 * with use_graph; membership changes need no new capture). task_capacity bounds each task list.
 * No reference counterpart (the reference solves one system per call). */
typedef struct yeti_batch yeti_batch;
int32_t yeti_batch_create(int32_t precision, int32_t chunk, int64_t task_capacity, yeti_batch **out);
int32_t yeti_batch_add(yeti_batch *b, yeti_ir *s);
int32_t yeti_batch_remove(yeti_batch *b, yeti_ir *s);
int32_t yeti_batch_size(const yeti_batch *b);
int32_t yeti_batch_step(yeti_batch *b, int32_t use_graph, void *stream);
/* statuses of the members in the order they were added (removed ones leave the order), into
 * host memory (pinned for an asynchronous copy); read it once the stream has reached this point */
int32_t yeti_batch_status_async(yeti_batch *b, yeti_ir_status *host_status, void *stream);
void    yeti_batch_destroy(yeti_batch *b);

/* ---- Displacement reconstruction ------------------------------------------------------------
 * Replaces f2py reconstructionSOL.reconstruction (src/yeti_iga/main/reconstruction.f:22-135,
 * called as rsol.reconstruction(**get_inputs4solution(x)) at overlay/iga_model.py:504 and
 * OPT_modelling.py:456,571): U (nb_cp*mcrd, dof-interleaved) = 0, then U[ind_dof_free[d]-1] =
 * u_inv[d] for d < nb_dof_free, then, boundary condition by boundary condition (later ones
 * overwrite earlier ones), U[(bc_target[.]-1)*mcrd + dir-1] = value for every condition whose
 * direction bc_values[2b] <= 3 (value bc_values[2b+1]); sol (nb_cp x mcrd, column-major) = U
 * reshaped. Device pointers: u_inv, ind_dof_free, bc_target (concatenated targets, 1-based CP
 * ids), sol, U. Host pointers: bc_target_nbelem[nb_bc], bc_values[2*nb_bc]. Out-of-range dof
 * indices are skipped on the device; callers validate them (the Python mirror does). */
/* ---- Displacement at parametric points ------------------------------------------------------
 * Replaces f2py postproc.evaldisp / evaldispmulti (src/yeti_iga/postprocessing/evaluate_disp.f:
 * 20-103, 105-189; called at OPT_modelling.py:573): for each point xi (3 per point) of one patch,
 * the element containing it (igaparmetrization.f90:321-372: per direction a binary search over
 * the interior knots, intervals [left, right), then the element whose spans match, element 1 when
 * none or several do), the rational basis there (evaluatenurbsfcts.f:94-201) and
 * disp = sum_a R_a sol(:, IEN(a, e)); disp has 3 values per point (beyond mcrd: 0). All pointers
 * are host memory; the evaluation runs on the device. Degrees up to 8 per direction. */
typedef struct yeti_nurbs_patch {
    int32_t p[3];                 /* degrees */
    int32_t nkv[3];               /* knot counts */
    const double *U[3];           /* knot vectors */
    int32_t nelem, nnode;
    const int32_t *nijk;          /* (3, nelem), 1-based spans */
    const int32_t *ien;           /* (nnode, nelem), 1-based CP ids */
    const double *weight;         /* (nnode, nelem) */
} yeti_nurbs_patch;
int32_t yeti_evaldisp(const yeti_nurbs_patch *patch, int32_t nb_cp, int32_t mcrd, const double *sol, int64_t npts,
                      const double *xi, double *disp);

int32_t yeti_reconstruction(const double *u_inv, const int32_t *ind_dof_free, int32_t nb_dof_free,
                            const int32_t *bc_target, const int32_t *bc_target_nbelem_host,
                            const double *bc_values_host, int32_t nb_bc, int32_t mcrd, int32_t nb_cp,
                            double *sol, double *U, void *stream);

/* VTU output of solid (U1) patches: generateVTU, postprocessing/postproc_vtu.f:22-297, called as
 * postproc.generatevtu(*IGAparametrization.get_inputs4postprocVTU(...))
 * (IGA_parametrization.py:717-752).
 *
 * yeti_svars_solid replaces compute_svars_Q1 (postprocessing/compute_svars_solid.f:24-188) over
 * the elements of one patch, on the GPU: host arrays in and out; coords3d and sol are (3, nb_cp)
 * column-major; nb_ref[d] in 1..8 gives 2^(nb_ref[d]-1)+1 vertices per direction; flags =
 * (disp, stress, strain) as FieldOutput_flag; E, nu the patch's MATERIAL_PROPERTIES. svars
 * receives nelem x nvertex x nsvint doubles (nsvint = 3 + 3 disp + 6 stress + 6 strain), the
 * reference's svars(nsvint*nb_vertice, nb_elem).
 *
 * yeti_vtu_write_piece replaces writePatch (postprocessing/writepatch.f) for a 3D patch: it
 * appends the <Piece> of those svars (points, hexahedron cells, disp, stress, von Mises,
 * strain) to the file at path. Host only. */
int32_t yeti_svars_solid(const yeti_nurbs_patch *patch, int32_t nb_cp, const double *coords3d, const double *sol,
                         const int32_t *nb_ref, const int32_t *flags, double E, double nu, double *svars);
int32_t yeti_vtu_write_piece(const char *path, const double *svars, int32_t nb_elem, const int32_t *nb_ref,
                             const int32_t *flags);

#ifdef __cplusplus
}
#endif
#endif /* YETI_AMD_H */
