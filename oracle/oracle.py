"""Python entry points of the oracle. TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only as the checker. Two independent CPU implementations are exposed:

* ``oracle_sys_linmat`` -- liboracle.so, our C restatement (oracle/yeti_oracle.c);
* ``oracle_reconstruction`` / ``oracle_evaldisp`` / ``oracle_mass_coo`` / ``oracle_svars_solid``
  / ``oracle_grad_vibration`` -- the same library's restatements of the displacement scatter,
  the displacement at parametric points, the consistent mass matrix, the VTU vertex fields and
  the natural-frequency shape gradient;
* ``ref_sys_linmat_omp`` / ``ref_sys_linmat_serial`` / ``ref_dersbasisfuns`` / ``ref_gauss`` /
  ``ref_grad_linelast_work`` / ``ref_reconstruction`` / ``ref_evaldisp`` / ``ref_build_cmassmatrix`` /
  ``ref_generate_vtu`` / ``ref_grad_vibration``
  -- oracle/_ref/libyetiref.so, YETI's own Fortran compiled from
  its sources by oracle/build_ref.sh (present only where the YETI sources were available at
  build time; callers must handle ``ref_available() == False``).
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle.inputs import YETI_LAYOUT_PACKED, YETI_LAYOUT_SLOTTED  # noqa: E402
from oracle.inputs import pack as pack_system_inputs  # noqa: E402
from oracle.inputs import to_dict as args_to_dict  # noqa: E402

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        _lib = C.CDLL(path)
        _lib.oracle_sys_linmat.restype = C.c_int
        _lib.oracle_sys_linmat.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p, _i32p, _i32p, _f64p]
        _lib.oracle_gauss.restype = C.c_int
        _lib.oracle_gauss.argtypes = [C.c_int, _f64p, _f64p]
        _lib.oracle_dersbasisfuns.restype = None
        _lib.oracle_dersbasisfuns.argtypes = [C.c_int, C.c_int, C.c_double, _f64p, _f64p, _f64p]
        _lib.oracle_element_matrix.restype = C.c_int
        _lib.oracle_element_matrix.argtypes = [C.c_void_p, C.c_int64, _f64p, _f64p]
        _lib.oracle_shape_at.restype = C.c_int
        _lib.oracle_shape_at.argtypes = [C.c_void_p, C.c_int64, _f64p, _f64p, _f64p, _f64p]
        _lib.oracle_mass_coo.restype = C.c_int
        _lib.oracle_mass_coo.argtypes = [C.c_void_p, _f64p, _i32p, _i32p]
        _lib.oracle_evaldisp.restype = C.c_int
        _lib.oracle_evaldisp.argtypes = [_i32p, _i32p, _f64p, _f64p, _f64p, C.c_int32, C.c_int32, _i32p, _i32p, _f64p,
                                         C.c_int32, _f64p, C.c_int64, _f64p, _f64p]
        _lib.oracle_reconstruction.restype = C.c_int
        _lib.oracle_reconstruction.argtypes = [_f64p, _i32p, C.c_int32, _i32p, _i32p, _f64p, C.c_int32, C.c_int32,
                                               C.c_int32, _f64p, _f64p]
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libyetiref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libyetiref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libyetiref.so is not built (needs the YETI sources)")
        _ref = C.CDLL(path)
    return _ref


def _p(a, t):
    return a.ctypes.data_as(t)


def _nb_data_for(s, layout):
    return int(s.nb_data)


# ---------------------------------------------------------------------------------------
def oracle_sys_linmat(args, layout=YETI_LAYOUT_PACKED, nthreads=None, kwargs=None):
    """Our C restatement of sys_linmat_lindef_static -> (Kdata, Krow, Kcol, F)."""
    d = args_to_dict(args, kwargs)
    s, keep = pack_system_inputs(d)
    nd = int(s.nb_data)
    K = np.zeros(nd, np.float64)
    r = np.zeros(nd, np.int32)
    c = np.zeros(nd, np.int32)
    F = np.zeros(int(s.nb_dof_tot), np.float64)
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    rc = lib().oracle_sys_linmat(C.byref(s), int(layout), int(nthreads), _p(K, _f64p),
                                 _p(r, _i32p), _p(c, _i32p), _p(F, _f64p))
    if rc != 0:
        raise RuntimeError("oracle_sys_linmat failed with code %d" % rc)
    return K, r, c, F


def oracle_element_matrix(args, elem):
    d = args_to_dict(args)
    s, keep = pack_system_inputs(d)
    n = int(np.asarray(d["NNODE"]).ravel()[0])
    A = np.zeros(9 * n * (n + 1) // 2)
    R = np.zeros(3 * n)
    rc = lib().oracle_element_matrix(C.byref(s), int(elem), _p(A, _f64p), _p(R, _f64p))
    if rc != 0:
        raise RuntimeError("oracle_element_matrix failed with code %d" % rc)
    return A.reshape(-1, 3, 3).transpose(0, 2, 1), R   # [pair][kk][ll]


def oracle_shape_at(args, elem, pt):
    d = args_to_dict(args)
    s, keep = pack_system_inputs(d)
    n = int(np.asarray(d["NNODE"]).ravel()[0])
    R = np.zeros(n)
    dR = np.zeros((n, 3))
    det = np.zeros(1)
    ptv = np.ascontiguousarray(pt, dtype=np.float64)
    rc = lib().oracle_shape_at(C.byref(s), int(elem), _p(ptv, _f64p), _p(R, _f64p),
                               _p(dR, _f64p), _p(det, _f64p))
    if rc != 0:
        raise RuntimeError("oracle_shape_at failed with code %d" % rc)
    return R, dR, float(det[0])


def oracle_gauss(n):
    x = np.zeros(n)
    w = np.zeros(n)
    if lib().oracle_gauss(int(n), _p(x, _f64p), _p(w, _f64p)):
        raise ValueError("unsupported Gauss order %d" % n)
    return x, w


def oracle_dersbasisfuns(span1, p, u, U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    N = np.zeros(p + 1)
    dN = np.zeros(p + 1)
    lib().oracle_dersbasisfuns(int(span1), int(p), float(u), _p(U, _f64p), _p(N, _f64p), _p(dN, _f64p))
    return N, dN


# ---------------------------------------------------------------------------------------
# The compiled reference
def ref_gauss(n, dim=1, face=0):
    """YETI's Gauss(norder, dim, GaussWtCoord, iFace) (main/GaussLegendre.f90:27)."""
    out = np.zeros((dim + 1, n ** dim), order="F")
    f = ref().gauss_
    f.restype = None
    f(C.byref(C.c_int(n)), C.byref(C.c_int(dim)), out.ctypes.data_as(_f64p), C.byref(C.c_int(face)))
    return out


def ref_dersbasisfuns(span1, p, u, U):
    """YETI's dersbasisfuns (main/dersbasisfuns.f90:30)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    N = np.zeros(p + 1)
    dN = np.zeros(p + 1)
    f = ref().dersbasisfuns_
    f.restype = None
    f(C.byref(C.c_int(span1)), C.byref(C.c_int(p)), C.byref(C.c_int(U.size)),
      C.byref(C.c_double(u)), _p(U, _f64p), _p(N, _f64p), _p(dN, _f64p))
    return N, dN


def _ival(v):
    return C.byref(C.c_int32(int(v)))


def ref_sys_linmat_omp(args, num_threads=1, kwargs=None):
    """YETI's sys_linmat_lindef_static_omp (build_system_elemStorage_OMP.f90:26), called with
    the Fortran calling convention (all by reference, hidden CHARACTER lengths appended)."""
    d = args_to_dict(args, kwargs)
    s, keep = pack_system_inputs(d)
    nd = int(s.nb_data)
    K = np.zeros(nd, np.float64)
    r = np.zeros(nd, np.int32)
    c = np.zeros(nd, np.int32)
    F = np.zeros(int(s.nb_dof_tot), np.float64)
    f = ref().sys_linmat_lindef_static_omp_
    f.restype = None
    nb_data64 = C.c_int64(nd)
    keep2 = []

    def iv(v):
        x = C.c_int32(int(v)); keep2.append(x); return C.byref(x)

    f(_p(K, _f64p), _p(r, _i32p), _p(c, _i32p), _p(F, _f64p),
      s.activeElement, C.byref(nb_data64), s.COORDS3D, s.IEN, s.nb_elem_patch, s.Nkv, s.Ukv,
      s.Nijk, s.weight, s.Jpqr, s.ELT_TYPE, s.PROPS, s.JPROPS, s.MATERIAL_PROPERTIES,
      s.n_mat_props, s.RHO, s.TENSOR, s.indDLoad, s.JDLType, s.ADLMAG, s.load_target_nbelem,
      s.load_additionalInfos, s.nb_load_additionalInfos, s.bc_values, iv(s.nb_bc), s.bc_target,
      s.bc_target_nbelem, s.ind_dof_free, iv(s.nb_dof_free), iv(s.MCRD), s.NBINT, iv(s.nb_load),
      iv(s.nb_patch), iv(s.nb_elem), s.NNODE, iv(s.nb_cp), iv(s.nb_dof_tot), s.nodal_dist,
      iv(s.nb_n_dist), iv(s.nb_cp_n_dist), iv(s.n_mat_props_max),
      iv(s.len_Ukv), iv(s.len_weight), iv(max(1, s.len_PROPS)), iv(s.len_IEN),
      iv(max(1, s.len_indDLoad)), iv(max(1, s.len_load_additionalInfos)), iv(s.nb_load),
      iv(max(1, s.len_bc_target)), iv(num_threads),
      C.c_size_t(s.len_ELT_TYPE), C.c_size_t(s.len_TENSOR))
    return K, r, c, F


def ref_sys_linmat_serial(args, kwargs=None):
    """YETI's serial sys_linmat_lindef_static (build_system_elemStorage.f90:24) through the
    bind(C) wrapper in oracle/ref_shim.f90. nb_data must fit int32 (as in the reference)."""
    d = args_to_dict(args, kwargs)
    s, keep = pack_system_inputs(d)
    nd = int(s.nb_data)
    if nd >= 2 ** 31:
        raise ValueError("the serial reference counts nb_data in int32")
    K = np.zeros(nd, np.float64)
    r = np.zeros(nd, np.int32)
    c = np.zeros(nd, np.int32)
    F = np.zeros(int(s.nb_dof_tot), np.float64)
    f = ref().ref_sys_linmat_serial
    f.restype = None
    keep2 = []

    def iv(v):
        x = C.c_int32(int(v)); keep2.append(x); return C.byref(x)

    f(_p(K, _f64p), _p(r, _i32p), _p(c, _i32p), _p(F, _f64p), s.activeElement, iv(nd),
      s.COORDS3D, s.IEN, iv(s.len_IEN), s.nb_elem_patch, s.Nkv, s.Ukv, iv(s.len_Ukv), s.Nijk,
      s.weight, iv(s.len_weight), s.Jpqr, s.ELT_TYPE, iv(s.len_ELT_TYPE), s.PROPS,
      iv(max(1, s.len_PROPS)), s.JPROPS, s.MATERIAL_PROPERTIES, s.n_mat_props, s.RHO, s.TENSOR,
      iv(s.len_TENSOR), s.indDLoad, iv(max(1, s.len_indDLoad)), s.JDLType, s.ADLMAG,
      s.load_target_nbelem, s.load_additionalInfos, iv(max(1, s.len_load_additionalInfos)),
      s.nb_load_additionalInfos, s.bc_values, s.bc_target, iv(max(1, s.len_bc_target)),
      s.bc_target_nbelem, s.ind_dof_free, iv(s.nb_dof_free), iv(s.MCRD), s.NBINT, s.NNODE,
      s.nodal_dist, iv(s.nb_bc), iv(s.nb_load), iv(s.nb_patch), iv(s.nb_elem), iv(s.nb_cp),
      iv(s.nb_dof_tot), iv(s.nb_n_dist), iv(s.nb_cp_n_dist), iv(s.n_mat_props_max))
    return K, r, c, F


def ref_grad_linelast_work(args, SOL, ADJ, compute_wint=True, compute_wext=True, active_dir=(1, 1, 1)):
    """YETI's gradLinElastWork_AN (optim/gradtotalwork_fullAN.f90:17) through
    oracle/ref_shim.f90. SOL: (3, nb_cp); ADJ: (nadj, 3, nb_cp). Returns (gradWint, gradWext),
    each (nadj, 3, nb_cp)."""
    d = args_to_dict(args)
    s, keep = pack_system_inputs(d)
    SOLf = np.asfortranarray(SOL, dtype=np.float64)
    ADJf = np.asfortranarray(ADJ, dtype=np.float64)
    nadj = ADJf.shape[0]
    nb_cp = int(s.nb_cp)
    gi = np.zeros((nadj, 3, nb_cp), order="F")
    ge = np.zeros((nadj, 3, nb_cp), order="F")
    ad = np.ascontiguousarray(active_dir, dtype=np.int32)
    matp = np.asfortranarray(np.asarray(d["MATERIAL_PROPERTIES"], np.float64)[:2])
    f = ref().ref_grad_linelast_work
    f.restype = None
    keep2 = []

    def iv(v):
        x = C.c_int32(int(v)); keep2.append(x); return C.byref(x)

    f(gi.ctypes.data_as(_f64p), ge.ctypes.data_as(_f64p), iv(1 if compute_wint else 0),
      iv(1 if compute_wext else 0), SOLf.ctypes.data_as(_f64p), ADJf.ctypes.data_as(_f64p),
      iv(nadj), s.activeElement, _p(ad, _i32p), s.COORDS3D, s.IEN, iv(s.len_IEN), s.nb_elem_patch,
      s.Nkv, s.Ukv, iv(s.len_Ukv), s.Nijk, s.weight, iv(s.len_weight), s.Jpqr, s.ELT_TYPE,
      iv(s.len_ELT_TYPE), s.PROPS, iv(max(1, s.len_PROPS)), s.JPROPS, matp.ctypes.data_as(_f64p),
      s.RHO, s.TENSOR, iv(s.len_TENSOR), s.indDLoad, iv(max(1, s.len_indDLoad)), s.JDLType,
      s.ADLMAG, s.load_target_nbelem, s.load_additionalInfos,
      iv(max(1, s.len_load_additionalInfos)), iv(s.MCRD), s.NBINT, iv(s.nb_load),
      iv(s.nb_patch), iv(s.nb_elem), s.NNODE, iv(nb_cp))
    return gi, ge


def _recon_arrays(u_inv, ind_dof_free, bc_target, bc_target_nbelem, bc_values, mcrd, nb_cp):
    u = np.ascontiguousarray(u_inv, dtype=np.float64)
    idf = np.ascontiguousarray(ind_dof_free, dtype=np.int32)
    bt = np.ascontiguousarray(bc_target, dtype=np.int32).reshape(-1)
    if bt.size == 0:
        bt = np.zeros(1, np.int32)
    bn = np.ascontiguousarray(bc_target_nbelem, dtype=np.int32).reshape(-1)
    bv = np.asfortranarray(np.asarray(bc_values, dtype=np.float64).reshape(2, -1, order="F"))
    return u, idf, bt, bn, bv


def oracle_reconstruction(u_inv, ind_dof_free, bc_target, bc_target_nbelem, bc_values, mcrd, nb_cp):
    """liboracle's restatement of reconstruction.f:22-135. Returns (sol (nb_cp, mcrd) Fortran
    order, U (nb_cp mcrd))."""
    u, idf, bt, bn, bv = _recon_arrays(u_inv, ind_dof_free, bc_target, bc_target_nbelem, bc_values, mcrd, nb_cp)
    sol = np.zeros((nb_cp, mcrd), order="F")
    U = np.zeros(nb_cp * mcrd)
    rc = lib().oracle_reconstruction(_p(u, _f64p), _p(idf, _i32p), int(u.size), _p(bt, _i32p), _p(bn, _i32p),
                                     _p(bv, _f64p), int(bn.size), int(mcrd), int(nb_cp), _p(sol, _f64p),
                                     _p(U, _f64p))
    if rc:
        raise ValueError("oracle_reconstruction: invalid dof index (code %d)" % rc)
    return sol, U


def ref_reconstruction(u_inv, coords, ind_dof_free, bc_target, bc_target_nbelem, bc_values, mcrd, nb_cp):
    """YETI's reconstruction (main/reconstruction.f:22) through oracle/ref_shim.f90; the output
    buffers start at zero."""
    u, idf, bt, bn, bv = _recon_arrays(u_inv, ind_dof_free, bc_target, bc_target_nbelem, bc_values, mcrd, nb_cp)
    X = np.asfortranarray(np.asarray(coords, dtype=np.float64).reshape(3, -1, order="F"))
    idf_full = np.zeros(nb_cp * mcrd, np.int32)
    idf_full[: idf.size] = idf
    sol = np.zeros((nb_cp, mcrd), order="F")
    U = np.zeros(nb_cp * mcrd)
    keep = []

    def iv(v):
        x = C.c_int32(int(v)); keep.append(x); return C.byref(x)

    f = ref().ref_reconstruction
    f.restype = None
    f(_p(sol, _f64p), _p(U, _f64p), _p(u, _f64p), _p(X, _f64p), _p(idf_full, _i32p), _p(bt, _i32p), iv(bt.size),
      _p(bn, _i32p), _p(bv, _f64p), iv(bn.size), iv(mcrd), iv(nb_cp), iv(u.size))
    return sol, U


def oracle_evaldisp(numpatch, xi_points, sol, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE, mcrd=3):
    """liboracle's restatement of evalDisp (evaluate_disp.f:20-103) at many points:
    xi_points (npts, 3) -> (npts, 3); sol (mcrd, nb_cp)."""
    from oracle.inputs import patch_slice
    keep = patch_slice(numpatch, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE)
    p, nkv = keep["p"], keep["nkv"]
    S = np.asfortranarray(np.asarray(sol, np.float64).reshape(mcrd, -1, order="F"))
    X = np.ascontiguousarray(np.asarray(xi_points, np.float64).reshape(-1, 3))
    out = np.zeros((X.shape[0], 3))
    U = keep["U"]
    rc = lib().oracle_evaldisp(_p(p, _i32p), _p(nkv, _i32p), _p(U[0], _f64p), _p(U[1], _f64p), _p(U[2], _f64p),
                               int(keep["nelem"]), int(keep["nnode"]), _p(keep["nijk"], _i32p), _p(keep["ien"], _i32p),
                               _p(keep["w"], _f64p), int(mcrd), _p(S, _f64p), X.shape[0], _p(X, _f64p),
                               _p(out, _f64p))
    if rc:
        raise ValueError("oracle_evaldisp failed (%d)" % rc)
    return out


def ref_evaldisp(d, numpatch, xi, sol):
    """YETI's evalDisp (postprocessing/evaluate_disp.f:20) through oracle/ref_shim.f90 for one point
    xi (3,), with the flattened model inputs of the dict d; sol (3, nb_cp)."""
    s, keep = pack_system_inputs(d)
    S = np.asfortranarray(np.asarray(sol, np.float64).reshape(3, -1, order="F"))
    X = np.ascontiguousarray(xi, np.float64)
    out = np.zeros(3)
    matp = np.asfortranarray(np.asarray(d["MATERIAL_PROPERTIES"], np.float64)[:2])
    keep2 = []

    def iv(v):
        x = C.c_int32(int(v)); keep2.append(x); return C.byref(x)

    f = ref().ref_evaldisp
    f.restype = None
    f(_p(out, _f64p), iv(numpatch), _p(X, _f64p), _p(S, _f64p), s.COORDS3D, s.IEN, iv(s.len_IEN), s.nb_elem_patch,
      s.Nkv, s.Ukv, iv(s.len_Ukv), s.Nijk, s.weight, iv(s.len_weight), s.Jpqr, s.ELT_TYPE, iv(s.len_ELT_TYPE),
      matp.ctypes.data_as(_f64p), s.TENSOR, iv(s.len_TENSOR), s.PROPS, iv(max(1, s.len_PROPS)), s.JPROPS, s.NNODE,
      iv(s.nb_patch), iv(s.nb_elem), iv(s.nb_cp), iv(3))
    return out


def oracle_mass_coo(args):
    """liboracle's restatement of build_CMASSMatrix (compute_CMASSMatrix.f:22) on the system
    inputs (dict or f2py positional list): (Mdata, Mrow, Mcol) in the packed layout."""
    d = args_to_dict(args) if not isinstance(args, dict) else args
    s, keep = pack_system_inputs(d)
    n = int(s.nb_data)
    Md = np.zeros(n)
    Mr = np.zeros(n, np.int32)
    Mc = np.zeros(n, np.int32)
    rc = lib().oracle_mass_coo(C.byref(s), _p(Md, _f64p), _p(Mr, _i32p), _p(Mc, _i32p))
    if rc:
        raise ValueError("oracle_mass_coo failed (%d)" % rc)
    return Md, Mr, Mc


def ref_build_cmassmatrix(args):
    """YETI's build_CMASSMatrix through oracle/ref_shim.f90 (every element active, as
    get_inputs4massmat gives for solid patches)."""
    d = args_to_dict(args) if not isinstance(args, dict) else args
    s, keep = pack_system_inputs(d)
    n = int(s.nb_data)
    Md = np.zeros(n)
    Mr = np.zeros(n, np.int32)
    Mc = np.zeros(n, np.int32)
    keep2 = []

    def iv(v):
        x = C.c_int32(int(v)); keep2.append(x); return C.byref(x)

    f = ref().ref_build_cmassmatrix
    f.restype = None
    f(_p(Md, _f64p), _p(Mr, _i32p), _p(Mc, _i32p), s.activeElement, iv(n), s.COORDS3D, s.IEN, iv(s.len_IEN),
      s.nb_elem_patch, s.Nkv, s.Ukv, iv(s.len_Ukv), s.Nijk, s.weight, iv(s.len_weight), s.Jpqr, s.ELT_TYPE,
      iv(s.len_ELT_TYPE), s.PROPS, iv(max(1, s.len_PROPS)), s.JPROPS, s.RHO, s.TENSOR, iv(s.len_TENSOR),
      s.ind_dof_free, iv(s.nb_dof_free), iv(s.MCRD), s.NBINT, iv(s.nb_patch), iv(s.nb_elem), s.NNODE, iv(s.nb_cp),
      iv(s.nb_dof_tot))
    return Md, Mr, Mc


def _vtu_counts(nb_ref, flags):
    nb = [2 ** max(int(r) - 1, 0) + 1 for r in nb_ref]
    nsv = 3 + 3 * bool(flags[0]) + 6 * bool(flags[1]) + 6 * bool(flags[2])
    return nb, nb[0] * nb[1] * nb[2], nsv


def oracle_svars_solid(d, ip, sol, nb_ref=(1, 1, 1), flags=(True, True, True)):
    """liboracle's restatement of compute_svars_Q1 (postprocessing/compute_svars_solid.f:24) over
    every element of solid patch ip (0-based); sol (nb_cp, 3). Returns (nelem, nvertex, nsvint)."""
    s, keep = pack_system_inputs(d)
    S = np.ascontiguousarray(np.asarray(sol, np.float64).reshape(-1, 3))
    ref_ = np.asarray(nb_ref, np.int32)
    fl = np.asarray([int(bool(f)) for f in flags], np.int32)
    _, nv, nsv = _vtu_counts(ref_, fl)
    nel = int(np.asarray(d["nb_elem_patch"]).reshape(-1)[ip])
    out = np.zeros((nel, nv, nsv))
    f = lib().oracle_svars_solid
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, _f64p, _i32p, _i32p, _f64p]
    rc = f(C.byref(s), int(ip), _p(S, _f64p), _p(ref_, _i32p), _p(fl, _i32p), _p(out, _f64p))
    if rc:
        raise ValueError("oracle_svars_solid failed (%d)" % rc)
    return out


def ref_generate_vtu(d, sol, outdir, filename="ref", nb_ref=(1, 1, 1), flags=(True, True, True)):
    """YETI's generateVTU (postprocessing/postproc_vtu.f:22) through oracle/ref_shim.f90; sol
    (nb_cp, 3). Writes outdir/filename.vtu and returns its path."""
    s, keep = pack_system_inputs(d)
    S = np.ascontiguousarray(np.asarray(sol, np.float64).reshape(-1, 3))
    matp = np.asfortranarray(np.asarray(d["MATERIAL_PROPERTIES"], np.float64)[:2])
    ref_ = np.asarray(nb_ref, np.int32)
    fl = np.asarray([int(bool(f)) for f in flags], np.int32)
    fn = filename.encode()
    op = str(outdir).encode()
    keep2 = []

    def iv(v):
        x = C.c_int32(int(v)); keep2.append(x); return C.byref(x)

    f = ref().ref_generate_vtu
    f.restype = None
    f(C.c_char_p(fn), iv(len(fn)), C.c_char_p(op), iv(len(op)), _p(fl, _i32p), _p(ref_, _i32p), _p(S, _f64p),
      s.COORDS3D, s.IEN, iv(s.len_IEN), s.nb_elem_patch, s.Nkv, s.Ukv, iv(s.len_Ukv), s.Nijk, s.weight,
      iv(s.len_weight), s.Jpqr, s.ELT_TYPE, iv(s.len_ELT_TYPE), matp.ctypes.data_as(_f64p), s.TENSOR,
      iv(s.len_TENSOR), s.PROPS, iv(max(1, s.len_PROPS)), s.JPROPS, s.NNODE, iv(s.nb_patch), iv(s.nb_elem),
      iv(s.nb_cp), iv(3))
    return os.path.join(str(outdir), filename + ".vtu")


def parse_vtu(path):
    """The pieces of a VTU file as written by writePatch (postprocessing/writepatch.f): a list of
    dicts with NumberOfPoints, NumberOfCells and one array per DataArray (points, connectivity,
    offsets, types, disp, stress, VM, strain). Tags may span lines (list-directed output wraps
    long records) and numbers are read whatever their formatting."""
    import re
    with open(path) as fh:
        text = fh.read()
    pieces = []
    for m in re.finditer(r"<Piece([^>]*)>(.*?)</Piece", text, re.S):
        nums = re.findall(r'"\s*(\d+)\s*"', m.group(1))
        cur = {"NumberOfPoints": int(nums[0]), "NumberOfCells": int(nums[1])}
        for a in re.finditer(r"<DataArray([^>]*)>(.*?)</DataArray", m.group(2), re.S):
            nm = re.search(r'Name="(\w+)"', a.group(1))
            toks = a.group(2).split()
            cur[nm.group(1) if nm else "points"] = np.array([float(t.replace("D", "E")) for t in toks])
        pieces.append(cur)
    return pieces


def oracle_grad_vibration(d, VECT, VALS):
    """liboracle's restatement of gradVibration_AN (optim/gradvibration_fullAN.f:22) over the
    active elements of d; VECT (nb_frq, 3, nb_cp), VALS (nb_frq,) -> gradV (nb_frq, 3, nb_cp)."""
    s, keep = pack_system_inputs(d)
    V = np.asfortranarray(np.asarray(VECT, np.float64))
    nf = V.shape[0]
    L = np.ascontiguousarray(np.asarray(VALS, np.float64).reshape(nf))
    out = np.zeros(V.shape, order="F")
    f = lib().oracle_grad_vibration
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int, _f64p]
    rc = f(C.byref(s), V.ctypes.data_as(_f64p), _p(L, _f64p), nf, out.ctypes.data_as(_f64p))
    if rc:
        raise ValueError("oracle_grad_vibration failed (%d)" % rc)
    return np.ascontiguousarray(out)


def ref_grad_vibration(d, VECT, VALS, activeDir=(1, 1, 1)):
    """YETI's gradVibration_AN through oracle/ref_shim.f90 over the active elements of d."""
    s, keep = pack_system_inputs(d)
    V = np.asfortranarray(np.asarray(VECT, np.float64))
    nf = V.shape[0]
    L = np.ascontiguousarray(np.asarray(VALS, np.float64).reshape(nf))
    out = np.zeros(V.shape, order="F")
    matp = np.asfortranarray(np.asarray(d["MATERIAL_PROPERTIES"], np.float64)[:2])
    adir = np.asarray(activeDir, np.int32)
    keep2 = []

    def iv(v):
        x = C.c_int32(int(v)); keep2.append(x); return C.byref(x)

    f = ref().ref_grad_vibration
    f.restype = None
    f(out.ctypes.data_as(_f64p), V.ctypes.data_as(_f64p), _p(L, _f64p), iv(nf), s.activeElement, _p(adir, _i32p),
      s.COORDS3D, s.IEN, iv(s.len_IEN), s.nb_elem_patch, s.Nkv, s.Ukv, iv(s.len_Ukv), s.Nijk, s.weight,
      iv(s.len_weight), s.Jpqr, s.ELT_TYPE, iv(s.len_ELT_TYPE), s.PROPS, iv(max(1, s.len_PROPS)), s.JPROPS,
      matp.ctypes.data_as(_f64p), s.RHO, s.TENSOR, iv(s.len_TENSOR), iv(3), s.NBINT, iv(s.nb_patch), iv(s.nb_elem),
      s.NNODE, iv(s.nb_cp))
    return np.ascontiguousarray(out)
