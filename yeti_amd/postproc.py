"""Drop-in for entry points of YETI's f2py module ``postproc``
(src/yeti_iga/postprocessing/CMakeLists.txt:18-38):

* ``evaldisp`` and ``evaldispmulti`` (evaluate_disp.f:20-189), called with the argument lists of
  ``IGAparametrization.get_inputs4evaldisp`` / ``get_inputs4evaldispmulti``
  (IGA_parametrization.py:908-975), e.g. OPT_modelling.py:573. The evaluation runs on the GPU
  (``yeti_evaldisp``); ``evaldisp_points`` takes many points at once;
* ``generatevtu`` (postproc_vtu.f:22-297), called with the argument list of
  ``IGAparametrization.get_inputs4postprocVTU`` (IGA_parametrization.py:717-752): the fields at
  each element's vertex grid are computed on the GPU (``yeti_svars_solid``), the text is written
  by native host code (``yeti_vtu_write_piece``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib


class _Patch(C.Structure):
    _fields_ = [("p", C.c_int32 * 3), ("nkv", C.c_int32 * 3), ("U", C.c_void_p * 3), ("nelem", C.c_int32),
                ("nnode", C.c_int32), ("nijk", C.c_void_p), ("ien", C.c_void_p), ("weight", C.c_void_p)]


def _patch(numpatch, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE):
    """Slices patch `numpatch` (1-based) out of the flattened inputs (the order of
    IGA_parametrization._flatten_data, IGA_parametrization.py:548-590)."""
    nep = np.asarray(elementsByPatch, np.int64).reshape(-1)
    npatch = nep.size
    ip = int(numpatch) - 1
    if not 0 <= ip < npatch:
        raise ValueError("numpatch should be between 1 and %d" % npatch)
    Nkv = np.asarray(Nkv, np.int64).reshape(3, -1, order="F")
    Jpqr = np.asarray(Jpqr, np.int64).reshape(3, -1, order="F")
    nnode = np.asarray(NNODE, np.int64).reshape(-1)
    Ukv = np.asarray(Ukv, np.float64).reshape(-1)
    IEN = np.asarray(IEN, np.int64).reshape(-1)
    weight = np.asarray(weight, np.float64).reshape(-1)
    Nijk = np.asarray(Nijk, np.int64).reshape(3, -1, order="F")
    kv = int(Nkv[:, :ip].sum())
    e0 = int(nep[:ip].sum())
    ie = int((nnode[:ip] * nep[:ip]).sum())
    ne, nn = int(nep[ip]), int(nnode[ip])
    keep = {"U": [np.ascontiguousarray(Ukv[kv + int(Nkv[:d, ip].sum()): kv + int(Nkv[:d + 1, ip].sum())])
                  for d in range(3)],
            "nijk": np.ascontiguousarray(Nijk[:, e0:e0 + ne].T, dtype=np.int32),       # (nelem, 3) = F (3, nelem)
            "ien": np.ascontiguousarray(IEN[ie:ie + nn * ne], dtype=np.int32),
            "w": np.ascontiguousarray(weight[ie:ie + nn * ne], dtype=np.float64)}
    s = _Patch()
    for d in range(3):
        s.p[d] = int(Jpqr[d, ip])
        s.nkv[d] = int(Nkv[d, ip])
        s.U[d] = keep["U"][d].ctypes.data
    s.nelem, s.nnode = ne, nn
    s.nijk = keep["nijk"].ctypes.data
    s.ien = keep["ien"].ctypes.data
    s.weight = keep["w"].ctypes.data
    return s, keep


def evaldisp_points(numpatch, xi_points, sol, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE,
                    mcrd=3):
    """Displacement at many parametric points: xi_points (npts, 3) -> (npts, 3). sol (mcrd, nb_cp)."""
    _lib.require_device()
    s, keep = _patch(numpatch, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE)
    S = np.asfortranarray(np.asarray(sol, np.float64).reshape(mcrd, -1, order="F"))
    X = np.ascontiguousarray(np.asarray(xi_points, np.float64).reshape(-1, 3))
    out = np.zeros((X.shape[0], 3))
    _lib.check(_lib.lib().yeti_evaldisp(C.byref(s), S.shape[1], int(mcrd), S.ctypes.data_as(C.c_void_p),
                                        X.shape[0], X.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)))
    return out


def evaldisp(numpatch, xi, sol, COORDS3D, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, ELT_TYPE,
             MATERIAL_PROPERTIES, TENSOR, PROPS, JPROPS, NNODE, nb_patch=None, nb_elem=None, nb_cp=None,
             mcrd=None):
    """f2py postproc.evaldisp: disp (3,) at xi (3,) of patch numpatch; sol (mcrd, nb_cp)."""
    m = int(mcrd) if mcrd is not None else np.asarray(sol).shape[0]
    return evaldisp_points(numpatch, np.asarray(xi, np.float64).reshape(1, 3), sol, IEN, elementsByPatch, Nkv, Ukv,
                           Nijk, weight, Jpqr, NNODE, m)[0]


def evaldispmulti(numpatch, n_xi, n_eta, n_zeta, sol, COORDS3D, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr,
                  ELT_TYPE, MATERIAL_PROPERTIES, TENSOR, PROPS, JPROPS, NNODE, nb_patch=None, nb_elem=None,
                  nb_cp=None, mcrd=None):
    """f2py postproc.evaldispmulti: disp (n_xi, n_eta, n_zeta, 3) on the uniform grid
    xi = (i - 1) / (n - 1) of [0, 1]^3 (evaluate_disp.f:166-171)."""
    m = int(mcrd) if mcrd is not None else np.asarray(sol).shape[0]
    # ((i - 1) * 1) / (n - 1) as the reference computes it (n = 1 divides by 0 there: NaN here)
    g = [(np.arange(n, dtype=np.float64) / (n - 1.0)) if n > 1 else np.array([np.nan]) for n in
         (n_xi, n_eta, n_zeta)]
    X = np.stack(np.meshgrid(g[0], g[1], g[2], indexing="ij"), axis=-1).reshape(-1, 3)
    d = evaldisp_points(numpatch, X, sol, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE, m)
    return np.asfortranarray(d.reshape(n_xi, n_eta, n_zeta, 3))


def _patch_types(ELT_TYPE, npatch):
    """Element type of each patch from the flattened string ("U1U1...", split at 'U' as
    m_nurbspatch.f90:186-201 does)."""
    if isinstance(ELT_TYPE, (bytes, np.bytes_)):
        s = ELT_TYPE.decode()
    elif isinstance(ELT_TYPE, str):
        s = ELT_TYPE
    else:   # an array of strings or characters
        s = "".join(x.decode() if isinstance(x, bytes) else str(x) for x in np.asarray(ELT_TYPE).reshape(-1))
    types = ["U" + t.strip() for t in s.split("U")[1:]]
    if len(types) < npatch:
        raise ValueError("ELT_TYPE %r names %d patches, %d expected" % (s, len(types), npatch))
    return types[:npatch]


def svars_solid(numpatch, nb_refinement, FieldOutput_flag, sol, COORDS3D, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight,
                Jpqr, NNODE, MATERIAL_PROPERTIES):
    """compute_svars_Q1 (compute_svars_solid.f:24-188) over patch numpatch (1-based) on the GPU:
    (nelem, nvertex, nsvint) with per vertex the point, [disp], [stress], [strain] as
    FieldOutput_flag selects. sol (3, nb_cp), COORDS3D (3, nb_cp)."""
    _lib.require_device()
    s, keep = _patch(numpatch, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE)
    X = np.asarray(COORDS3D, np.float64)
    ncp = X.shape[1]
    S = np.asarray(sol, np.float64)
    if X.shape != (3, ncp) or S.shape != (3, ncp):
        raise ValueError("COORDS3D and sol must be (3, nb_cp), got %s and %s" % (X.shape, S.shape))
    Xc = np.ascontiguousarray(X.T)
    Sc = np.ascontiguousarray(S.T)
    ref = np.maximum(np.asarray(nb_refinement, np.int32).reshape(3), 1).astype(np.int32)
    fl = np.asarray([int(bool(f)) for f in np.asarray(FieldOutput_flag).reshape(3)], np.int32)
    nb = 2 ** np.maximum(ref - 1, 0) + 1
    nsv = 3 + 3 * fl[0] + 6 * fl[1] + 6 * fl[2]
    out = np.empty((int(s.nelem), int(np.prod(nb)), int(nsv)))
    mat = np.asarray(MATERIAL_PROPERTIES, np.float64).reshape(2, -1, order="F")
    ip = int(numpatch) - 1
    _lib.check(_lib.lib().yeti_svars_solid(C.byref(s), ncp, Xc.ctypes.data_as(C.c_void_p),
                                           Sc.ctypes.data_as(C.c_void_p), ref.ctypes.data_as(C.c_void_p),
                                           fl.ctypes.data_as(C.c_void_p), float(mat[0, ip]), float(mat[1, ip]),
                                           out.ctypes.data_as(C.c_void_p)))
    return out


def generatevtu(FILENAME, output_path, FieldOutput_flag, nb_refinement, sol, COORDS3D, IEN, nb_elem_patch, Nkv, Ukv,
                Nijk, weight, Jpqr, ELT_TYPE, MATERIAL_PROPERTIES, TENSOR, PROPS, JPROPS, NNODE, nb_patch=None,
                nb_elem=None, nb_cp=None, MCRD=None):
    """f2py postproc.generatevtu: writes output_path/FILENAME.vtu with one <Piece> per solid (U1)
    patch (postproc_vtu.f:22-297). sol is SOL(MCRD, nb_cp) (callers pass SOL.transpose()).
    Patches of shell / embedded types (U2, U3, U10, U30), which the reference also writes, raise
    NotImplementedError; patches of any other type are skipped, as there. Numbers are written in
    the shortest form that reads back to the same double."""
    nep = np.asarray(nb_elem_patch).reshape(-1)
    npatch = int(nb_patch) if nb_patch is not None else nep.size
    mcrd = int(MCRD) if MCRD is not None else np.asarray(sol).shape[0]
    if mcrd != 3:
        raise NotImplementedError("generatevtu: only MCRD = 3 (solid) models are supported")
    types = _patch_types(ELT_TYPE, npatch)
    for t in types:
        if t in ("U2", "U3", "U10", "U30"):
            raise NotImplementedError("generatevtu: element type %s (shell / embedded) is not supported" % t)
    ref = np.maximum(np.asarray(nb_refinement, np.int32).reshape(3), 1).astype(np.int32)
    fl = np.asarray([int(bool(f)) for f in np.asarray(FieldOutput_flag).reshape(3)], np.int32)
    path = os.path.join(str(output_path), str(FILENAME) + ".vtu")
    with open(path, "w") as fh:
        fh.write(' <VTKFile type="UnstructuredGrid"  version="0.1"   >\n <UnstructuredGrid>\n')
    for ip, t in enumerate(types):
        if t != "U1":
            continue
        sv = svars_solid(ip + 1, ref, fl, sol, COORDS3D, IEN, nb_elem_patch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE,
                         MATERIAL_PROPERTIES)
        _lib.check(_lib.lib().yeti_vtu_write_piece(path.encode(), sv.ctypes.data_as(C.c_void_p), sv.shape[0],
                                                   ref.ctypes.data_as(C.c_void_p), fl.ctypes.data_as(C.c_void_p)))
    with open(path, "a") as fh:
        fh.write(" </UnstructuredGrid>\n </VTKFile>\n")
    return path
