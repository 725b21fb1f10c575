"""Drop-in for the displacement evaluations of YETI's f2py module ``postproc``
(src/yeti_iga/postprocessing/CMakeLists.txt:18-38): ``evaldisp`` and ``evaldispmulti``
(evaluate_disp.f:20-189), called with the argument lists of
``IGAparametrization.get_inputs4evaldisp`` / ``get_inputs4evaldispmulti``
(IGA_parametrization.py:908-975), e.g. OPT_modelling.py:573. The evaluation runs on the GPU
(``yeti_evaldisp``); ``evaldisp_points`` takes many points at once.
"""
from __future__ import annotations

import ctypes as C

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
