"""ctypes mirror of ``include/yeti_amd.h`` and the marshalling of YETI's f2py argument list.

``pack_system_inputs`` takes the positional argument list of
``stiffmtrx_elemstorage.sys_linmat_lindef_static`` exactly as
``IGAparametrization.get_inputs4system_elemStorage`` builds it
(src/yeti_iga/preprocessing/igaparametrization/IGA_parametrization.py:1705-1716) and returns a
``YetiSystemInputs`` structure whose pointers reference contiguous int32 / float64 copies in
Fortran order, which is what f2py itself does before calling the Fortran routine
(intent(in) arrays are cast and made Fortran-contiguous).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)

YETI_OK = 0
YETI_ERR_INVALID_ARGUMENT = 1
YETI_ERR_UNSUPPORTED_ELEMENT = 2
YETI_ERR_UNSUPPORTED_DEGREE = 3
YETI_ERR_SIZE_OVERFLOW = 4
YETI_ERR_DEVICE = 5
YETI_ERR_NOT_CONVERGED = 6
YETI_ERR_BUDGET = 7

YETI_LAYOUT_PACKED = 0
YETI_LAYOUT_SLOTTED = 1

ERROR_NAMES = {
    YETI_ERR_INVALID_ARGUMENT: "invalid argument",
    YETI_ERR_UNSUPPORTED_ELEMENT: "unsupported element type",
    YETI_ERR_UNSUPPORTED_DEGREE: "unsupported degree or Gauss order",
    YETI_ERR_SIZE_OVERFLOW: "size overflow",
    YETI_ERR_DEVICE: "HIP device error",
    YETI_ERR_NOT_CONVERGED: "iterative solve did not converge",
    YETI_ERR_BUDGET: "rejected by HBM budget",
}


class YetiSystemInputs(C.Structure):
    _fields_ = [
        ("activeElement", _i32p),
        ("nb_data", C.c_int64),
        ("COORDS3D", _f64p),
        ("IEN", _i32p),
        ("len_IEN", C.c_int64),
        ("nb_elem_patch", _i32p),
        ("Nkv", _i32p),
        ("Ukv", _f64p),
        ("len_Ukv", C.c_int64),
        ("Nijk", _i32p),
        ("weight", _f64p),
        ("len_weight", C.c_int64),
        ("Jpqr", _i32p),
        ("ELT_TYPE", C.c_char_p),
        ("len_ELT_TYPE", C.c_int64),
        ("PROPS", _f64p),
        ("len_PROPS", C.c_int64),
        ("JPROPS", _i32p),
        ("MATERIAL_PROPERTIES", _f64p),
        ("n_mat_props", _i32p),
        ("RHO", _f64p),
        ("TENSOR", C.c_char_p),
        ("len_TENSOR", C.c_int64),
        ("indDLoad", _i32p),
        ("len_indDLoad", C.c_int64),
        ("JDLType", _i32p),
        ("ADLMAG", _f64p),
        ("load_target_nbelem", _i32p),
        ("load_additionalInfos", _f64p),
        ("len_load_additionalInfos", C.c_int64),
        ("nb_load_additionalInfos", _i32p),
        ("bc_values", _f64p),
        ("bc_target", _i32p),
        ("len_bc_target", C.c_int64),
        ("bc_target_nbelem", _i32p),
        ("ind_dof_free", _i32p),
        ("nb_dof_free", C.c_int32),
        ("MCRD", C.c_int32),
        ("NBINT", _i32p),
        ("NNODE", _i32p),
        ("nodal_dist", _f64p),
        ("nb_bc", C.c_int32),
        ("nb_load", C.c_int32),
        ("nb_patch", C.c_int32),
        ("nb_elem", C.c_int32),
        ("nb_cp", C.c_int32),
        ("nb_dof_tot", C.c_int32),
        ("nb_n_dist", C.c_int32),
        ("nb_cp_n_dist", C.c_int32),
        ("n_mat_props_max", C.c_int32),
    ]


# Positional argument names of sys_linmat_lindef_static as f2py exposes it: the required
# arguments in Fortran order, then the optional (inferable) sizes
# (build_system_elemStorage.f90:24-33).
F2PY_ARGS = (
    "activeElement", "nb_data", "COORDS3D", "IEN", "nb_elem_patch", "Nkv", "Ukv", "Nijk",
    "weight", "Jpqr", "ELT_TYPE", "PROPS", "JPROPS", "MATERIAL_PROPERTIES", "n_mat_props",
    "RHO", "TENSOR", "indDLoad", "JDLType", "ADLMAG", "load_target_nbelem",
    "load_additionalInfos", "nb_load_additionalInfos", "bc_values", "bc_target",
    "bc_target_nbelem", "ind_dof_free", "nb_dof_free", "MCRD", "NBINT", "NNODE", "nodal_dist",
)
F2PY_OPTIONAL = ("nb_bc", "nb_load", "nb_patch", "nb_elem", "nb_cp", "nb_dof_tot",
                 "nb_n_dist", "nb_cp_n_dist", "n_mat_props_max")


def _f(a, dtype, ndmin=1):
    """Fortran-contiguous copy with at least one element (f2py passes a valid pointer even
    for empty arrays)."""
    arr = np.asarray(a, dtype=dtype)
    if arr.ndim < ndmin:
        arr = arr.reshape(arr.shape + (1,) * (ndmin - arr.ndim))
    arr = np.asfortranarray(arr)
    return arr


def _flat(a, dtype):
    arr = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))
    n = arr.size
    if n == 0:
        arr = np.zeros(1, dtype=dtype)
    return arr, n


def args_to_dict(args, kwargs=None):
    """Map the f2py positional (and keyword) arguments to a name -> value dict."""
    kwargs = dict(kwargs or {})
    if len(args) > len(F2PY_ARGS) + len(F2PY_OPTIONAL):
        raise TypeError("sys_linmat_lindef_static() takes at most %d positional arguments (%d given)"
                        % (len(F2PY_ARGS) + len(F2PY_OPTIONAL), len(args)))
    names = F2PY_ARGS + F2PY_OPTIONAL
    d = {}
    for name, val in zip(names, args):
        d[name] = val
    lower = {n.lower(): n for n in names}
    for k, v in kwargs.items():
        key = lower.get(k.lower())
        if key is None:
            raise TypeError("sys_linmat_lindef_static() got an unexpected keyword argument %r" % k)
        if key in d:
            raise TypeError("sys_linmat_lindef_static() got multiple values for argument %r" % k)
        d[key] = v
    missing = [n for n in F2PY_ARGS if n not in d]
    if missing:
        raise TypeError("sys_linmat_lindef_static() missing required argument(s): %s"
                        % ", ".join(missing))
    return d


def pack_system_inputs(d):
    """Build (YetiSystemInputs, keepalive) from a name -> value dict of the f2py arguments.

    Sizes missing from ``d`` are inferred from the array shapes, as f2py does."""
    keep = []

    def ptr_i(name):
        arr, n = _flat(d[name], np.int32)
        keep.append(arr)
        return arr.ctypes.data_as(_i32p), n

    def ptr_f(name):
        arr, n = _flat(d[name], np.float64)
        keep.append(arr)
        return arr.ctypes.data_as(_f64p), n

    def chars(name):
        s = d[name]
        if isinstance(s, np.ndarray):
            s = s.tobytes().decode() if s.dtype.kind in "SU" else str(s)
        b = str(s).encode("ascii")
        buf = C.create_string_buffer(b, len(b) + 1)
        keep.append(buf)
        return C.cast(buf, C.c_char_p), len(b)

    coords = _f(d["COORDS3D"], np.float64, 2)
    if coords.shape[0] != 3:
        raise ValueError("COORDS3D must have shape (3, nb_cp)")
    nkv = _f(d["Nkv"], np.int32, 2)
    nijk = _f(d["Nijk"], np.int32, 2)
    bcv = _f(d["bc_values"], np.float64, 2)
    matp = _f(d["MATERIAL_PROPERTIES"], np.float64, 2)
    nd = _f(d["nodal_dist"], np.float64, 2)

    s = YetiSystemInputs()
    s.activeElement, _ = ptr_i("activeElement")
    s.nb_data = int(d["nb_data"])
    arr, _ = _flat(coords, np.float64); keep.append(arr); s.COORDS3D = arr.ctypes.data_as(_f64p)
    s.IEN, s.len_IEN = ptr_i("IEN")
    s.nb_elem_patch, _ = ptr_i("nb_elem_patch")
    arr, _ = _flat(nkv, np.int32); keep.append(arr); s.Nkv = arr.ctypes.data_as(_i32p)
    s.Ukv, s.len_Ukv = ptr_f("Ukv")
    arr, _ = _flat(nijk, np.int32); keep.append(arr); s.Nijk = arr.ctypes.data_as(_i32p)
    s.weight, s.len_weight = ptr_f("weight")
    s.Jpqr, _ = ptr_i("Jpqr")
    s.ELT_TYPE, s.len_ELT_TYPE = chars("ELT_TYPE")
    s.PROPS, s.len_PROPS = ptr_f("PROPS")
    s.JPROPS, _ = ptr_i("JPROPS")
    arr, _ = _flat(matp, np.float64); keep.append(arr); s.MATERIAL_PROPERTIES = arr.ctypes.data_as(_f64p)
    s.n_mat_props, _ = ptr_i("n_mat_props")
    s.RHO, _ = ptr_f("RHO")
    s.TENSOR, s.len_TENSOR = chars("TENSOR")
    s.indDLoad, s.len_indDLoad = ptr_i("indDLoad")
    s.JDLType, _ = ptr_i("JDLType")
    s.ADLMAG, _ = ptr_f("ADLMAG")
    s.load_target_nbelem, _ = ptr_i("load_target_nbelem")
    s.load_additionalInfos, s.len_load_additionalInfos = ptr_f("load_additionalInfos")
    s.nb_load_additionalInfos, _ = ptr_i("nb_load_additionalInfos")
    arr, _ = _flat(bcv, np.float64); keep.append(arr); s.bc_values = arr.ctypes.data_as(_f64p)
    s.bc_target, s.len_bc_target = ptr_i("bc_target")
    s.bc_target_nbelem, _ = ptr_i("bc_target_nbelem")
    s.ind_dof_free, _ = ptr_i("ind_dof_free")
    s.nb_dof_free = int(d["nb_dof_free"])
    s.MCRD = int(d["MCRD"])
    s.NBINT, _ = ptr_i("NBINT")
    s.NNODE, _ = ptr_i("NNODE")
    arr, _ = _flat(nd, np.float64); keep.append(arr); s.nodal_dist = arr.ctypes.data_as(_f64p)

    s.nb_bc = int(d.get("nb_bc", bcv.shape[1]))
    s.nb_load = int(d.get("nb_load", np.asarray(d["ADLMAG"]).size))
    s.nb_patch = int(d.get("nb_patch", nkv.shape[1]))
    s.nb_elem = int(d.get("nb_elem", nijk.shape[1]))
    s.nb_cp = int(d.get("nb_cp", coords.shape[1]))
    s.nb_dof_tot = int(d.get("nb_dof_tot", np.asarray(d["ind_dof_free"]).size))
    s.nb_n_dist = int(d.get("nb_n_dist", nd.shape[0]))
    s.nb_cp_n_dist = int(d.get("nb_cp_n_dist", nd.shape[1]))
    s.n_mat_props_max = int(d.get("n_mat_props_max", matp.shape[0]))
    return s, keep
