"""The oracle's own marshalling of sys_linmat_lindef_static's arguments into the C struct
`yeti_system_inputs` (test infrastructure: only the oracle uses it).

It shares nothing with the product's marshaller (yeti_amd/_abi.py): the struct's fields, their
C types and their order are read from include/yeti_amd.h itself, the one definition the C
oracle is compiled against, and every field is filled by a rule of its own from the argument
list, so a packing mistake on either side shows up as a disagreement between the two (the
tests compare them field by field).

Argument order: the positional list IGAparametrization.get_inputs4system_elemStorage returns
(IGA_parametrization.py:1705-1716), i.e. build_system_elemStorage.f90:24-33 with the sizes f2py
infers moved to the end.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "yeti_amd.h")

# positional arguments (IGA_parametrization.py:1705-1716), then the optional sizes
POSITIONAL = [
    "activeElement", "nb_data", "COORDS3D", "IEN", "nb_elem_patch", "Nkv", "Ukv", "Nijk", "weight", "Jpqr",
    "ELT_TYPE", "PROPS", "JPROPS", "MATERIAL_PROPERTIES", "n_mat_props", "RHO", "TENSOR", "indDLoad", "JDLType",
    "ADLMAG", "load_target_nbelem", "load_additionalInfos", "nb_load_additionalInfos", "bc_values", "bc_target",
    "bc_target_nbelem", "ind_dof_free", "nb_dof_free", "MCRD", "NBINT", "NNODE", "nodal_dist",
]
OPTIONAL_SIZES = ["nb_bc", "nb_load", "nb_patch", "nb_elem", "nb_cp", "nb_dof_tot", "nb_n_dist", "nb_cp_n_dist",
                  "n_mat_props_max"]

_CTYPES = {"int32_t": C.c_int32, "int64_t": C.c_int64, "double": C.c_double, "char": C.c_char}


def header_fields(path=HEADER):
    """[(name, kind, ctype)] of `yeti_system_inputs` in declaration order; kind is "ptr" or
    "val"."""
    txt = open(path).read()
    body = txt[txt.index("typedef struct yeti_system_inputs {"):txt.index("} yeti_system_inputs;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for m in re.finditer(r"(?:const\s+)?(int32_t|int64_t|double|char)\s*(\*?)\s*([A-Za-z_][A-Za-z0-9_]*)\s*;", body):
        base, star, name = m.groups()
        out.append((name, "ptr" if star else "val", base))
    return out


def _struct_type():
    fields = []
    for name, kind, base in header_fields():
        ct = _CTYPES[base]
        fields.append((name, C.POINTER(ct) if kind == "ptr" else ct))
    return type("OracleSystemInputs", (C.Structure,), {"_fields_": fields})


OracleSystemInputs = _struct_type()


def header_define(name, path=HEADER):
    m = re.search(r"#define\s+%s\s+(\d+)" % name, open(path).read())
    if m is None:
        raise KeyError(name)
    return int(m.group(1))


YETI_LAYOUT_PACKED = header_define("YETI_LAYOUT_PACKED")
YETI_LAYOUT_SLOTTED = header_define("YETI_LAYOUT_SLOTTED")


def to_dict(args=(), kwargs=None):
    """Arguments by name from the positional list and/or keywords (keywords are matched without
    regard to case, as f2py's lower-case names)."""
    names = POSITIONAL + OPTIONAL_SIZES
    if len(args) > len(names):
        raise TypeError("too many arguments (%d)" % len(args))
    d = dict(zip(names, args))
    by_lower = {n.lower(): n for n in names}
    for k, v in (kwargs or {}).items():
        n = by_lower.get(k.lower())
        if n is None or n in d:
            raise TypeError("bad or repeated argument %r" % k)
        d[n] = v
    lacking = [n for n in POSITIONAL if n not in d]
    if lacking:
        raise TypeError("missing arguments: %s" % lacking)
    return d


def _column_major(a, dtype):
    """The values of a (any shape) in Fortran order, as a contiguous 1D array."""
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1, order="F"))


def _as_text(v):
    if isinstance(v, bytes):
        return v
    if isinstance(v, np.ndarray):
        return v.tobytes() if v.dtype.kind == "S" else str(v.item() if v.ndim == 0 else "".join(map(str, v))).encode()
    return str(v).encode("ascii")


def pack(d):
    """(OracleSystemInputs, keepalive) from the arguments by name."""
    s = OracleSystemInputs()
    keep = []
    arrays = {}

    def put(name, base):
        v = d[name]
        if base == "char":
            raw = _as_text(v)
            buf = C.create_string_buffer(raw, len(raw) + 1)
            keep.append(buf)
            setattr(s, name, C.cast(buf, C.POINTER(C.c_char)))
            arrays[name] = len(raw)
            return
        dt = np.int32 if base == "int32_t" else np.float64 if base == "double" else np.int64
        a = _column_major(v, dt)
        n = a.size
        if n == 0:                       # a valid pointer even for an empty array
            a = np.zeros(1, dt)
        keep.append(a)
        setattr(s, name, a.ctypes.data_as(C.POINTER(_CTYPES[base])))
        arrays[name] = n

    fields = header_fields()
    for name, kind, base in fields:
        if kind == "ptr":
            put(name, base)

    def shape2(name, rows_hint=None):
        a = np.asarray(d[name])
        if a.ndim >= 2:
            return a.shape[0], int(np.prod(a.shape[1:]))
        return (rows_hint or a.size), (a.size // rows_hint if rows_hint else 1)

    # scalar fields: given values, else the size of the first argument the Fortran declares
    # with that dimension (build_system_elemStorage.f90:44-104), which is the one f2py reads
    infer = {
        "nb_bc": lambda: shape2("bc_values", 2)[1],
        "nb_load": lambda: np.asarray(d["JDLType"]).size,
        "nb_patch": lambda: np.asarray(d["nb_elem_patch"]).size,
        "nb_elem": lambda: np.asarray(d["activeElement"]).size,
        "nb_cp": lambda: np.asarray(d["COORDS3D"]).size // 3,
        "nb_dof_tot": lambda: np.asarray(d["ind_dof_free"]).size,
        "nb_n_dist": lambda: shape2("nodal_dist")[0],
        "nb_cp_n_dist": lambda: shape2("nodal_dist")[1],
        "n_mat_props_max": lambda: shape2("MATERIAL_PROPERTIES")[0],
    }
    for name, kind, base in fields:
        if kind != "val":
            continue
        if name.startswith("len_"):
            setattr(s, name, int(arrays[name[4:]]))
        elif name in d:
            setattr(s, name, int(d[name]))
        else:
            setattr(s, name, int(infer[name]()))
    return s, keep


def patch_slice(numpatch, IEN, elementsByPatch, Nkv, Ukv, Nijk, weight, Jpqr, NNODE):
    """Patch `numpatch` (1-based) cut out of the flattened model arrays, whose per-patch
    blocks follow one another in patch order (IGA_parametrization.py:548-590): a dict with the
    degrees p (3,), knot counts nkv (3,), the three knot vectors U, the element count and CPs
    per element, Nijk as (nelem, 3), the 1-based IEN block and its weights."""
    counts = [int(v) for v in np.asarray(elementsByPatch).reshape(-1)]
    ip = int(numpatch) - 1
    if ip < 0 or ip >= len(counts):
        raise ValueError("numpatch should be between 1 and %d" % len(counts))
    nkv_all = np.asarray(Nkv).reshape(-1, order="F").reshape(-1, 3)          # row j = patch j (F order (3, npatch))
    deg_all = np.asarray(Jpqr).reshape(-1, order="F").reshape(-1, 3)
    nnode_all = [int(v) for v in np.asarray(NNODE).reshape(-1)]
    knots = np.asarray(Ukv, np.float64).reshape(-1)
    k0 = int(nkv_all[:ip].sum())
    U = []
    for n in nkv_all[ip]:
        U.append(np.ascontiguousarray(knots[k0:k0 + int(n)]))
        k0 += int(n)
    e0 = sum(counts[:ip])
    i0 = sum(c * n for c, n in zip(counts[:ip], nnode_all[:ip]))
    ne, nn = counts[ip], nnode_all[ip]
    spans = np.asarray(Nijk).reshape(-1, order="F").reshape(-1, 3)           # row e = element e
    return {"p": np.array(deg_all[ip], np.int32), "nkv": np.array(nkv_all[ip], np.int32), "U": U,
            "nelem": ne, "nnode": nn,
            "nijk": np.ascontiguousarray(spans[e0:e0 + ne], np.int32),
            "ien": np.ascontiguousarray(np.asarray(IEN).reshape(-1)[i0:i0 + ne * nn], np.int32),
            "w": np.ascontiguousarray(np.asarray(weight, np.float64).reshape(-1)[i0:i0 + ne * nn])}
