"""Drop-in for YETI's f2py module ``yeti_iga.stiffmtrx_elemstorage``.

``sys_linmat_lindef_static`` takes the same positional arguments as the f2py routine
(src/yeti_iga/main/build_system_elemStorage.f90:24-33, marshalled by
IGA_parametrization.py:1705-1716) and returns the same four arrays::

    Kdata, Krow, Kcol, F = sys_linmat_lindef_static(*igapara.get_inputs4system_elemStorage())

Kdata float64[nb_data], Krow/Kcol int32[nb_data] (0-based dof ids, element-major upper-triangle
blocks with halved diagonals, inactive elements packed out), F float64[nb_dof_tot].

The work runs on the GPU through libyeti_hip.so; there is no CPU fallback. Where the reference
prints a message and carries on (unsupported element type, build_system_elemStorage.f90:255-257),
this raises ``YetiError``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._abi import YETI_LAYOUT_PACKED, args_to_dict, pack_system_inputs

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


def _run(d, layout):
    _lib.require_device()
    s, keep = pack_system_inputs(d)
    nd = int(s.nb_data)
    if nd < 0:
        raise ValueError("nb_data must be >= 0")
    K = np.zeros(max(nd, 1), np.float64)
    r = np.zeros(max(nd, 1), np.int32)
    c = np.zeros(max(nd, 1), np.int32)
    F = np.zeros(max(int(s.nb_dof_tot), 1), np.float64)
    _lib.check(_lib.lib().yeti_sys_linmat_lindef_static(
        C.byref(s), int(layout), K.ctypes.data_as(_f64p), r.ctypes.data_as(_i32p),
        c.ctypes.data_as(_i32p), F.ctypes.data_as(_f64p)))
    return K[:nd], r[:nd], c[:nd], F[: int(s.nb_dof_tot)]


def sys_linmat_lindef_static(*args, **kwargs):
    """GPU replacement of ``stiffmtrx_elemstorage.sys_linmat_lindef_static``."""
    return _run(args_to_dict(args, kwargs), YETI_LAYOUT_PACKED)
