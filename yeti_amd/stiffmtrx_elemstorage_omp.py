"""Drop-in for YETI's f2py module ``yeti_iga.stiffmtrx_elemstorage_omp``.

``sys_linmat_lindef_static_omp(**kwargs)`` accepts the keyword arguments of the OpenMP routine
(src/yeti_iga/main/build_system_elemStorage_OMP.f90:26-38, as
IGA_parametrization.py:1611-1653 builds them) and returns (Kdata, Krow, Kcol, F) in that
routine's slot layout: element e of a patch owns COO slots
``count + e * nb_data_elem ...`` (build_system_elemStorage_OMP.f90:229), inactive slots stay 0.
The ``len_*`` sizes and ``num_threads`` are accepted and ignored (sizes are taken from the
arrays; the work runs on the GPU).
"""
from __future__ import annotations

from ._abi import YETI_LAYOUT_SLOTTED, args_to_dict
from .stiffmtrx_elemstorage import _run

_IGNORED = {"len_ukv", "len_weight", "len_props", "len_ien", "len_inddload",
            "len_load_additionalinfos", "len_nb_load_additionalinfos", "len_bc_target",
            "num_threads"}


def sys_linmat_lindef_static_omp(**kwargs):
    """GPU replacement of ``stiffmtrx_elemstorage_omp.sys_linmat_lindef_static_omp`` (keyword
    call, as the reference's callers use it: benchs/omp/test_build_stifmatrix_omp.py:63-64)."""
    kw = {k: v for k, v in kwargs.items() if k.lower() not in _IGNORED}
    return _run(args_to_dict((), kw), YETI_LAYOUT_SLOTTED)
