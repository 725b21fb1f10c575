"""Loader of libyeti_hip.so (the HIP kernels behind include/yeti_amd.h).

There is no CPU fallback: every entry point of this package fails loudly when the library
is missing or when no HIP device is visible.
"""
from __future__ import annotations

import ctypes as C
import os

from ._abi import ERROR_NAMES, YETI_ERR_NOT_CONVERGED, YetiSystemInputs  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyeti_hip.so")

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol include/yeti_amd.h declares
SIGNATURES = {
    "yeti_last_error": (C.c_char_p, []),
    "yeti_version": (C.c_int32, []),
    "yeti_device_count": (C.c_int32, []),
    "yeti_sys_linmat_lindef_static": (C.c_int32, [C.POINTER(YetiSystemInputs), C.c_int32, _f64p,
                                                  _i32p, _i32p, _f64p]),
    "yeti_asm_create": (C.c_int32, [C.POINTER(YetiSystemInputs), C.c_int32, C.POINTER(C.c_void_p)]),
    "yeti_asm_nb_data": (C.c_int64, [C.c_void_p]),
    "yeti_asm_run_coo": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "yeti_asm_run_coo_matrix": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_asm_run_coo_values": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_asm_run_rhs": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_asm_run_mass_coo": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_asm_strain_energy_grad": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_asm_volume_grad": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_asm_run_vib_grad": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "yeti_asm_update_coords": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_asm_load_shape_grad": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_grad_lin_elast_work": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                             C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_asm_set_timing": (C.c_int32, [C.c_void_p, C.c_int32]),
    "yeti_asm_kernel_times": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "yeti_asm_destroy": (None, [C.c_void_p]),
    "yeti_csr_create": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "yeti_csr_nnz": (C.c_int64, [C.c_void_p]),
    "yeti_csr_nrows": (C.c_int32, [C.c_void_p]),
    "yeti_csr_rowptr": (C.c_void_p, [C.c_void_p]),
    "yeti_csr_colind": (C.c_void_p, [C.c_void_p]),
    "yeti_csr_pattern_to_host": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_csr_values": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_csr_spmv": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_csr_spmv_abs": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_csr_pcg": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int32,
                                 C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p]),
    "yeti_csr_destroy": (None, [C.c_void_p]),
    "yeti_mf_create": (C.c_int32, [C.POINTER(YetiSystemInputs), C.c_int32, C.POINTER(C.c_void_p)]),
    "yeti_mf_nb_cp": (C.c_int32, [C.c_void_p]),
    "yeti_mf_apply_f64": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_mf_apply_f32": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_mf_diag_blocks_f64": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_mf_strain_energy_grad_f64": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_mf_update_coords": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_mf_destroy": (None, [C.c_void_p]),
    "yeti_mf_pcg_f64": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int32,
                                    C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p]),
    "yeti_mf_solve_ir": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                     C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p]),
    "yeti_mf_create_in": (C.c_int32, [C.POINTER(YetiSystemInputs), C.c_int32, C.c_void_p, C.c_size_t, C.c_void_p,
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "yeti_asm_create_in": (C.c_int32, [C.POINTER(YetiSystemInputs), C.c_int32, C.c_void_p, C.c_size_t, C.c_void_p,
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "yeti_ir_create": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                   C.c_int32, C.c_int32, C.c_void_p, C.c_size_t, C.c_void_p, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_size_t)]),
    "yeti_ir_set_csr": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_ir_begin": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_ir_step": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p]),
    "yeti_ir_status_async": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_ir_destroy": (None, [C.c_void_p]),
    "yeti_batch_create": (C.c_int32, [C.c_int32, C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]),
    "yeti_batch_add": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "yeti_batch_remove": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "yeti_batch_size": (C.c_int32, [C.c_void_p]),
    "yeti_batch_step": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p]),
    "yeti_batch_status_async": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_batch_destroy": (None, [C.c_void_p]),
    "yeti_csr_drop_zeros": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "yeti_evaldisp": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "yeti_reconstruction": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "yeti_svars_solid": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_double, C.c_double, C.c_void_p]),
    "yeti_vtu_write_piece": (C.c_int32, [C.c_char_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "yeti_pcg_csr": (C.c_int32, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_double, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_double), C.c_void_p]),
    "yeti_pcg_csr_noalloc": (C.c_int32, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_double, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_double), C.c_void_p]),
}

_lib = None


class YetiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d): %s" % (ERROR_NAMES.get(code, "error"), code, msg))
        self.code = code


def lib():
    """The loaded library. Raises when it is not built."""
    global _lib
    if _lib is None:
        # torch first: its HIP runtime (same soname as the system's) is then the one this library
        # binds to, so the process holds one runtime whichever module the caller touches first
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libyeti_hip.so is not built: run `python -m yeti_amd.csrc.build` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().yeti_last_error().decode(errors="replace")
        raise YetiError(rc, msg)


def require_device():
    if lib().yeti_device_count() < 1:
        raise RuntimeError("no HIP device visible: the YETI MI355X kernels need a GPU")
