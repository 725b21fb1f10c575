"""Matrix-free NURBS elasticity operator and the mixed-precision solves built on it.

``MatrixFree`` wraps ``yeti_mf_*`` (include/yeti_amd.h): a sum-factorised apply of
K_side + K_side^T in FP64 or FP32, generalising YETI's scalar matrix-free operator
mf_iga_get_ku_3d (iga_wq_mf/fsrc/mf_iga_solver.f90:292-555) to 3-component elasticity on NURBS
patches; ``pcg_f64`` is the plain FP64 solve and ``solve_ir`` the FP64 iterative refinement with
FP32 inner PCG.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._abi import args_to_dict, pack_system_inputs
from .assembly import _stream_handle

F64 = 1
F32 = 2


class MatrixFree:
    def __init__(self, args, precisions=F64 | F32):
        _lib.require_device()
        d = args if isinstance(args, dict) else args_to_dict(args)
        s, keep = pack_system_inputs(d)
        h = C.c_void_p()
        _lib.check(_lib.lib().yeti_mf_create(C.byref(s), int(precisions), C.byref(h)))
        self._h = h
        self.nb_cp = int(_lib.lib().yeti_mf_nb_cp(h))
        self.ndof = 3 * self.nb_cp

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().yeti_mf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, x, y=None, stream=None):
        if y is None:
            y = torch.empty_like(x)
        if x.dtype == torch.float64:
            f = _lib.lib().yeti_mf_apply_f64
        elif x.dtype == torch.float32:
            f = _lib.lib().yeti_mf_apply_f32
        else:
            raise TypeError("x must be float64 or float32")
        if x.numel() != self.ndof or y.dtype != x.dtype or not x.is_cuda:
            raise ValueError("x, y must be CUDA tensors of %d entries and the same dtype" % self.ndof)
        _lib.check(f(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), _stream_handle(stream)))
        return y

    def diag_blocks(self, stream=None):
        out = torch.empty(9 * self.nb_cp, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().yeti_mf_diag_blocks_f64(self._h, C.c_void_p(out.data_ptr()), _stream_handle(stream)))
        return out.view(self.nb_cp, 3, 3)

    def strain_energy_grad(self, u, out=None, stream=None):
        """g[3 I + j] = -u^T (dK/dX_Ij) u (FP64), the stiffness part of the compliance shape
        gradient (optim/gradtotalwork_fullAN.f90:17-300 with adjoint u)."""
        if u.dtype != torch.float64 or u.numel() != self.ndof or not u.is_cuda:
            raise ValueError("u must be a CUDA float64 tensor of %d entries" % self.ndof)
        if out is None:
            out = torch.empty_like(u)
        _lib.check(_lib.lib().yeti_mf_strain_energy_grad_f64(self._h, C.c_void_p(u.data_ptr()),
                                                              C.c_void_p(out.data_ptr()), _stream_handle(stream)))
        return out

    def update_coords(self, coords, stream=None):
        """New control-point coordinates (CUDA float64, nb_cp x 3 or 3 nb_cp, COORDS3D order)."""
        c = coords.reshape(-1)
        if c.dtype != torch.float64 or c.numel() != self.ndof or not c.is_cuda or not c.is_contiguous():
            raise ValueError("coords must be a contiguous CUDA float64 tensor of %d entries" % self.ndof)
        _lib.check(_lib.lib().yeti_mf_update_coords(self._h, C.c_void_p(c.data_ptr()), _stream_handle(stream)))

    def pcg_f64(self, b, free_mask, rtol=1e-10, maxit=200000, check_every=20, norm="2", stream=None):
        """FP64 PCG (block-Jacobi) on the free dofs: (x, iterations, relative residual). norm "2"
        stops at ||r||_2 <= rtol ||b||_2, norm "inf" at max|r| <= rtol max|b|, the rule of
        iga_mf_cg_3d (mf_iga_solver.f90:779,935)."""
        if norm not in ("2", "inf"):
            raise ValueError("norm must be '2' or 'inf'")
        x = torch.empty(self.ndof, dtype=torch.float64, device=b.device)
        it, rel = C.c_int32(0), C.c_double(0.0)
        _lib.check(_lib.lib().yeti_mf_pcg_f64(self._h, C.c_void_p(b.data_ptr()), C.c_void_p(free_mask.data_ptr()),
                                              C.c_void_p(x.data_ptr()), float(rtol), int(maxit), int(check_every),
                                              2 if norm == "2" else 0, C.byref(it), C.byref(rel),
                                              _stream_handle(stream)))
        return x, int(it.value), float(rel.value)

    def solve_ir(self, b, free_mask, rtol=1e-10, inner_rtol=1e-4, max_outer=50, max_inner=100000,
                 check_every=20, stream=None):
        x = torch.empty(self.ndof, dtype=torch.float64, device=b.device)
        oi, ii, rel = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
        _lib.check(_lib.lib().yeti_mf_solve_ir(self._h, C.c_void_p(b.data_ptr()), C.c_void_p(free_mask.data_ptr()),
                                               C.c_void_p(x.data_ptr()), float(rtol), float(inner_rtol),
                                               int(max_outer), int(max_inner), int(check_every), C.byref(oi),
                                               C.byref(ii), C.byref(rel), _stream_handle(stream)))
        return x, int(oi.value), int(ii.value), float(rel.value)
