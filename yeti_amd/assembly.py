"""Device-resident assembly handle (``yeti_asm_*`` of include/yeti_amd.h) over torch-owned HBM.

Mirrors what ``sys_linmat_lindef_static`` (build_system_elemStorage.f90:24-315) computes, split
into a one-time set-up that copies the model to the GPU and runs that write the COO triplets
and F into device tensors. torch is used for device memory and streams only.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._abi import YETI_LAYOUT_PACKED, args_to_dict, pack_system_inputs


def _stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Assembler:
    """Holds one model resident on the current HIP device.

    ``args`` is the positional f2py argument list of sys_linmat_lindef_static (what
    ``IGAparametrization.get_inputs4system_elemStorage()`` returns) or a dict of those
    arguments by name."""

    def __init__(self, args, layout=YETI_LAYOUT_PACKED, kwargs=None):
        _lib.require_device()
        d = args if isinstance(args, dict) else args_to_dict(args, kwargs)
        self._inputs, self._keep = pack_system_inputs(d)
        self.nb_dof_tot = int(self._inputs.nb_dof_tot)
        self.nb_cp = int(self._inputs.nb_cp)
        h = C.c_void_p()
        _lib.check(_lib.lib().yeti_asm_create(C.byref(self._inputs), int(layout), C.byref(h)))
        self._h = h
        self.nb_data = int(_lib.lib().yeti_asm_nb_data(h))
        # the host copies are no longer needed once the model is resident
        self._keep = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().yeti_asm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------------------
    def update_coords(self, coords, stream=None):
        """New control-point coordinates (CUDA float64, nb_cp x 3 or 3 nb_cp, COORDS3D order)."""
        c = coords.reshape(-1)
        if c.dtype != torch.float64 or c.numel() != 3 * self.nb_cp or not c.is_cuda or not c.is_contiguous():
            raise ValueError("coords must be a contiguous CUDA float64 tensor of %d entries" % (3 * self.nb_cp))
        _lib.check(_lib.lib().yeti_asm_update_coords(self._h, C.c_void_p(c.data_ptr()), _stream_handle(stream)))

    def set_timing(self, on=True):
        """Bracket each element-kernel launch of later runs with HIP events on its stream."""
        _lib.check(_lib.lib().yeti_asm_set_timing(self._h, 1 if on else 0))

    def kernel_times(self, cap=4096):
        """Durations (ms) of the element-kernel launches of the last run (timing on)."""
        buf = (C.c_float * cap)()
        n = C.c_int32(0)
        _lib.check(_lib.lib().yeti_asm_kernel_times(self._h, buf, cap, C.byref(n)))
        return [float(buf[i]) for i in range(min(n.value, cap))]

    def load_shape_grad(self, u, out=None, stream=None):
        """g[3 I + j] = 2 u^T dF/dX_Ij for the face and body loads (FP64)."""
        if u.dtype != torch.float64 or u.numel() != 3 * self.nb_cp or not u.is_cuda:
            raise ValueError("u must be a CUDA float64 tensor of %d entries" % (3 * self.nb_cp))
        if out is None:
            out = torch.empty_like(u)
        _lib.check(_lib.lib().yeti_asm_load_shape_grad(self._h, C.c_void_p(u.data_ptr()),
                                                       C.c_void_p(out.data_ptr()), _stream_handle(stream)))
        return out

    # ------------------------------------------------------------------------------
    def alloc_coo(self, with_indices=True, device="cuda"):
        n = max(self.nb_data, 1)
        K = torch.empty(n, dtype=torch.float64, device=device)
        if not with_indices:
            return K, None, None
        r = torch.empty(n, dtype=torch.int32, device=device)
        c = torch.empty(n, dtype=torch.int32, device=device)
        return K, r, c

    def run_coo(self, Kdata=None, Krow=None, Kcol=None, F=None, stream=None):
        if Kdata is None:
            Kdata, Krow, Kcol = self.alloc_coo(True)
        if F is None:
            F = torch.empty(self.nb_dof_tot, dtype=torch.float64, device=Kdata.device)
        self._check_out(Kdata, Krow, Kcol, F)
        _lib.check(_lib.lib().yeti_asm_run_coo(self._h, C.c_void_p(Kdata.data_ptr()),
                                               C.c_void_p(Krow.data_ptr()),
                                               C.c_void_p(Kcol.data_ptr()),
                                               C.c_void_p(F.data_ptr()), _stream_handle(stream)))
        return Kdata, Krow, Kcol, F

    def run_matrix(self, Kdata, Krow, Kcol, stream=None):
        self._check_out(Kdata, Krow, Kcol, None)
        _lib.check(_lib.lib().yeti_asm_run_coo_matrix(self._h, C.c_void_p(Kdata.data_ptr()),
                                                      C.c_void_p(Krow.data_ptr()),
                                                      C.c_void_p(Kcol.data_ptr()),
                                                      _stream_handle(stream)))
        return Kdata, Krow, Kcol

    def run_values(self, Kdata, stream=None):
        self._check_out(Kdata, None, None, None)
        _lib.check(_lib.lib().yeti_asm_run_coo_values(self._h, C.c_void_p(Kdata.data_ptr()),
                                                      _stream_handle(stream)))
        return Kdata

    def run_rhs(self, F=None, stream=None):
        if F is None:
            F = torch.empty(self.nb_dof_tot, dtype=torch.float64, device="cuda")
        if F.dtype != torch.float64 or not F.is_cuda or F.numel() < self.nb_dof_tot:
            raise ValueError("F must be a CUDA float64 tensor of at least nb_dof_tot entries")
        _lib.check(_lib.lib().yeti_asm_run_rhs(self._h, C.c_void_p(F.data_ptr()), _stream_handle(stream)))
        return F

    def run_mass(self, M=None, r=None, c=None, with_indices=True, stream=None):
        """Consistent mass matrix (compute_CMASSMatrix.f) in the stiffness COO layout:
        (Mdata, Mrow, Mcol) device tensors (Mrow/Mcol None when with_indices is False)."""
        if M is None:
            M, r0, c0 = self.alloc_coo(with_indices)
            r = r if r is not None else r0
            c = c if c is not None else c0
        self._check_out(M, r, c, None)
        _lib.check(_lib.lib().yeti_asm_run_mass_coo(self._h, C.c_void_p(M.data_ptr()),
                                                    C.c_void_p(r.data_ptr()) if r is not None else None,
                                                    C.c_void_p(c.data_ptr()) if c is not None else None,
                                                    _stream_handle(stream)))
        return M, r, c

    def run_vib_grad(self, vect, vals, gradV=None, stream=None):
        """Natural-frequency shape gradient (gradvibration_fullAN.f:22-189): vect holds nb_frq
        eigenvectors (nb_frq, 3, nb_cp) (numpy or CUDA; column-major like the f2py argument, or
        any array of that shape), vals their eigenvalues. Returns the CUDA tensor of
        d lambda / d P in the same column-major layout, viewed as (nb_frq, 3, nb_cp)."""
        vals = np.ascontiguousarray(np.asarray(vals, np.float64).reshape(-1))
        nf = vals.size
        if isinstance(vect, torch.Tensor):
            v = vect.detach().to(device="cuda", dtype=torch.float64)
        else:
            v = torch.as_tensor(np.asarray(vect, np.float64), device="cuda")
        if tuple(v.shape) != (nf, 3, self.nb_cp):
            raise ValueError("vect must be (nb_frq, 3, nb_cp) = (%d, 3, %d), got %s" % (nf, self.nb_cp, tuple(v.shape)))
        vf = v.permute(2, 1, 0).contiguous()           # column-major (nb_frq, 3, nb_cp) in memory
        if gradV is None:
            gradV = torch.empty((self.nb_cp, 3, nf), dtype=torch.float64, device="cuda")
        if gradV.dtype != torch.float64 or not gradV.is_cuda or gradV.numel() != nf * 3 * self.nb_cp:
            raise ValueError("gradV must be a CUDA float64 tensor of nb_frq * 3 * nb_cp entries")
        _lib.check(_lib.lib().yeti_asm_run_vib_grad(self._h, C.c_void_p(vf.data_ptr()), vals.ctypes.data_as(C.c_void_p),
                                                    nf, C.c_void_p(gradV.data_ptr()), _stream_handle(stream)))
        if stream is not None:
            vf.record_stream(stream)   # the temporary stays allocated until that stream has used it
        return gradV.reshape(self.nb_cp, 3, nf).permute(2, 1, 0)

    def grad_lin_elast_work(self, sol, adj, compute_wint=True, compute_wext=True, stream=None):
        """(gradWint, gradWext), CUDA float64 tensors of nadj x 3 nb_cp (adjoint-major, dof
        3 I + d fastest), for the dof-interleaved state ``sol`` (3 nb_cp) and adjoints ``adj``
        (nadj x 3 nb_cp): -adj_k^T dK/dX sol and adj_k^T dF/dX with the reference's external terms
        (centrifugal loads). ``yeti_grad_lin_elast_work`` works on the (nadj, 3, nb_cp)
        column-major layout, which is the transpose of these tensors."""
        n3 = 3 * self.nb_cp
        sol = sol.detach().to(device="cuda", dtype=torch.float64).reshape(-1).contiguous()
        adj = adj.detach().to(device="cuda", dtype=torch.float64).reshape(-1, n3)
        if sol.numel() != n3:
            raise ValueError("sol must hold 3 nb_cp = %d values" % n3)
        nadj = adj.shape[0]
        adj_cm = adj.t().contiguous()                       # (3 nb_cp, nadj) = (nadj, 3, nb_cp) col-major
        gi = torch.empty(n3, nadj, dtype=torch.float64, device="cuda")
        ge = torch.empty(n3, nadj, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().yeti_grad_lin_elast_work(self._h, C.c_void_p(sol.data_ptr()), C.c_void_p(adj_cm.data_ptr()),
                                                       int(nadj), int(bool(compute_wint)), int(bool(compute_wext)),
                                                       C.c_void_p(gi.data_ptr()), C.c_void_p(ge.data_ptr()),
                                                       _stream_handle(stream)))
        return gi.t(), ge.t()

    def strain_energy_grad(self, u, g=None, stream=None):
        """-d(u^T K u)/dX (3 nb_cp, COORDS3D order) for the dof-interleaved CUDA vector u, over the
        active solid elements (any degree triple the stiffness accepts)."""
        u = u.detach().to(device="cuda", dtype=torch.float64).contiguous()
        if u.numel() != 3 * self.nb_cp:
            raise ValueError("u must hold 3 nb_cp = %d values" % (3 * self.nb_cp))
        if g is None:
            g = torch.empty(3 * self.nb_cp, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().yeti_asm_strain_energy_grad(self._h, C.c_void_p(u.data_ptr()), C.c_void_p(g.data_ptr()),
                                                          _stream_handle(stream)))
        if stream is not None:
            u.record_stream(stream)
        return g

    def volume_grad(self, stream=None):
        """(V, dV/dX): volume of the active solid elements (0-d CUDA tensor) and its derivative
        with respect to the control points (3 nb_cp, COORDS3D order)."""
        vol = torch.empty((), dtype=torch.float64, device="cuda")
        g = torch.empty(3 * self.nb_cp, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().yeti_asm_volume_grad(self._h, C.c_void_p(vol.data_ptr()), C.c_void_p(g.data_ptr()),
                                                   _stream_handle(stream)))
        return vol, g

    def _check_out(self, K, r, c, F):
        if K.dtype != torch.float64 or not K.is_cuda or not K.is_contiguous() or K.numel() < self.nb_data:
            raise ValueError("Kdata must be a contiguous CUDA float64 tensor with >= %d entries" % self.nb_data)
        for t in (r, c):
            if t is not None and (t.dtype != torch.int32 or not t.is_cuda or t.numel() < self.nb_data):
                raise ValueError("Krow/Kcol must be CUDA int32 tensors with >= %d entries" % self.nb_data)
        if F is not None and (F.dtype != torch.float64 or not F.is_cuda or F.numel() < self.nb_dof_tot):
            raise ValueError("F must be a CUDA float64 tensor with >= %d entries" % self.nb_dof_tot)
