"""Device CSR of K_side + K_side^T and the FP64 Jacobi-PCG on it.

``StiffnessCSR`` mirrors what the reference's callers do after ``sys_linmat_lindef_static``:
``(coo_matrix(...).tocsc() + transpose)`` (overlay/iga_model.py:391-395) and the solve on the
free dofs (benchs/volume_force/test_volume_force_3d.py:67-79), with the pattern built once on
the GPU and values gathered deterministically from the COO values.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .assembly import Assembler, _stream_handle


class StiffnessCSR:
    def __init__(self, asm: Assembler, stream=None):
        self.asm = asm
        h = C.c_void_p()
        _lib.check(_lib.lib().yeti_csr_create(asm._h, _stream_handle(stream), C.byref(h)))
        self._h = h
        self.nnz = int(_lib.lib().yeti_csr_nnz(h))
        self.nrows = int(_lib.lib().yeti_csr_nrows(h))
        self.rowptr_ptr = _lib.lib().yeti_csr_rowptr(h)
        self.colind_ptr = _lib.lib().yeti_csr_colind(h)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().yeti_csr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def values(self, Kdata, out=None, stream=None):
        if out is None:
            out = torch.empty(max(self.nnz, 1), dtype=torch.float64, device=Kdata.device)
        _lib.check(_lib.lib().yeti_csr_values(self._h, C.c_void_p(Kdata.data_ptr()),
                                              C.c_void_p(out.data_ptr()), _stream_handle(stream)))
        return out

    def spmv(self, vals, x, y=None, stream=None):
        if y is None:
            y = torch.empty(self.nrows, dtype=torch.float64, device=x.device)
        _lib.check(_lib.lib().yeti_csr_spmv(self._h, C.c_void_p(vals.data_ptr()), C.c_void_p(x.data_ptr()),
                                            C.c_void_p(y.data_ptr()), _stream_handle(stream)))
        return y

    def pattern_host(self):
        """(indptr int64, indices int32) copied to the host."""
        rp = np.empty(self.nrows + 1, np.int64)
        ci = np.empty(max(self.nnz, 1), np.int32)
        _lib.check(_lib.lib().yeti_csr_pattern_to_host(self._h, C.c_void_p(rp.ctypes.data),
                                                       C.c_void_p(ci.ctypes.data)))
        return rp, ci[: self.nnz]

    def canonical(self, vals, stream=None):
        """scipy's canonical K_side + K_side^T on the device: (rowptr int64, colind int32, values)
        with the entries that sum to exactly 0 removed, as scipy's sparse add does."""
        rp = torch.empty(self.nrows + 1, dtype=torch.int64, device=vals.device)
        ci = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=vals.device)
        v = torch.empty(max(self.nnz, 1), dtype=torch.float64, device=vals.device)
        n = C.c_int64(0)
        _lib.check(_lib.lib().yeti_csr_drop_zeros(self._h, C.c_void_p(vals.data_ptr()), C.c_void_p(rp.data_ptr()),
                                                  C.c_void_p(ci.data_ptr()), C.c_void_p(v.data_ptr()), C.byref(n),
                                                  _stream_handle(stream)))
        return rp, ci[: n.value], v[: n.value]

    def to_scipy(self, vals, canonical=False, fmt="csr"):
        """The matrix as a scipy sparse matrix; canonical=True drops exact zeros (scipy's form of
        K_side + K_side^T); fmt "csc" gives the reference's type (the matrix is symmetric, so its
        CSC arrays are the CSR ones)."""
        import scipy.sparse as sp
        if canonical:
            rp, ci, v = (t.cpu().numpy() for t in self.canonical(vals))
        else:
            rp, ci = self.pattern_host()
            v = vals[: self.nnz].cpu().numpy()
        if fmt == "csc":
            return sp.csc_matrix((v, ci, rp), shape=(self.nrows, self.nrows))
        return sp.csr_matrix((v, ci, rp), shape=(self.nrows, self.nrows))


    def pcg(self, vals, b, free_mask, rtol=1e-10, maxit=100000, check_every=20, norm="2", stream=None,
            raise_on_maxit=True):
        """Jacobi-PCG on A[free][:, free]; returns (x, iterations, relative residual). norm "2"
        stops at ||r||_2 <= rtol ||b||_2, norm "inf" at max|r| <= rtol max|b| (the stopping rule
        of iga_mf_cg_3d, mf_iga_solver.f90:779,935). With raise_on_maxit False a solve that ends
        at maxit returns its iterate instead of raising."""
        if norm not in ("2", "inf"):
            raise ValueError("norm must be '2' or 'inf'")
        x = torch.zeros(self.nrows, dtype=torch.float64, device=b.device)
        it = C.c_int32(0)
        rel = C.c_double(0.0)
        rc = _lib.lib().yeti_csr_pcg(self._h, C.c_void_p(vals.data_ptr()), C.c_void_p(b.data_ptr()),
                                     C.c_void_p(free_mask.data_ptr()), C.c_void_p(x.data_ptr()), float(rtol),
                                     int(maxit), int(check_every), 2 if norm == "2" else 0, C.byref(it),
                                     C.byref(rel), _stream_handle(stream))
        if rc == _lib.YETI_ERR_NOT_CONVERGED and not raise_on_maxit:
            return x, int(it.value), float(rel.value)
        _lib.check(rc)
        return x, int(it.value), float(rel.value)

    def spmv_abs(self, vals, x, y=None, stream=None):
        """y = |A| |x|: eps ||y|| bounds the FP64 rounding of A x (the residual floor)."""
        if y is None:
            y = torch.empty(self.nrows, dtype=torch.float64, device=x.device)
        _lib.check(_lib.lib().yeti_csr_spmv_abs(self._h, C.c_void_p(vals.data_ptr()), C.c_void_p(x.data_ptr()),
                                                C.c_void_p(y.data_ptr()), _stream_handle(stream)))
        return y


def free_mask_from_args(args_dict, device="cuda"):
    idf = np.asarray(args_dict["ind_dof_free"]).astype(np.int64)
    nfree = int(args_dict["nb_dof_free"])
    ndof = idf.size
    m = np.zeros(ndof, np.uint8)
    m[idf[:nfree] - 1] = 1
    return torch.from_numpy(m).to(device)


def build_and_solve(args, rtol=1e-10, maxit=200000):
    """Assemble, form the CSR and solve the free-dof system on the GPU.
    Returns (u_full (device), iterations, relative residual, handles)."""
    from ._abi import args_to_dict
    d = args if isinstance(args, dict) else args_to_dict(args)
    asm = Assembler(d)
    K, r, c, F = asm.run_coo()
    csr = StiffnessCSR(asm)
    vals = csr.values(K)
    mask = free_mask_from_args(d)
    x, it, rel = csr.pcg(vals, F, mask, rtol=rtol, maxit=maxit)
    return x, it, rel, (asm, csr, K, vals, F, mask)


def build_stiffness_matrix(args):
    """Drop-in for IgaModel.build_stiffness_matrix (overlay/iga_model.py:375-395): the assembly
    and K_side + K_side^T on the GPU, returned as the reference returns them (scipy CSC in
    canonical form, rhs as numpy)."""
    from ._abi import args_to_dict
    d = args if isinstance(args, dict) else args_to_dict(args)
    asm = Assembler(d)
    K, r, c, F = asm.run_coo()
    csr = StiffnessCSR(asm)
    vals = csr.values(K)
    A = csr.to_scipy(vals, canonical=True, fmt="csc")
    return A, F.cpu().numpy()
