"""Drop-in for YETI's f2py module ``reconstructionSOL`` (src/yeti_iga/CMakeLists.txt:83-85).

``reconstruction`` rebuilds the full displacement from the free-dof solution on the GPU
(``yeti_reconstruction``, reconstruction.f:22-135); it is called like the reference, e.g.
``sol, U = reconstruction(**iga.get_inputs4solution(x))`` (overlay/iga_model.py:504).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .assembly import _stream_handle


def reconstruction(u_inv, coords, ind_dof_free, bc_target, bc_target_nbelem, bc_values, couplg_flag=False,
                   nb_bc=None, mcrd=None, nb_cp=None, nb_dof_free=None):
    """Returns ``(sol, U)``: sol (nb_cp, mcrd) in Fortran order, U (nb_cp*mcrd,) dof-interleaved.

    Numpy inputs give numpy outputs; a CUDA ``u_inv`` gives CUDA tensors (the solve's output
    stays on the device). Entries neither free nor prescribed are 0. ``couplg_flag`` and
    ``coords`` are accepted for the reference's signature; its coupling branch is commented out
    (reconstruction.f:76-129), so they do not change the result."""
    _lib.require_device()
    on_dev = isinstance(u_inv, torch.Tensor) and u_inv.is_cuda
    X = np.asarray(coords, dtype=np.float64)
    n_cp = int(nb_cp) if nb_cp is not None else (X.shape[1] if X.ndim == 2 else X.size // 3)
    idf = np.asarray(ind_dof_free.cpu() if isinstance(ind_dof_free, torch.Tensor) else ind_dof_free).astype(np.int64)
    m = int(mcrd) if mcrd is not None else idf.size // max(n_cp, 1)
    n_free = int(nb_dof_free) if nb_dof_free is not None else int(u_inv.numel() if on_dev else np.size(u_inv))
    bn = np.ascontiguousarray(np.asarray(bc_target_nbelem, dtype=np.int64).reshape(-1)[: (nb_bc if nb_bc is not None
                                                                                             else None)], dtype=np.int32)
    n_bc = int(bn.size)
    bv = np.asfortranarray(np.asarray(bc_values, dtype=np.float64).reshape(2, -1, order="F")[:, :n_bc])
    bt = np.asarray(bc_target, dtype=np.int64).reshape(-1)
    ndof = n_cp * m
    # the reference writes through these indices without checks; reject what would be out of range
    if n_free > idf.size or (n_free and (idf[:n_free].min() < 1 or idf[:n_free].max() > ndof)):
        raise ValueError("reconstruction: ind_dof_free has entries outside 1..nb_cp*mcrd")
    need = int(bn.sum())
    if bt.size < need:
        raise ValueError("reconstruction: bc_target holds %d ids, bc_target_nbelem asks for %d" % (bt.size, need))
    off = 0
    for b in range(n_bc):
        if bv[0, b] <= 3:
            t = bt[off:off + bn[b]]
            if t.size and (t.min() < 1 or t.max() > n_cp or bv[0, b] < 1):
                raise ValueError("reconstruction: boundary condition %d targets a CP or direction out of range" % b)
        off += int(bn[b])
    dev = torch.device("cuda")
    u = (u_inv.to(torch.float64) if on_dev else torch.as_tensor(np.asarray(u_inv, np.float64)).to(dev)).contiguous()
    ind = torch.as_tensor(idf[:max(n_free, 1)].astype(np.int32)).to(dev)
    tgt = torch.as_tensor(bt[:max(need, 1)].astype(np.int32) if bt.size else np.zeros(1, np.int32)).to(dev)
    sol = torch.empty(ndof, dtype=torch.float64, device=dev)
    U = torch.empty(ndof, dtype=torch.float64, device=dev)
    bvc = np.ascontiguousarray(bv.T.reshape(-1))    # [dir, value] per condition
    _lib.check(_lib.lib().yeti_reconstruction(
        C.c_void_p(u.data_ptr()), C.c_void_p(ind.data_ptr()), n_free, C.c_void_p(tgt.data_ptr()),
        bn.ctypes.data_as(C.c_void_p), bvc.ctypes.data_as(C.c_void_p), n_bc, m, n_cp,
        C.c_void_p(sol.data_ptr()), C.c_void_p(U.data_ptr()), _stream_handle(None)))
    if on_dev:
        return sol.view(m, n_cp).t(), U     # (nb_cp, mcrd) view of the column-major buffer
    torch.cuda.current_stream().synchronize()
    return np.asfortranarray(sol.cpu().numpy().reshape(n_cp, m, order="F")), U.cpu().numpy()
