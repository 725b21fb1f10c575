"""Compliance J = F . u of a NURBS solid and its shape gradient dJ/dX on the GPU, behind
``torch.autograd.Function``.

Reference: OPTmodelling.compute_compliance_discrete / compute_gradCompliance_AN
(OPT_modelling.py:534-566) with gradLinElastWork_AN (optim/gradtotalwork_fullAN.f90:17-300).
The compliance is self-adjoint, so the adjoint state is u itself:

    dJ/dX = -u^T (dK/dX) u + 2 u^T (dF/dX)

The first term is one pass of the sum-factorised operator with the quadrature tensor
M = 2 H^T sigma - (sigma : H) I (``MatrixFree.strain_energy_grad``); the second is the
face-load term (``Assembler.load_shape_grad``), which YETI's own gradient leaves as a stub
(optim/gradUELMAT_solid.f90:463-477). YETI's J_YETI = 1/2 F . u, so dJ/dX = 2 dJ_YETI/dX.
Design variables are the control-point coordinates themselves (the identity map of
_propagateSensitivityAnalysis, OPT_modelling.py:439-445). Homogeneous Dirichlet conditions only.
"""
from __future__ import annotations

import warnings

import numpy as np
import torch

from ._abi import (YETI_ERR_INVALID_ARGUMENT, YETI_ERR_UNSUPPORTED_DEGREE, YETI_ERR_UNSUPPORTED_ELEMENT,
                   args_to_dict)
from ._lib import YetiError
from .assembly import Assembler
from .matfree import F64, MatrixFree
from .solver import StiffnessCSR, free_mask_from_args


class CompliancePatch:
    """One model resident on the GPU whose control points can move.

    solver: "csr" (device CSR of K + K^T and FP64 Jacobi-PCG) or "mf" (FP64 matrix-free PCG)."""

    def __init__(self, args, solver="csr", rtol=1e-12, maxit=400000):
        d = args if isinstance(args, dict) else args_to_dict(args)
        bcv = np.asarray(d["bc_values"], dtype=np.float64)   # (2, nb_bc): direction, value
        if bcv.size and np.any(bcv.reshape(2, -1)[1] != 0.0):
            raise ValueError("compliance gradient: only homogeneous Dirichlet conditions are supported")
        if solver not in ("csr", "mf"):
            raise ValueError("solver must be 'csr' or 'mf'")
        self.solver = solver
        self.rtol, self.maxit = float(rtol), int(maxit)
        self.asm = Assembler(d)
        # the matrix-free operator covers full element grids of equal degrees with p + 1 points;
        # other models take the element-wise strain-energy gradient of the assembler and the CSR
        # solve
        try:
            self.mf = MatrixFree(d, precisions=F64)
        except YetiError as e:
            model_errors = (YETI_ERR_INVALID_ARGUMENT, YETI_ERR_UNSUPPORTED_ELEMENT, YETI_ERR_UNSUPPORTED_DEGREE)
            if solver == "mf" or e.code not in model_errors:   # device and budget errors propagate
                raise
            self.mf = None
        self.csr = StiffnessCSR(self.asm) if solver == "csr" else None
        self.mask = free_mask_from_args(d)
        self.nb_cp = self.asm.nb_cp
        self.coords0 = torch.from_numpy(np.asarray(d["COORDS3D"], dtype=np.float64).reshape(-1, order="F").copy())
        # coordinates the resident operator was last built at (device, 3 nb_cp)
        self.coords = self.coords0.to("cuda")
        self.iters = 0
        self.relres = 0.0
        self.converged = True

    def set_coords(self, coords):
        c = coords.detach().reshape(-1).to(device="cuda", dtype=torch.float64).contiguous()
        if c.numel() != 3 * self.nb_cp:
            raise ValueError("expected %d coordinates, got %d" % (3 * self.nb_cp, c.numel()))
        self.asm.update_coords(c)
        if self.mf is not None:
            self.mf.update_coords(c)
        self.coords = c.clone()

    def solve(self):
        """(u, F) at the current coordinates; u is 0 on the fixed dofs. A solve that stops at
        maxit above rtol sets ``converged`` False and warns: J and its gradient then come from an
        unconverged u."""
        if self.solver == "csr":
            K, r, c, F = self.asm.run_coo()
            vals = self.csr.values(K)
            u, it, rel = self.csr.pcg(vals, F, self.mask, rtol=self.rtol, maxit=self.maxit)
        else:
            F = torch.empty(3 * self.nb_cp, dtype=torch.float64, device="cuda")
            self.asm.run_rhs(F)
            u, it, rel = self.mf.pcg_f64(F, self.mask, rtol=self.rtol, maxit=self.maxit)
        self.iters, self.relres = it, rel
        self.converged = bool(rel <= self.rtol)
        if not self.converged:
            warnings.warn("compliance solve stopped after %d iterations at relative residual %.3g > rtol %.3g"
                          % (it, rel, self.rtol), RuntimeWarning, stacklevel=2)
        return u, F

    def shape_gradient(self, u):
        """dJ/dX (3 nb_cp, COORDS3D order) for the state u at the current coordinates."""
        gint = self.mf.strain_energy_grad(u) if self.mf is not None else self.asm.strain_energy_grad(u)
        return gint + self.asm.load_shape_grad(u)


class ComplianceFunction(torch.autograd.Function):
    """J(X) = F(X) . u(X) with K(X) u = F(X) on the free dofs; X = control-point coordinates."""

    @staticmethod
    def forward(ctx, coords, patch):
        patch.set_coords(coords)
        u, F = patch.solve()
        J = torch.dot(F, u)
        ctx.patch = patch
        ctx.save_for_backward(u, patch.coords.clone())
        ctx.coords_shape = coords.shape
        return J

    @staticmethod
    def backward(ctx, gJ):
        # the patch may have moved since this forward (a later forward, as in a line search):
        # the gradient is taken at the coordinates this J was computed at, then the patch is put
        # back where it was
        u, X = ctx.saved_tensors
        patch = ctx.patch
        moved = not torch.equal(patch.coords, X)
        if moved:
            current = patch.coords
            patch.set_coords(X)
        try:
            g = patch.shape_gradient(u)
        finally:
            if moved:
                patch.set_coords(current)
        return (gJ * g).reshape(ctx.coords_shape).to(gJ.device), None


def compliance(coords, patch: CompliancePatch):
    """J = F . u at the control points ``coords`` (CUDA float64, nb_cp x 3 or 3 nb_cp, COORDS3D
    order), differentiable with respect to ``coords``."""
    return ComplianceFunction.apply(coords, patch)
