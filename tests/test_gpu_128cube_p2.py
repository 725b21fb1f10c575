"""The precision workload at its size: the 128^3-element, degree-2 cantilever (2,097,152
elements, 6,591,000 dofs).

The FP64 oracle is the assembled matrix: the device COO (6.96e9 entries) turned into the CSR of
K + K^T, solved in FP64 (fast-diagonalisation PCG whose every product is the assembled CSR) to
1e-10 or its FP64 rounding floor. The mixed-precision solve (FP32 matrix-free apply, FP32
preconditioner and vectors, FP64 residual replacement with the FP64 matrix-free apply) must give
a displacement and a compliance J = F . u within 1e-8 relative of the oracle's; its own FP64
residual against the assembled matrix is checked as well.
"""
import numpy as np
import pytest

from yeti_amd.geometry import make_box_patch

pytestmark = pytest.mark.gpu


def test_128cube_p2_mixed_precision_against_assembled_fp64(gpu):
    import torch
    from yeti_amd.assembly import Assembler
    from yeti_amd.matfree import MatrixFree
    from yeti_amd.serving import solve_session
    from yeti_amd.solver import StiffnessCSR, free_mask_from_args
    print("[case] 128^3 p=2: mixed solve against the assembled FP64 solve", flush=True)
    bp = make_box_patch(nel=(128, 128, 128), p=2)
    asm = Assembler(bp.args)
    F = asm.run_rhs()
    K = torch.empty(asm.nb_data, dtype=torch.float64, device="cuda")
    asm.run_values(K)
    csr = StiffnessCSR(asm)
    vals = csr.values(K)
    del K
    torch.cuda.empty_cache()
    print("  CSR of K + K^T: %d rows, %d nonzeros" % (csr.nrows, csr.nnz), flush=True)
    mask = free_mask_from_args(bp.args)
    maskd = mask.to(torch.float64)
    bnorm = float((F * maskd).norm())
    mf = MatrixFree(bp.args)
    u0, st0 = solve_session(mf, F, mask, precision=64, rtol=1e-10, csr=csr, csr_vals=vals, max_iters=20000)
    r0 = float(((F - csr.spmv(vals, u0)) * maskd).norm()) / bnorm
    floor = float(np.finfo(np.float64).eps) * float((csr.spmv_abs(vals, u0) * maskd).norm()) / bnorm
    print("  FP64 assembled: %d iterations, relres %.2e (recomputed %.2e), floor %.2e, at floor %d"
          % (st0["iters"], st0["relres"], r0, floor, st0["floor"]), flush=True)
    assert not st0["failed"] and (st0["relres"] <= 1e-10 or st0["floor"])
    assert r0 <= max(1e-10, 20 * floor)
    u1, st1 = solve_session(mf, F, mask, precision=32, rtol=1e-10, max_iters=20000)
    du = float((u1 - u0).norm() / u0.norm())
    J0, J1 = float(torch.dot(F, u0)), float(torch.dot(F, u1))
    dJ = abs(J1 - J0) / abs(J0)
    r1 = float(((F - csr.spmv(vals, u1)) * maskd).norm()) / bnorm
    print("  mixed: %d iterations, %d replacements, |du|/|u| %.2e, |dJ|/J %.2e, assembled residual %.2e"
          % (st1["iters"], st1["updates"], du, dJ, r1), flush=True)
    assert not st1["failed"]
    assert du <= 1e-8 and dJ <= 1e-8
    assert r1 <= max(1e-9, 50 * floor)
    mf.close()
    csr.close()
    asm.close()
    del vals, u0, u1, F
    torch.cuda.empty_cache()
