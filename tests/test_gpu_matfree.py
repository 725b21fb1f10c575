"""GPU parity of the sum-factorised matrix-free operator and the mixed-precision solves.

Oracle: K = K_side + K_side^T built by scipy from the oracle's COO (the reference's own
post-processing, overlay/iga_model.py:391-395). Bars: FP64 apply within 1e-12 of ||K x||_inf;
FP32 apply within 2e-5 (FP32 arithmetic); 3x3 diagonal blocks within 1e-12; FP64 matrix-free
PCG and FP64-refined FP32 solves reach ||b - K u|| <= 1e-10 ||b|| measured with the oracle's K
(a relative error of the displacement against scipy's direct solve is printed and bounded by
the conditioning of the beam)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from yeti_amd.geometry import make_box_patch

pytestmark = pytest.mark.gpu


def oracle_K(args):
    K, r, c, F = O.oracle_sys_linmat((), kwargs=args)
    n = int(args["nb_dof_tot"])
    A = sp.coo_matrix((K, (r, c)), shape=(n, n)).tocsr()
    return (A + A.T).tocsr(), F


@pytest.mark.parametrize("nel,p", [((5, 4, 3), 2), ((4, 3, 3), 3), ((3, 2, 2), 4), ((4, 3, 2), 1), ((13, 7, 9), 2),
                                   ((3, 5, 10), 4), ((5, 3, 11), 3), ((9, 10, 17), 1)])
def test_mf_apply_matches_assembled(gpu, nel, p):
    import torch
    from yeti_amd.matfree import MatrixFree
    print("[case] mf apply", nel, p, flush=True)
    bp = make_box_patch(nel=nel, p=p, seed=50 + p)
    A, F = oracle_K(bp.args)
    mf = MatrixFree(bp.args)
    x = np.random.default_rng(2).standard_normal(bp.nb_dof_tot)
    y0 = A @ x
    y = mf.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.abs(y - y0).max() <= 1e-12 * np.abs(y0).max()
    y32 = mf.apply(torch.from_numpy(x).float().cuda()).cpu().numpy()
    assert np.abs(y32 - y0).max() <= 2e-5 * np.abs(y0).max()
    # bit-reproducible
    y2 = mf.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(y, y2)
    # diagonal blocks
    D = mf.diag_blocks().cpu().numpy()
    Ad = A.toarray() if bp.nb_dof_tot < 4000 else None
    if Ad is not None:
        for cp in range(0, bp.nb_cp, 7):
            blk = Ad[3 * cp:3 * cp + 3, 3 * cp:3 * cp + 3]
            assert np.abs(D[cp] - blk).max() <= 1e-12 * np.abs(Ad).max()


def test_mf_two_patches(gpu):
    """Two patches joined by a face (the refined centrifugal two-patch bench, p = 2): the apply,
    the diagonal blocks and the strain-energy gradient sum both patches' contributions on the
    shared CPs and start from zero on the others. Against the oracle's assembled K (1e-12), its
    3x3 diagonal blocks, and the assembler's element-wise strain-energy gradient (1e-12),
    starting from output buffers filled with NaN."""
    import os
    import torch
    from yeti_amd._abi import args_to_dict
    from yeti_amd.assembly import Assembler
    from yeti_amd.matfree import MatrixFree
    from yeti_amd.model import IGAparametrization
    print("[case] mf two patches", flush=True)
    m = IGAparametrization(filename=os.path.join(os.path.dirname(__file__), "golden", "models",
                                                 "centrif_two_patches", "centrif_U1_U1_1load"))
    m.refine(np.full((3, 2), 1, np.intp), np.full((3, 2), 1, np.intp))
    d = args_to_dict(m.get_inputs4system_elemStorage())
    A, _ = oracle_K(d)
    mf = MatrixFree(d)
    x = np.random.default_rng(4).standard_normal(m.nb_dof_tot)
    xt = torch.from_numpy(x).cuda()
    y = torch.full_like(xt, float("nan"))
    mf.apply(xt, y)
    y = y.cpu().numpy()
    y0 = A @ x
    assert np.abs(y - y0).max() <= 1e-12 * np.abs(y0).max()
    D = mf.diag_blocks().cpu().numpy()
    Ad = A.toarray()
    for cp in range(m.nb_cp):
        assert np.abs(D[cp] - Ad[3 * cp:3 * cp + 3, 3 * cp:3 * cp + 3]).max() <= 1e-12 * np.abs(Ad).max()
    asm = Assembler(d)
    try:
        g0 = asm.strain_energy_grad(xt).cpu().numpy()
    finally:
        asm.close()
    g = mf.strain_energy_grad(xt).cpu().numpy()
    assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()


def test_mf_solves_cantilever(gpu):
    import torch
    from yeti_amd.matfree import MatrixFree
    from yeti_amd.solver import free_mask_from_args
    bp = make_box_patch(nel=(16, 4, 4), p=2)
    A, F = oracle_K(bp.args)
    free = bp.ind_dof_free0
    u_ref = spla.spsolve(A[free][:, free].tocsc(), F[free])
    mf = MatrixFree(bp.args)
    b = torch.from_numpy(F).cuda()
    mask = free_mask_from_args(bp.args)
    x, it, rel = mf.pcg_f64(b, mask, rtol=1e-10)
    u = x.cpu().numpy()
    res = np.linalg.norm(F[free] - A[free][:, free] @ u[free]) / np.linalg.norm(F[free])
    err = np.linalg.norm(u[free] - u_ref) / np.linalg.norm(u_ref)
    print("[mf pcg f64] it %d rel %.2e true residual %.2e err %.2e" % (it, rel, res, err))
    assert res <= 2e-10 and err <= 1e-6
    xi, outer, inner, reli = mf.solve_ir(b, mask, rtol=1e-10, inner_rtol=1e-4)
    ui = xi.cpu().numpy()
    resi = np.linalg.norm(F[free] - A[free][:, free] @ ui[free]) / np.linalg.norm(F[free])
    erri = np.linalg.norm(ui[free] - u_ref) / np.linalg.norm(u_ref)
    print("[mf ir] outer %d inner %d rel %.2e true residual %.2e err %.2e" % (outer, inner, reli, resi, erri))
    assert resi <= 2e-10 and erri <= 1e-6
    fixed = np.setdiff1d(np.arange(bp.nb_dof_tot), free)
    assert np.all(ui[fixed] == 0.0) and np.all(u[fixed] == 0.0)
