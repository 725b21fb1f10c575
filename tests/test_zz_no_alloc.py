"""yeti_pcg_csr_noalloc (the solve of yeti_pcg_csr with its work vectors in the library's
file-scope device scratch) against yeti_pcg_csr, both called through the C-ABI: the same
kernels in the same order, so the solutions, iteration counts and residuals are bit-identical.
Two calls in a row, the second on a smaller system (the scratch keeps the first call's
contents), then a system beyond the scratch maximum, which is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from yeti_amd.geometry import make_box_patch

pytestmark = pytest.mark.gpu


def _solve(entry, csr, vals, F, mask, rtol=1e-10, maxit=20000):
    import torch
    from yeti_amd import _lib
    x = torch.full((csr.nrows,), 7.0, dtype=torch.float64, device="cuda")   # overwritten by the solve
    it = C.c_int32(0)
    rel = C.c_double(0.0)
    rc = getattr(_lib.lib(), entry)(csr.nrows, C.c_void_p(csr.rowptr_ptr), C.c_void_p(csr.colind_ptr),
                                    C.c_void_p(vals.data_ptr()), C.c_void_p(F.data_ptr()),
                                    C.c_void_p(mask.data_ptr()), C.c_void_p(x.data_ptr()), float(rtol), int(maxit),
                                    10, C.byref(it), C.byref(rel), None)
    torch.cuda.synchronize()
    return rc, x, it.value, rel.value


def test_pcg_noalloc_matches_pcg(gpu):
    import torch
    from yeti_amd import _lib
    from yeti_amd.assembly import Assembler
    from yeti_amd.solver import StiffnessCSR, free_mask_from_args
    for nel, p in (((12, 4, 4), 2), ((4, 2, 2), 3)):
        print("[case] pcg no-alloc vs pcg:", nel, "p =", p, flush=True)
        bp = make_box_patch(nel=nel, p=p, seed=9)
        asm = Assembler(bp.args)
        K, r, c, F = asm.run_coo()
        csr = StiffnessCSR(asm)
        vals = csr.values(K)
        mask = free_mask_from_args(bp.args)
        rc0, x0, it0, rel0 = _solve("yeti_pcg_csr", csr, vals, F, mask)
        rc1, x1, it1, rel1 = _solve("yeti_pcg_csr_noalloc", csr, vals, F, mask)
        print("  %d rows: %d iterations, relative residual %.3e" % (csr.nrows, it1, rel1), flush=True)
        assert rc0 == 0 and rc1 == 0, _lib.lib().yeti_last_error()
        assert it0 == it1 and rel0 == rel1 and rel1 <= 1e-10
        assert torch.equal(x0, x1)
        csr.close()
        asm.close()


def test_pcg_noalloc_refuses_oversize_systems(gpu):
    from yeti_amd import _lib
    from yeti_amd._abi import YETI_ERR_SIZE_OVERFLOW
    L = _lib.lib()
    d = C.c_void_p(16)   # never dereferenced: the size check comes first
    it = C.c_int32(0)
    rel = C.c_double(0.0)
    rc = L.yeti_pcg_csr_noalloc((1 << 20) + 1, d, d, d, d, d, d, 1e-10, 10, 10, C.byref(it), C.byref(rel), None)
    assert rc == YETI_ERR_SIZE_OVERFLOW
    assert "scratch maximum" in L.yeti_last_error().decode()
