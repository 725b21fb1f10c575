"""GPU parity of the assembly workload at its full size: the 64^3-element, degree-3 cantilever
(262,144 elements, 300,763 control points, 4,857,004,032 COO entries = 77.7 GB with indices).

* Element slices of the device COO (PACKED layout, every element active) against the oracle's
  COO of the same elements (the oracle run with only those elements active, so its packed output
  holds them back to back): indices bit-exact, values within 1e-12 of the largest magnitude. The
  slices are the first 256 elements, the 256 elements around entry offset 2^31 (element 115,904
  holds it: where the reference's OpenMP routine overflows its int32 slot offset,
  build_system_elemStorage_OMP.f90:229; the device offsets are int64) and the last 256.
* The CSR pattern of K + K^T: every row's length against the closed form (3 x the product over
  directions of the CPs within +-p along each grid line) and the total
  9 (67 x 7 - 12)^3 = 858,995,937.
* SpMV: the CSR product with a seeded vector against the same product summed from the COO
  triplets (y[r] += v x[c], y[c] += v x[r]) over all 4.86e9 entries, within 1e-12 relative.
"""
import numpy as np
import pytest

from yeti_amd.geometry import make_box_patch

pytestmark = pytest.mark.gpu

NEL, P = 64, 3
NDE = 9 * 64 * 63 // 2 + 6 * 64          # 18,528 entries per element


def _slices():
    ne = NEL ** 3
    mid = (2 ** 31) // NDE                # 115,904: the element whose COO slots hold entry 2^31
    return [np.arange(0, 256), np.arange(mid - 128, mid + 128), np.arange(ne - 256, ne)]


def test_64cube_p3_full_size_assembly_csr_spmv(gpu):
    import torch
    from oracle import oracle as O
    from yeti_amd.assembly import Assembler
    from yeti_amd.solver import StiffnessCSR
    print("[case] 64^3 p=3 full-size assembly, CSR pattern and SpMV", flush=True)
    bp = make_box_patch(nel=(NEL, NEL, NEL), p=P)
    ne = bp.nb_elem
    asm = Assembler(bp.args)
    assert asm.nb_data == ne * NDE == 4857004032
    K, r, c, F = asm.run_coo()
    torch.cuda.synchronize()
    print("  assembled %d entries" % asm.nb_data, flush=True)

    # element slices against the oracle
    sel = np.concatenate(_slices())
    d = dict(bp.args)
    act = np.zeros(ne, np.int32)
    act[sel] = 1
    d["activeElement"] = act
    d["nb_data"] = float(sel.size * NDE)
    K0, r0, c0, _ = O.oracle_sys_linmat((), kwargs=d)
    kmax = 0.0
    worst = 0.0
    for k, e in enumerate(sel):
        g = slice(int(e) * NDE, (int(e) + 1) * NDE)
        o = slice(k * NDE, (k + 1) * NDE)
        assert np.array_equal(r[g].cpu().numpy(), r0[o]), "row ids of element %d" % e
        assert np.array_equal(c[g].cpu().numpy(), c0[o]), "column ids of element %d" % e
        kv = K[g].cpu().numpy()
        kmax = max(kmax, float(np.abs(K0[o]).max()))
        worst = max(worst, float(np.abs(kv - K0[o]).max()))
    print("  %d sampled elements (offsets up to %.3e): indices exact, max |dK| %.2e of %.2e"
          % (sel.size, float(sel.max()) * NDE, worst, kmax), flush=True)
    assert worst <= 1e-12 * kmax

    # CSR pattern: closed-form row lengths
    csr = StiffnessCSR(asm)
    n = NEL + P
    assert csr.nnz == 9 * (n * (2 * P + 1) - P * (P + 1)) ** 3 == 858995937
    rp, _ = csr.pattern_host()
    i = np.arange(n)
    per_dir = np.minimum(i + P, n - 1) - np.maximum(i - P, 0) + 1
    cp_len = 3 * (per_dir[:, None, None] * per_dir[None, :, None] * per_dir[None, None, :])   # [k, j, i]
    row_len = np.repeat(cp_len.transpose(2, 1, 0).reshape(-1, order="F"), 3)
    assert np.array_equal(np.diff(rp), row_len)

    # SpMV against the COO triplets
    vals = csr.values(K)
    gen = torch.Generator(device="cuda").manual_seed(20261202)
    x = torch.rand(3 * bp.nb_cp, dtype=torch.float64, device="cuda", generator=gen) - 0.5
    y = csr.spmv(vals, x)
    del vals
    yc = torch.zeros_like(x)
    step = 1 << 28
    for s0 in range(0, asm.nb_data, step):
        s1 = min(asm.nb_data, s0 + step)
        rr = r[s0:s1].long()
        cc = c[s0:s1].long()
        kv = K[s0:s1]
        yc.index_add_(0, rr, kv * x[cc])
        yc.index_add_(0, cc, kv * x[rr])
        del rr, cc, kv
    rel = float((y - yc).abs().max() / yc.abs().max())
    print("  SpMV vs COO-summed product: max relative difference %.2e" % rel, flush=True)
    assert rel <= 1e-12
    csr.close()
    asm.close()
    del K, r, c, F, x, y, yc
    torch.cuda.empty_cache()
