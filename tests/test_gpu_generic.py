"""The general-degree element path (elem_generic_kernel): degree triples and Gauss orders the
MFMA element kernel is not built for -- unequal degrees per direction, NBINT other than
(p + 1)^3 -- against the oracle (COO indices bit-exact, values and F within 1e-12 of their
maximum), in the serial (packed) and OpenMP (slotted) layouts of the drop-ins, and the
consistent mass matrix, the natural-frequency gradient (run-time-degree instance of the
gradient kernel), the VTU vertex fields and the displacement at parametric points on the same
cases.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from yeti_amd.geometry import make_box_patch
from yeti_amd.model import IGAparametrization

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


def box(p, ng):
    return make_box_patch(nel=(3, 2, 2), p=p, seed=40 + 10 * p + ng, ngauss=ng).f2py_args()


def refined(name, deg, ref):
    m = IGAparametrization(filename=os.path.join(MODELS, name))
    m.refine(np.array(ref), np.array(deg))
    return m.get_inputs4system_elemStorage()


CASES = {
    "box p=2 ng=4": lambda: box(2, 4),
    "box p=1 ng=3": lambda: box(1, 3),
    "box p=3 ng=3": lambda: box(3, 3),
    "beam degrees (2,3,1)": lambda: refined("beam-dist", (1, 2, 0), (1, 1, 2)),
    "beam degrees (1,1,2)": lambda: refined("load_combination/beam-centrif-dloadF2-dloadF4", (0, 0, 1), (0, 0, 0)),
    "beam degrees (3,2,2)": lambda: refined("volume_force_3D", (2, 1, 1), (1, 1, 1)),
}


def test_cases_take_the_general_path():
    """Every case here is one the MFMA kernel is not built for (equal degrees, p + 1 points)."""
    for name, make in CASES.items():
        a = make()
        from yeti_amd._abi import args_to_dict
        d = args_to_dict(a)
        p = np.asarray(d["Jpqr"]).reshape(-1)[:3]
        nbint = int(np.asarray(d["NBINT"]).reshape(-1)[0])
        ng = int(round(nbint ** (1.0 / 3.0)))
        assert not (p[0] == p[1] == p[2] and ng == p[0] + 1), name
        assert np.prod(p + 1) <= 64, name


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_general_degrees_match_oracle(name):
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static
    from yeti_amd.stiffmtrx_elemstorage_omp import sys_linmat_lindef_static_omp
    from yeti_amd._abi import args_to_dict
    from oracle.oracle import YETI_LAYOUT_SLOTTED
    print("[case] general degrees: %s" % name, flush=True)
    args = CASES[name]()
    K, r, c, F = sys_linmat_lindef_static(*args)
    Ko, ro, co, Fo = O.oracle_sys_linmat(args)
    assert np.array_equal(r, ro) and np.array_equal(c, co)
    assert np.abs(K - Ko).max() <= 1e-12 * np.abs(Ko).max()
    assert np.abs(F - Fo).max() <= 1e-12 * max(np.abs(Fo).max(), 1e-300)
    kw = {k.lower(): v for k, v in args_to_dict(args).items()}
    K2, r2, c2, F2 = sys_linmat_lindef_static_omp(**kw)
    Ks, rs, cs, _ = O.oracle_sys_linmat([], layout=YETI_LAYOUT_SLOTTED, kwargs=kw)
    assert np.array_equal(r2, rs) and np.array_equal(c2, cs)
    assert np.abs(K2 - Ks).max() <= 1e-12 * np.abs(Ks).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_general_degrees_mass_match_oracle(name):
    from yeti_amd._abi import args_to_dict
    from yeti_amd.assembly import Assembler
    print("[case] general degrees mass: %s" % name, flush=True)
    d = args_to_dict(CASES[name]())
    d["RHO"] = np.full(int(np.asarray(d["nb_elem_patch"]).size), 7.8)
    asm = Assembler(d)
    try:
        M, r, c = asm.run_mass()
        Mo, ro, co = O.oracle_mass_coo(d)
        n = Mo.size
        assert np.array_equal(r[:n].cpu().numpy(), ro) and np.array_equal(c[:n].cpu().numpy(), co)
        assert np.abs(M[:n].cpu().numpy() - Mo).max() <= 1e-12 * np.abs(Mo).max()
    finally:
        asm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_general_degrees_vib_grad_match_oracle(name):
    from yeti_amd._abi import args_to_dict
    from yeti_amd.assembly import Assembler
    print("[case] general degrees vibration gradient: %s" % name, flush=True)
    d = args_to_dict(CASES[name]())
    d["RHO"] = np.full(int(np.asarray(d["nb_elem_patch"]).size), 7.8e-9)
    ncp = int(np.asarray(d["COORDS3D"]).shape[1])
    V = np.random.default_rng(5).standard_normal((5, 3, ncp))
    L = np.linspace(1e6, 4e7, 5)
    asm = Assembler(d)
    try:
        g = asm.run_vib_grad(V, L).cpu().numpy()
    finally:
        asm.close()
    go = O.oracle_grad_vibration(d, V, L)
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()


def _energy(d, u):
    """u^T K u from the oracle's packed COO (upper blocks, diagonal halved: K = C + C^T)."""
    K, r, c, _ = O.oracle_sys_linmat([], kwargs={k.lower(): v for k, v in d.items()})
    return 2.0 * float(np.sum(K * u[r] * u[c]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_general_degrees_strain_energy_grad(name):
    """-d(u^T K u)/dX of Assembler.strain_energy_grad against (1) the oracle's natural-frequency
    gradient with eigenvalue 0 times the mode mass u^T M u (that gradient divides by it), within
    1e-11 of the maximum, and (2) central differences of the oracle's u^T K(X) u along random
    directions (step 1e-5 of the model size, within 1e-6 of |g| |d|)."""
    import torch
    from yeti_amd._abi import args_to_dict
    from yeti_amd.assembly import Assembler
    print("[case] general degrees strain energy gradient: %s" % name, flush=True)
    d = args_to_dict(CASES[name]())
    d["RHO"] = np.full(int(np.asarray(d["nb_elem_patch"]).size), 7.8)
    X = np.asarray(d["COORDS3D"], dtype=np.float64)
    ncp = X.shape[1]
    V = np.random.default_rng(6).standard_normal((1, 3, ncp))
    u = np.ascontiguousarray(V[0].T.reshape(-1))
    asm = Assembler(d)
    try:
        g = asm.strain_energy_grad(torch.from_numpy(u).cuda()).cpu().numpy()
    finally:
        asm.close()
    M, rm, cm = O.oracle_mass_coo(d)
    mass = 2.0 * float(np.sum(M * u[rm] * u[cm]))
    go = -O.oracle_grad_vibration(d, V, np.zeros(1))[0].T.reshape(-1) * mass
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()
    size = float(np.ptp(X, axis=1).max())
    rng = np.random.default_rng(7)
    for k in range(2):
        dX = rng.standard_normal(X.shape)
        h = 1e-5 * size
        dp, dm = dict(d), dict(d)
        dp["COORDS3D"] = np.asfortranarray(X + h * dX)
        dm["COORDS3D"] = np.asfortranarray(X - h * dX)
        fd = -(_energy(dp, u) - _energy(dm, u)) / (2 * h)
        an = float(g @ dX.T.reshape(-1))
        assert abs(fd - an) <= 1e-6 * np.linalg.norm(g) * np.linalg.norm(dX), (fd, an)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_general_degrees_postprocessing_match_oracle(name):
    from yeti_amd._abi import args_to_dict
    from yeti_amd.postproc import evaldisp_points, svars_solid
    print("[case] general degrees post-processing: %s" % name, flush=True)
    d = args_to_dict(CASES[name]())
    ncp = int(np.asarray(d["COORDS3D"]).shape[1])
    sol = np.random.default_rng(6).standard_normal((ncp, 3)) * 1e-3
    kw = dict(IEN=d["IEN"], elementsByPatch=d["nb_elem_patch"], Nkv=d["Nkv"], Ukv=d["Ukv"], Nijk=d["Nijk"],
              weight=d["weight"], Jpqr=d["Jpqr"], NNODE=d["NNODE"])
    sv = svars_solid(1, (2, 2, 2), (True, True, True), sol.T, d["COORDS3D"], MATERIAL_PROPERTIES=d["MATERIAL_PROPERTIES"],
                     **kw)
    so = O.oracle_svars_solid(d, 0, sol, (2, 2, 2), (True, True, True))
    assert sv.shape == so.shape
    assert np.abs(sv[:, :, :6] - so[:, :, :6]).max() <= 1e-13 * np.abs(so[:, :, :6]).max()
    assert np.abs(sv[:, :, 6:] - so[:, :, 6:]).max() <= 1e-11 * np.abs(so[:, :, 6:]).max()
    xi = np.random.default_rng(7).uniform(0.0, 1.0, (256, 3))
    u = evaldisp_points(1, xi, sol.T, **kw)
    uo = O.oracle_evaldisp(1, xi, sol.T, **kw)
    assert np.abs(u - uo).max() <= 1e-13 * np.abs(uo).max()
