"""Natural-frequency shape gradient (gradVibration_AN, optim/gradvibration_fullAN.f:22-189 with
gradUELMASSMAT1): oracle pins on the CPU, the GPU kernel and the vibration.gradvibration_an
drop-in against the oracle.

Pins of the oracle (oracle_grad_vibration, a restatement of the reference's covariant-basis
form):
* YETI's gradVibration_AN compiled from its sources (oracle/_ref): within 1e-13 of max |gradV|
  (degrees 1-3, two modes, with inactive elements);
* central differences of the eigenvalues: K and M of a small jittered cantilever are assembled
  by the oracle, the lowest eigenpairs of K u = lambda M u on the free dofs come from a dense
  generalised eigensolver, and d lambda / d X at a control-point coordinate must match
  (lambda(X + h) - lambda(X - h)) / 2h for the M-normalised modes. A dropped metric term, a sign
  of the mass term or a missing factor 2 on the strain term fails it.
The GPU kernel uses the Cartesian form (W' I - 2 grad(u)^T sig) : grad(R) of the same
derivative, so GPU-vs-oracle agreement also checks the two forms against each other.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from oracle import oracle as O
from yeti_amd.geometry import make_box_patch

RHO = 7.8e-9


def model(nel=(3, 2, 2), p=2, seed=1, inactive=(), L=(10.0, 1.0, 1.0)):
    d = dict(make_box_patch(nel=nel, p=p, seed=seed, rho=RHO, L=L).args)
    if inactive:
        act = np.ones(int(d["nb_elem"]), np.int32)
        act[list(inactive)] = 0
        d["activeElement"] = act
    return d


def random_modes(d, nf, seed=0):
    ncp = int(d["nb_cp"])
    V = np.random.default_rng(seed).standard_normal((nf, 3, ncp))
    L = np.linspace(1.0e6, 4.0e7, nf)
    return V, L


def lowest_modes(d, k):
    """Lowest k eigenpairs of K u = lambda M u on the free dofs, K and M from the oracle; the
    modes come back M-normalised and scattered to (k, 3, nb_cp)."""
    ndof = int(d["nb_dof_tot"])
    K, r, c, _ = O.oracle_sys_linmat([], kwargs=d)
    Ks = sp.coo_matrix((K, (r, c)), shape=(ndof, ndof)).toarray()
    M, mr, mc = O.oracle_mass_coo(d)
    Ms = sp.coo_matrix((M, (mr, mc)), shape=(ndof, ndof)).toarray()
    Kf, Mf = Ks + Ks.T, Ms + Ms.T
    idx = np.asarray(d["ind_dof_free"])[: int(d["nb_dof_free"])] - 1
    w, v = sla.eigh(Kf[np.ix_(idx, idx)], Mf[np.ix_(idx, idx)], subset_by_index=[0, k - 1])
    full = np.zeros((k, ndof))
    full[:, idx] = v.T
    return w, full.reshape(k, -1, 3).transpose(0, 2, 1)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("p", [1, 2, 3])
def test_oracle_grad_vibration_matches_reference(p):
    d = model(p=p, seed=p, inactive=(1, 4))
    V, L = random_modes(d, 2, seed=p)
    go = O.oracle_grad_vibration(d, V, L)
    gr = O.ref_grad_vibration(d, V, L)
    assert np.abs(go - gr).max() <= 1e-13 * np.abs(gr).max()
    # CPs touched only by inactive elements get nothing
    assert np.abs(go).max() > 0


def test_oracle_grad_vibration_central_differences():
    # a rectangular section keeps the two bending modes apart (a square one makes them a
    # near-degenerate pair, whose individual eigenvalues are not smooth enough for differences)
    d = model(nel=(3, 1, 1), p=2, seed=11, L=(10.0, 1.0, 2.5))
    w, V = lowest_modes(d, 3)
    assert np.all(np.diff(w) > 0.5 * w[0])          # well-separated simple eigenvalues
    g = O.oracle_grad_vibration(d, V, w)
    gmax = np.abs(g).max()
    X0 = np.array(d["COORDS3D"], np.float64, copy=True)
    ncp = X0.shape[1]
    # the eigensolver's rounding dominates below h ~ 1e-4 (differences at 1e-5 / 1e-6 land at
    # 1e-6 of max |g|); at 1e-4 the truncation and rounding errors are both ~1e-7
    h = 1e-4
    for cp, comp in ((ncp - 1, 0), (ncp - 1, 2), (ncp // 2, 1), (7, 0), (3, 2)):
        lam = []
        for s in (+1, -1):
            X = X0.copy()
            X[comp, cp] += s * h
            d["COORDS3D"] = np.asfortranarray(X)
            lam.append(lowest_modes(d, 3)[0])
        d["COORDS3D"] = np.asfortranarray(X0)
        fd = (lam[0] - lam[1]) / (2 * h)
        assert np.abs(g[:, comp, cp] - fd).max() <= 2e-6 * gmax, (cp, comp, g[:, comp, cp], fd)
    # planted mistake: dropping the mass term moves the gradient far outside that bar
    assert np.abs(O.oracle_grad_vibration(d, V, 0.0 * w) - g).max() > 1e-2 * gmax


# --------------------------------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("nf", [1, 2, 5])
def test_gpu_vib_grad_matches_oracle(p, nf):
    from yeti_amd.assembly import Assembler
    print("[case] vib grad p=%d nf=%d" % (p, nf), flush=True)
    d = model(nel=(3, 2, 2), p=p, seed=30 + p, inactive=(0, 5))
    V, L = random_modes(d, nf, seed=nf)
    asm = Assembler(d)
    try:
        g = asm.run_vib_grad(V, L).cpu().numpy()
        g2 = asm.run_vib_grad(V, L).cpu().numpy()     # second call reuses the set-up
    finally:
        asm.close()
    go = O.oracle_grad_vibration(d, V, L)
    assert g.shape == go.shape
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()
    assert np.array_equal(g, g2)


@pytest.mark.gpu
def test_gpu_gradvibration_an_dropin():
    """vibration.gradvibration_an with the argument list of get_inputs4gradVibration
    (IGA_parametrization.py:1383-1389), on eigenpairs of the model."""
    from yeti_amd.vibration import gradvibration_an
    print("[case] gradvibration_an", flush=True)
    d = model(nel=(3, 1, 1), p=2, seed=11)
    w, V = lowest_modes(d, 2)
    act = np.ones(int(d["nb_elem"]), np.intp)
    mat = np.vstack([np.asarray(d["MATERIAL_PROPERTIES"])[:2], np.asarray(d["RHO"]).reshape(1, -1)])
    args = [V, w, act, np.ones(3, np.intp), d["COORDS3D"], d["IEN"], d["nb_elem_patch"], d["Nkv"], d["Ukv"],
            d["Nijk"], d["weight"], d["Jpqr"], d["ELT_TYPE"], d["PROPS"], d["JPROPS"], mat[:2, :], mat[2, :],
            d["TENSOR"], 3, d["NBINT"], d["NNODE"], 2, 1, int(d["nb_elem"]), int(d["nb_cp"])]
    g = gradvibration_an(*args)
    go = O.oracle_grad_vibration(d, V, w)
    assert g.shape == (2, 3, int(d["nb_cp"])) and g.flags.f_contiguous
    assert np.abs(g - go).max() <= 1e-11 * np.abs(go).max()
