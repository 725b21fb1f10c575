"""GPU parity of the compliance shape gradient (MatrixFree.strain_energy_grad,
Assembler.load_shape_grad and the torch.autograd.Function around them).

Oracles (tests/golden/grad_4x2x2_p2.npz and unit_cube_p1.npz, written from YETI's compiled
Fortran by tests/golden/make_golden.py):
* -a^T (dK/dX) u is bilinear; the kernel computes q(v) = -v^T (dK/dX) v, so
  (q(u + a) - q(u - a)) / 4 must equal gradLinElastWork_AN's gradWint for (SOL = u, ADJ = a)
  (optim/gradtotalwork_fullAN.f90:17-300), within 1e-10 of max |gradWint|;
* the full dJ/dX, J = F . u, against central differences of the reference's own J (serial
  assembly + direct solve): within 1e-9 + 1e-6 |g| at step 1e-4 and 1e-7 + 1e-4 |g| at the
  1e-6 step (where the ~1e-14 noise of J dominates);
* invariants: translation (sum over control points of dJ/dX_j = 0) and homogeneity
  (sum X . dJ/dX = 3 J: K ~ s, traction F ~ s^2, u ~ s).
"""
import hashlib

import numpy as np
import pytest

from yeti_amd.geometry import make_box_patch

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fixture():
    d = np.load("tests/golden/grad_4x2x2_p2.npz")
    bp = make_box_patch(nel=(4, 2, 2), p=2, seed=11)
    geo = sha(np.concatenate([bp.args["COORDS3D"].ravel(order="F"), bp.args["weight"]])) + sha(bp.args["IEN"])
    assert geo == str(d["geometry_sha"]), "generator changed: regenerate the fixture"
    return d, bp


def q_grad(mf, v):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(v.T.reshape(-1))).cuda()
    return mf.strain_energy_grad(t).cpu().numpy().reshape(-1, 3).T


@pytest.mark.parametrize("impl", ["matrix-free", "assembler"])
@pytest.mark.parametrize("name", ["grad_4x2x2_p2", "unit_cube_p1"])
def test_strain_energy_grad_matches_reference(gpu, name, impl):
    from yeti_amd.assembly import Assembler
    from yeti_amd.matfree import F64, MatrixFree
    print("[case] strain energy gradient vs gradLinElastWork_AN", name, impl, flush=True)
    if name == "unit_cube_p1":
        from tests.golden.make_golden import unit_cube_args
        args = unit_cube_args()
        d = np.load("tests/golden/unit_cube_p1.npz")
    else:
        d, bp = fixture()
        args = bp.args
    mf = MatrixFree(args, precisions=F64) if impl == "matrix-free" else Assembler(args)
    SOL, ADJ, gi = d["SOL"], d["ADJ"], d["gradWint"]
    for k in range(ADJ.shape[0]):
        g = 0.25 * (q_grad(mf, SOL + ADJ[k]) - q_grad(mf, SOL - ADJ[k]))
        err = np.abs(g - gi[k]).max()
        print("  adjoint %d: max |gradWint| %.3e, max err %.3e" % (k, np.abs(gi[k]).max(), err))
        assert err <= 1e-10 * np.abs(gi[k]).max()


def test_compliance_gradient_matches_finite_differences(gpu):
    import torch
    from yeti_amd.gradient import CompliancePatch, compliance
    print("[case] compliance gradient vs central differences", flush=True)
    d, bp = fixture()
    for solver in ("csr", "mf"):
        patch = CompliancePatch(bp.args, solver=solver, rtol=1e-13)
        X = patch.coords0.clone().cuda().requires_grad_(True)
        J = compliance(X, patch)
        J.backward()
        g = X.grad.detach().cpu().numpy().reshape(-1, 3).T
        print("  %s: J %.15e (reference %.15e), %d iterations" % (solver, J.item(), float(d["J"]), patch.iters))
        assert abs(J.item() - float(d["J"])) <= 1e-10 * abs(float(d["J"]))
        gsel = g[d["fd_dir"], d["fd_cp"]]
        e4 = np.abs(gsel - d["fd_val_h1e4"])
        e6 = np.abs(gsel - d["fd_val"])
        print("  max |fd(1e-4) - g| %.2e, max |fd(1e-6) - g| %.2e, max |g| %.2e" % (e4.max(), e6.max(), np.abs(g).max()))
        assert np.all(e4 <= 1e-9 + 1e-6 * np.abs(gsel))
        assert np.all(e6 <= 1e-7 + 1e-4 * np.abs(gsel))
        # the stiffness part is twice YETI's internal term for the adjoint u / 2
        u = torch.from_numpy(np.ascontiguousarray(d["SOL"].T.reshape(-1))).cuda()
        gint = patch.mf.strain_energy_grad(u).cpu().numpy().reshape(-1, 3).T
        assert np.abs(gint - 2 * d["gradWint"][0]).max() <= 1e-10 * np.abs(d["gradWint"][0]).max() * 2
        # invariants
        X0 = patch.coords0.numpy().reshape(-1, 3).T
        assert np.abs(g.sum(axis=1)).max() <= 1e-9 * np.abs(g).max() * g.shape[1]
        hom = float((X0 * g).sum())
        assert abs(hom - 3 * J.item()) <= 1e-8 * abs(J.item())


@pytest.mark.parametrize("name", ["degrees (2,3,1), pressure", "two patches, centrifugal"])
def test_compliance_gradient_general_degrees(gpu, name):
    """Models the matrix-free operator does not cover (degrees (2, 3, 1) with pressure on a face;
    two patches with a centrifugal load): CompliancePatch takes the assembler's element-wise
    strain-energy gradient. Checked against central differences of its own J along random
    directions (within 1e-6 of |g| |d|, step 1e-5 of the model size) and homogeneity about the
    origin, which the rotation axis passes through: sum X . dJ/dX = 3 J for the pressure
    (K ~ s, F ~ s^2) and 7 J for the centrifugal load (F ~ s^4); translation invariance for the
    pressure."""
    import os
    import torch
    from yeti_amd.gradient import CompliancePatch, compliance
    from yeti_amd.model import IGAparametrization
    print("[case] compliance gradient,", name, flush=True)
    if name.startswith("degrees"):
        m = IGAparametrization(filename=os.path.join(os.path.dirname(__file__), "golden", "models", "beam-dist"))
        m.refine(np.array((1, 1, 2)), np.array((1, 2, 0)))
    else:
        m = _body_load_models()["two patches, centrifugal"]()
    patch = CompliancePatch(m.get_inputs4system_elemStorage(), solver="csr", rtol=1e-14)
    if name.startswith("degrees"):
        assert patch.mf is None
    X0 = patch.coords0.clone().cuda()
    X = X0.clone().requires_grad_(True)
    J = compliance(X, patch)
    J.backward()
    g = X.grad.detach().cpu().numpy()
    size = float(np.ptp(X0.cpu().numpy().reshape(-1, 3), axis=0).max())
    rng = np.random.default_rng(3)
    for k in range(3):
        dX = rng.standard_normal(g.size)
        h = 1e-5 * size
        t = torch.from_numpy(dX).cuda()
        Jp = compliance(X0 + h * t, patch).item()
        Jm = compliance(X0 - h * t, patch).item()
        fd, an = (Jp - Jm) / (2 * h), float(g @ dX)
        print("  direction %d: fd %.12e, gradient %.12e" % (k, fd, an))
        assert abs(fd - an) <= 1e-6 * np.linalg.norm(g) * np.linalg.norm(dX)
    G = g.reshape(-1, 3)
    if name.startswith("degrees"):
        assert np.abs(G.sum(axis=0)).max() <= 1e-9 * np.abs(G).max() * G.shape[0]
    hom = float((X0.cpu().numpy() * g).sum())
    assert abs(hom - (3 if name.startswith("degrees") else 7) * J.item()) <= 1e-8 * abs(J.item())


MODELS = "tests/golden/models"


def _body_load_models():
    from yeti_amd.model import IGAparametrization

    def unit_cube():
        return IGAparametrization(filename=MODELS + "/grad_solid/unit_cube_U1")

    def volume_force():
        m = IGAparametrization(filename=MODELS + "/volume_force_3D")
        m.refine(np.array([1, 1, 1]), np.array([1, 1, 1]))
        return m

    def centrif_beam():
        m = IGAparametrization(filename=MODELS + "/load_combination/beam-centrif-dloadF2-dloadF4")
        m.refine(np.array([1, 1, 2]), np.array([1, 1, 1]))
        return m

    def two_patches():
        m = IGAparametrization(filename=MODELS + "/centrif_two_patches/centrif_U1_U1_1load")
        m.refine(np.full((3, 2), 1, np.intp), np.full((3, 2), 1, np.intp))
        return m
    return {"unit cube, centrifugal": unit_cube, "volume force (102)": volume_force,
            "beam, centrifugal + pressure": centrif_beam, "two patches, centrifugal": two_patches}


@pytest.mark.parametrize("name", ["unit cube, centrifugal", "volume force (102)", "beam, centrifugal + pressure",
                                  "two patches, centrifugal"])
def test_load_shape_grad_with_body_loads(gpu, name):
    """g = 2 u^T dF/dX with body loads (101 centrifugal, 102 constant; with face loads on the
    beam) against central differences of u . F(X) from the oracle's assembly along random
    directions (step 1e-5 of the model size, within 1e-6 of |g| |d|)."""
    import torch
    from oracle import oracle as O
    from yeti_amd._abi import args_to_dict
    from yeti_amd.assembly import Assembler
    print("[case] load shape gradient with body loads:", name, flush=True)
    d = args_to_dict(_body_load_models()[name]().get_inputs4system_elemStorage())
    X = np.asarray(d["COORDS3D"], np.float64)
    rng = np.random.default_rng(8)
    u = rng.standard_normal(X.size)
    asm = Assembler(d)
    try:
        g = asm.load_shape_grad(torch.from_numpy(u).cuda()).cpu().numpy()
    finally:
        asm.close()
    assert np.abs(g).max() > 0

    def uF(Xs):
        dd = dict(d)
        dd["COORDS3D"] = np.asfortranarray(Xs)
        return float(u @ O.oracle_sys_linmat([], kwargs={k.lower(): v for k, v in dd.items()})[3])
    h = 1e-5 * float(np.ptp(X, axis=1).max())
    for k in range(2):
        dX = rng.standard_normal(X.shape)
        fd = 2.0 * (uF(X + h * dX) - uF(X - h * dX)) / (2 * h)
        an = float(g @ dX.T.reshape(-1))
        print("  direction %d: fd %.12e, gradient %.12e" % (k, fd, an))
        assert abs(fd - an) <= 1e-6 * np.linalg.norm(g) * np.linalg.norm(dX)


def test_compliance_gradient_centrifugal_unit_cube(gpu):
    """benchs/grad_solid/test_grad_solid_unit_cube_centrif.py: the analytical compliance gradient
    of the centrifugally loaded unit cube against finite differences (the bench's rtol 1e-5),
    here of the oracle's J = F . K^-1 F (scipy solve, step 1e-6) for every coordinate of every
    control point; J itself within 1e-10 of the oracle's."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from oracle import oracle as O
    from yeti_amd._abi import args_to_dict
    from yeti_amd.gradient import CompliancePatch, compliance
    print("[case] compliance gradient, centrifugal unit cube", flush=True)
    m = _body_load_models()["unit cube, centrifugal"]()
    d = args_to_dict(m.get_inputs4system_elemStorage())
    patch = CompliancePatch(d, solver="csr", rtol=1e-14)
    X = patch.coords0.clone().cuda().requires_grad_(True)
    J = compliance(X, patch)
    J.backward()
    g = X.grad.detach().cpu().numpy()
    idof = m.ind_dof_free[:m.nb_dof_free] - 1
    n = m.nb_dof_tot

    def J_oracle(Xflat):
        dd = dict(d)
        dd["COORDS3D"] = np.asfortranarray(Xflat.reshape(-1, 3).T)
        K, r, c, F = O.oracle_sys_linmat([], kwargs={k.lower(): v for k, v in dd.items()})
        Ks = sp.coo_matrix((K, (r, c)), shape=(n, n)).tocsc()
        Kt = (Ks + Ks.T)[idof, :][:, idof]
        return float(F[idof] @ spl.spsolve(Kt, F[idof]))
    X0 = patch.coords0.numpy()
    J0 = J_oracle(X0)
    assert abs(J.item() - J0) <= 1e-10 * abs(J0)
    h = 1e-6
    fd = np.array([(J_oracle(X0 + h * e) - J_oracle(X0 - h * e)) / (2 * h) for e in np.eye(X0.size)])
    print("  max |g - fd| %.3e, max |g| %.3e" % (np.abs(g - fd).max(), np.abs(g).max()))
    assert np.all(np.abs(g - fd) <= 1e-5 * np.abs(fd) + 1e-6 * np.abs(fd).max())


@pytest.mark.parametrize("name", ["unit cube", "two patches", "beam degrees (2,3,1)", "box p=3 ng=3"])
def test_volume_and_gradient(gpu, name):
    """Assembler.volume_grad against (1) the closed form on the unit cube (V = 1; a corner control
    point moves the three faces through it: dV/dX = +-1/4 per direction) and (2) on every model,
    V = e^T M e for the oracle's consistent mass matrix at density 1 and the translation e = e_x
    (within 1e-12), and central differences of that volume along random directions (step 1e-5
    of the model size, within 1e-7 of |g| |d|)."""
    import torch
    from oracle import oracle as O
    from tests.test_gpu_generic import CASES
    from yeti_amd._abi import args_to_dict
    from yeti_amd.assembly import Assembler
    from yeti_amd.model import IGAparametrization
    print("[case] volume and its gradient:", name, flush=True)
    if name == "unit cube":
        args = IGAparametrization(filename=MODELS + "/grad_solid/unit_cube_U1").get_inputs4system_elemStorage()
    elif name == "two patches":
        args = _body_load_models()["two patches, centrifugal"]().get_inputs4system_elemStorage()
    else:
        args = CASES[name]()
    d = args_to_dict(args)
    d["RHO"] = np.ones(int(np.asarray(d["nb_elem_patch"]).size))
    X = np.asarray(d["COORDS3D"], np.float64)
    asm = Assembler(d)
    try:
        v, g = asm.volume_grad()
        V, g = v.item(), g.cpu().numpy()
    finally:
        asm.close()
    e = np.zeros(X.size)
    e[0::3] = 1.0

    def vol(Xs):
        dd = dict(d)
        dd["COORDS3D"] = np.asfortranarray(Xs)
        M, r, c = O.oracle_mass_coo(dd)
        return 2.0 * float(np.sum(M * e[r] * e[c]))
    assert abs(V - vol(X)) <= 1e-12 * abs(V)
    if name == "unit cube":
        assert abs(V - 1.0) <= 1e-14
        G = g.reshape(-1, 3)
        assert np.abs(np.abs(G) - 0.25).max() <= 1e-14
        assert np.array_equal(np.sign(G), np.where(X.T > 0.5, 1.0, -1.0))
    rng = np.random.default_rng(9)
    h = 1e-5 * float(np.ptp(X, axis=1).max())
    for k in range(2):
        dX = rng.standard_normal(X.shape)
        fd = (vol(X + h * dX) - vol(X - h * dX)) / (2 * h)
        an = float(g @ dX.T.reshape(-1))
        assert abs(fd - an) <= 1e-7 * np.linalg.norm(g) * np.linalg.norm(dX), (fd, an)


@pytest.mark.parametrize("name", ["unit_cube_p1", "grad_two_patches_p2"])
def test_gradients_match_reference_fixtures(gpu, name):
    """gradLinElastWork_AN (optim/gradtotalwork_fullAN.f90:17-300) on the centrifugally loaded
    unit cube and on the two-patch model: its external term for the adjoint a is a^T dF/dX =
    load_shape_grad(a) / 2 (the centrifugal term of gradUELMAT1adj), its internal term the
    polarisation (q(u + a) - q(u - a)) / 4 of the assembler's strain-energy gradient; both within
    1e-10 of their maximum."""
    import hashlib
    import torch
    from yeti_amd._abi import args_to_dict
    from yeti_amd.assembly import Assembler
    print("[case] gradients vs gradLinElastWork_AN:", name, flush=True)
    f = np.load("tests/golden/%s.npz" % name)
    if name == "unit_cube_p1":
        from tests.golden.make_golden import unit_cube_args
        d = unit_cube_args()
    else:
        from yeti_amd.model import IGAparametrization
        m = IGAparametrization(filename=MODELS + "/centrif_two_patches/centrif_U1_U1_1load")
        m.refine(np.full((3, 2), 1, np.intp), np.full((3, 2), 1, np.intp))
        d = args_to_dict(m.get_inputs4system_elemStorage())
        md = m.get_model_dict()
        geo = hashlib.sha256(np.ascontiguousarray(np.concatenate([np.asarray(md["COORDS3D"]).ravel(order="F"),
                                                                  np.asarray(md["weight"], np.float64)])).tobytes())
        geo = geo.hexdigest() + hashlib.sha256(np.ascontiguousarray(np.asarray(md["IEN"], np.int64)).tobytes()).hexdigest()
        assert geo == str(f["geometry_sha"]), "refinement changed: regenerate the fixture"
    asm = Assembler(d)
    try:
        def dev(v):
            return torch.from_numpy(np.ascontiguousarray(v.T.reshape(-1))).cuda()

        def host(t):
            return t.cpu().numpy().reshape(-1, 3).T
        SOL, ADJ = f["SOL"], f["ADJ"]
        for k in range(ADJ.shape[0]):
            ge = 0.5 * host(asm.load_shape_grad(dev(ADJ[k])))
            gi = 0.25 * (host(asm.strain_energy_grad(dev(SOL + ADJ[k]))) - host(asm.strain_energy_grad(dev(SOL - ADJ[k]))))
            ee = np.abs(ge - f["gradWext"][k]).max() / np.abs(f["gradWext"][k]).max()
            ei = np.abs(gi - f["gradWint"][k]).max() / np.abs(f["gradWint"][k]).max()
            print("  adjoint %d: relative error external %.2e, internal %.2e" % (k, ee, ei))
            assert ee <= 1e-10 and ei <= 1e-10
    finally:
        asm.close()


def test_coordinate_update_round_trip(gpu):
    """Moving the control points and moving them back reproduces J and the gradient bit for bit."""
    import torch
    from yeti_amd.gradient import CompliancePatch, compliance
    print("[case] coordinate update round trip", flush=True)
    d, bp = fixture()
    patch = CompliancePatch(bp.args, solver="csr", rtol=1e-13)
    X0 = patch.coords0.clone().cuda()
    X = X0.clone().requires_grad_(True)
    J0 = compliance(X, patch)
    J0.backward()
    g0 = X.grad.clone()
    Xs = (X0 * 1.5).requires_grad_(True)
    Js = compliance(Xs, patch)
    assert abs(Js.item() - 1.5 ** 3 * J0.item()) <= 1e-9 * abs(Js.item())
    X1 = X0.clone().requires_grad_(True)
    J1 = compliance(X1, patch)
    J1.backward()
    assert J1.item() == J0.item()
    assert torch.equal(X1.grad, g0)


def test_backward_after_a_later_forward(gpu):
    """J1 = J(X1), then J2 = J(X2), then J1.backward(): the gradient is taken at X1 (the
    coordinates J1 was computed at), bit-identical to a fresh forward + backward at X1, and the
    patch is left at X2 (a following backward of J2 equals a fresh one at X2)."""
    import torch
    from yeti_amd.gradient import CompliancePatch, compliance
    print("[case] backward after a later forward", flush=True)
    d, bp = fixture()
    patch = CompliancePatch(bp.args, solver="csr", rtol=1e-13)
    X0 = patch.coords0.clone().cuda()
    rng = np.random.default_rng(3)
    X2v = X0 + torch.from_numpy(rng.uniform(-0.01, 0.01, X0.numel())).cuda()
    Xa = X0.clone().requires_grad_(True)
    Ja = compliance(Xa, patch)
    Jb = compliance(X2v.clone().requires_grad_(True), patch)
    Ja.backward()
    assert patch.converged
    ref = CompliancePatch(bp.args, solver="csr", rtol=1e-13)
    Xr = X0.clone().requires_grad_(True)
    Jr = compliance(Xr, ref)
    Jr.backward()
    assert Ja.item() == Jr.item()
    assert torch.equal(Xa.grad, Xr.grad)
    assert torch.equal(patch.coords, X2v)
    Xc = X2v.clone().requires_grad_(True)
    Jc = compliance(Xc, ref)
    Jc.backward()
    assert Jb.item() == Jc.item()


def _dropin_cases():
    return ["grad_4x2x2_p2", "unit_cube_p1", "grad_two_patches_p2", "load_combination_pf", "load_combination_fp"]


@pytest.mark.parametrize("name", _dropin_cases())
def test_gradlinelastwork_an_dropin_matches_reference(gpu, name):
    """gradadjointwork.gradlinelastwork_an, called with the reference's argument list
    (_get_inputs4gradTotalwork order), returns the compiled gradLinElastWork_AN's (gradWint,
    gradWext) within 1e-10 of their maximum: traction only (external term 0 in the reference),
    the centrifugal unit cube, the refined two-patch model, and two beams with a nodal pressure
    field and a centrifugal load in both orders with half of the elements active (the reference
    reads the centrifugal axis from the start of load_additionalInfos)."""
    from yeti_amd._abi import args_to_dict
    from yeti_amd.gradadjointwork import gradlinelastwork_an, inputs4gradtotalwork
    print("[case] gradlinelastwork_an drop-in:", name, flush=True)
    act = None
    if name == "grad_4x2x2_p2":
        f, bp = fixture()
        d = dict(bp.args)
        SOL, ADJ, gi_ref, ge_ref = f["SOL"], f["ADJ"], f["gradWint"], f["gradWext"]
    elif name == "unit_cube_p1":
        from tests.golden.make_golden import unit_cube_args
        f = np.load("tests/golden/unit_cube_p1.npz")
        d = unit_cube_args()
        SOL, ADJ, gi_ref, ge_ref = f["SOL"], f["ADJ"], f["gradWint"], f["gradWext"]
    else:
        from yeti_amd.model import IGAparametrization
        if name == "grad_two_patches_p2":
            m = IGAparametrization(filename=MODELS + "/centrif_two_patches/centrif_U1_U1_1load")
            m.refine(np.full((3, 2), 1, np.intp), np.full((3, 2), 1, np.intp))
            f = np.load("tests/golden/grad_two_patches_p2.npz")
            SOL, ADJ, gi_ref, ge_ref = f["SOL"], f["ADJ"], f["gradWint"], f["gradWext"]
        else:
            tag = name.split("_")[-1]
            model = {"pf": "beam-dloadF2-centrif", "fp": "beam-centrif-dloadF2"}[tag]
            m = IGAparametrization(filename=MODELS + "/load_combination/" + model)
            m.refine(np.array([1, 1, 1]), np.array([1, 1, 1]))
            f = np.load("tests/golden/grad_load_combination_p2.npz")
            SOL, ADJ = f[tag + "_SOL"], f[tag + "_ADJ"]
            gi_ref, ge_ref, act = f[tag + "_gradWint"], f[tag + "_gradWext"], f[tag + "_active"]
        d = args_to_dict(m.get_inputs4system_elemStorage())
    gi, ge = gradlinelastwork_an(*inputs4gradtotalwork(d, SOL, ADJ, activeElem=act))
    assert gi.shape == gi_ref.shape and ge.shape == ge_ref.shape and gi.flags.f_contiguous
    ei = np.abs(gi - gi_ref).max() / np.abs(gi_ref).max()
    print("  internal: max %.3e, relative error %.2e" % (np.abs(gi_ref).max(), ei))
    assert ei <= 1e-10
    if np.abs(ge_ref).max() == 0.0:
        assert np.abs(ge).max() == 0.0
    else:
        ee = np.abs(ge - ge_ref).max() / np.abs(ge_ref).max()
        print("  external: max %.3e, relative error %.2e" % (np.abs(ge_ref).max(), ee))
        assert ee <= 1e-10
    # the flags switch either term off (zeros, as the reference's initialisation)
    gi0, ge0 = gradlinelastwork_an(*inputs4gradtotalwork(d, SOL, ADJ, activeElem=act, computeWint=False))
    assert np.abs(gi0).max() == 0.0 and np.array_equal(ge0, ge)
    gi1, ge1 = gradlinelastwork_an(*inputs4gradtotalwork(d, SOL, ADJ, activeElem=act, computeWext=False))
    assert np.abs(ge1).max() == 0.0 and np.array_equal(gi1, gi)


def test_32cube_p2_gradient(gpu):
    """The differentiation workload at its size: 32^3 elements, p = 2 (39,304 control points).
    * gradlinelastwork_an (element-wise gradient kernel, polarised) and the matrix-free MODE 2
      strain-energy gradient against the compiled gradLinElastWork_AN (tests/golden/
      grad_32cube_p2.npz: 8,192 sampled control points x 3 directions x 2 adjoints, and the full
      arrays' Frobenius norms): within 1e-10 of the maximum;
    * translation invariance of the reference's own array (its sum is ~1e-15 of its norm);
    * the compliance gradient dJ/dX (J = F . u, CSR PCG to 1e-12) against central differences of
      J at 8 seeded coordinates: step 1e-4 within 1e-5 max|g|, step 1e-6 within 1e-4 max|g|."""
    import torch
    from tests.golden.make_golden import c5_fields
    from yeti_amd.gradadjointwork import gradlinelastwork_an, inputs4gradtotalwork
    from yeti_amd.gradient import CompliancePatch, compliance
    from yeti_amd.matfree import F64, MatrixFree
    print("[case] 32^3 p=2 gradient", flush=True)
    f = np.load("tests/golden/grad_32cube_p2.npz")
    bp = make_box_patch(nel=(32, 32, 32), p=2)
    geo = sha(np.concatenate([bp.args["COORDS3D"].ravel(order="F"), bp.args["weight"]])) + sha(bp.args["IEN"])
    assert geo == str(f["geometry_sha"]), "generator changed: regenerate the fixture"
    X = np.asarray(bp.args["COORDS3D"], np.float64)
    SOL, ADJ = c5_fields(X)
    cps = f["sample_cp"]
    ref = f["gradWint_sample"]
    scale = f["max_abs"][:, None, None]
    assert np.all(np.abs(f["total"]) <= 1e-12 * f["fro"])
    gi, ge = gradlinelastwork_an(*inputs4gradtotalwork(dict(bp.args), SOL, ADJ))
    err = (np.abs(gi[:, :, cps] - ref) / scale).max()
    fro = np.sqrt((gi ** 2).sum(axis=(1, 2)))
    print("  drop-in: sampled relative error %.2e, Frobenius %s vs %s" % (err, fro, f["fro"]), flush=True)
    assert err <= 1e-10 and np.all(np.abs(fro - f["fro"]) <= 1e-10 * f["fro"])
    assert np.abs(ge).max() == 0.0 == float(f["gradWext_max"])
    mf = MatrixFree(bp.args, precisions=F64)
    q = q_grad(mf, SOL)                       # -u^T dK/dX u = 2 gradWint_0 for ADJ_0 = u / 2
    err_mf = (np.abs(0.5 * q[:, cps] - ref[0]) / scale[0]).max()
    print("  matrix-free MODE 2: sampled relative error %.2e" % err_mf, flush=True)
    assert err_mf <= 1e-10
    mf = None
    patch = CompliancePatch(bp.args, solver="csr", rtol=1e-12)
    X0 = patch.coords0.clone().cuda()
    Xg = X0.clone().requires_grad_(True)
    J = compliance(Xg, patch)
    J.backward()
    assert patch.converged
    g = Xg.grad.detach().cpu().numpy()
    gmax = np.abs(g).max()
    rng = np.random.default_rng(50)
    interior = np.flatnonzero(np.abs(g) > 1e-3 * gmax)
    for h, tol in ((1e-4, 1e-5), (1e-6, 1e-4)):
        worst = 0.0
        for i in rng.choice(interior, 4, replace=False):
            Xp, Xm = X0.clone(), X0.clone()
            Xp[i] += h
            Xm[i] -= h
            fd = (compliance(Xp, patch).item() - compliance(Xm, patch).item()) / (2 * h)
            worst = max(worst, abs(fd - g[i]) / gmax)
        print("  central differences, step %g: max error %.2e of max|g|" % (h, worst), flush=True)
        assert worst <= tol
