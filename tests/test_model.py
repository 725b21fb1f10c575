"""Model files and k-refinement (yeti_amd.model.IGAparametrization): the reference's own bench
models end to end.

Fixtures (tests/golden/models/) are the input files of the reference's benchmarks
benchs/dload/{beam-dist, beam-dist-field, beam-triangle-1, beam-triangle-2, beam-dloadF2F4}.{inp,NB},
benchs/volume_force/volume_force_3D.{inp,NB}, benchs/platevib/plateVolume.{inp,NB} and
benchs/load_combination/inputs_beam/* (in load_combination/), copied as data. Pins:
* refinement is exact: the refined NURBS map equals the coarse one at random parametric points
  (geometry and weights; a wrong knot multiplicity, Greville point or homogeneous/Cartesian mix-up
  moves it);
* the benchmarks' own published checks, with the reference's refinement and tolerances:
  benchs/dload/test_dload_beam_{dist, dist_field, triangular_1, triangular_2}.py (deflexion at
  CP (20, 0, 0) within 2e-2 of the analytic value; pressure U40 on the face of one coarse
  element, mapped to the refined elements on that face, or U44 with a nodal *DISTRIBUTION
  refined as a rational field; the uniform field must give exactly the uniform pressure's
  answer), benchs/volume_force/test_volume_force_3d.py (displacement at xi = (1, 1, 0)
  within 2e-3 of the Abaqus reference; body force U102 with a direction) and
  benchs/platevib/test_platevib.py (the 5 lowest eigenfrequencies of K u = lambda M u within
  1e-3 of the reference's values; stiffness and consistent mass) and
  benchs/load_combination/test_beam_dload_triangle_centrif.py (the right-hand side of two
  nodal-distribution pressures and a centrifugal load is the sum of the single loads, whatever
  the order of declaration, to 1e-12). On the CPU the
  assembly is the oracle's; on the GPU it is the drop-in stack (stiffness on the GPU, the
  callers' scipy solve, reconstruction and evaldisp on the GPU).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import oracle as O
from yeti_amd.model import IGAparametrization

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")
VF3D_REF = np.array([-9.96639E-5, -1.01212, -6.42538])   # benchs/volume_force/test_volume_force_3d.py


# benchs/dload: model -> analytic deflexion at CP (20, 0, 0)
DLOAD = {"beam-dist": -1142.857, "beam-dist-field": -1142.86, "beam-triangle-1": -838.10,
         "beam-triangle-2": -304.76}


def dload_model(name):
    m = IGAparametrization(filename=os.path.join(MODELS, name))
    m.refine(np.array([2, 2, 5]), np.array([1, 1, 1]))
    return m


def beam_dist():
    return dload_model("beam-dist")


def volume_force_3d():
    m = IGAparametrization(filename=os.path.join(MODELS, "volume_force_3D"))
    nb_deg = np.zeros((3, 1), np.intp)
    nb_ref = np.zeros((3, 1), np.intp)
    nb_deg[:, 0] = 1
    nb_ref[:, 0] = (3, 4, 3)
    m.refine(nb_ref, nb_deg)
    return m


def solve(m, K, r, c, F):
    n = m.nb_dof_tot
    Ks = sp.coo_matrix((K, (r, c)), shape=(n, n)).tocsc()
    Kt = Ks + Ks.T
    idof = m.ind_dof_free[:m.nb_dof_free] - 1
    return spl.spsolve(Kt[idof, :][:, idof], F[idof])


def full_sol(m, x):
    U = np.zeros(m.nb_dof_tot)
    U[m.ind_dof_free[:m.nb_dof_free] - 1] = x
    return U.reshape(-1, 3)


def geometry_at(d, xi):
    """x(xi) of patch 1 as the NURBS interpolation of the control-point coordinates."""
    X = np.asarray(d["COORDS3D"], np.float64)
    return O.oracle_evaldisp(1, xi, X, d["IEN"], d["nb_elem_patch"], d["Nkv"], d["Ukv"], d["Nijk"], d["weight"],
                             d["Jpqr"], d["NNODE"])


def test_reader_coarse_models():
    m = IGAparametrization(filename=os.path.join(MODELS, "beam-dist"))
    d = m.get_model_dict()
    assert m.nb_cp == 8 and m.nb_elem == 1 and d["ELT_TYPE"] == "U1"
    assert np.array_equal(np.asarray(d["IEN"]), np.arange(8, 0, -1))
    assert np.allclose(d["MATERIAL_PROPERTIES"].ravel(), [210000.0, 0.3])
    assert list(d["JDLType"]) == [40] and list(d["ADLMAG"]) == [1000.0]
    assert d["nb_bc"] == 3 and m.nb_dof_free == 12
    v = IGAparametrization(filename=os.path.join(MODELS, "volume_force_3D")).get_model_dict()
    assert list(v["JDLType"]) == [102] and np.allclose(v["load_additionalInfos"], [0.0, 0.0, -1.0])
    assert np.allclose(v["RHO"], [7.8e-9])


@pytest.mark.parametrize("build", [beam_dist, volume_force_3d])
def test_refinement_is_exact(build):
    name = build.__name__
    coarse = IGAparametrization(filename=os.path.join(MODELS, "beam-dist" if name == "beam_dist" else "volume_force_3D"))
    fine = build()
    xi = np.random.default_rng(3).uniform(0.0, 1.0, (200, 3))
    xc = geometry_at(coarse.get_model_dict(), xi)
    xf = geometry_at(fine.get_model_dict(), xi)
    assert np.abs(xf - xc).max() <= 1e-12 * np.abs(xc).max()
    d = fine.get_model_dict()
    assert d["nb_elem"] == int(np.prod([np.unique(u).size - 1 for u in fine._Ukv[0]]))
    assert np.asarray(d["NBINT"])[0] == 27


def tip_deflexion(m, sol):
    idx = np.nonzero(np.all(m.coords == np.array([[20.0], [0.0], [0.0]]), axis=0))[0]
    assert idx.size == 1
    return sol[idx[0], 2]


@pytest.mark.parametrize("name", sorted(DLOAD))
def test_bench_dload_oracle(name):
    m = dload_model(name)
    K, r, c, F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())
    assert np.allclose(tip_deflexion(m, full_sol(m, solve(m, K, r, c, F))), DLOAD[name], rtol=2e-2)


def test_uniform_field_equals_uniform_pressure():
    a, b = dload_model("beam-dist"), dload_model("beam-dist-field")
    Fa = O.oracle_sys_linmat(a.get_inputs4system_elemStorage())[3]
    Fb = O.oracle_sys_linmat(b.get_inputs4system_elemStorage())[3]
    assert np.abs(Fa - Fb).max() <= 1e-12 * np.abs(Fa).max()


def test_bench_volume_force_3d_oracle():
    m = volume_force_3d()
    d = m.get_model_dict()
    K, r, c, F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())
    sol = full_sol(m, solve(m, K, r, c, F))
    disp = O.oracle_evaldisp(1, np.array([[1.0, 1.0, 0.0]]), sol.T, d["IEN"], d["nb_elem_patch"], d["Nkv"], d["Ukv"],
                             d["Nijk"], d["weight"], d["Jpqr"], d["NNODE"])[0]
    assert np.linalg.norm(disp - VF3D_REF) / np.linalg.norm(VF3D_REF) < 2e-3


PLATEVIB_REF = np.array([0.2548, 0.6217, 1.9546, 2.3870, 2.8585])   # benchs/platevib/test_platevib.py


def platevib():
    m = IGAparametrization(filename=os.path.join(MODELS, "plateVolume"))
    m.refine(nb_refinementByDirection=np.array([[2, 2, 1]]).T, nb_degreeElevationByDirection=np.array([[1, 1, 1]]).T)
    return m


def frequencies(m, K, r, c, M, mr, mc, k=5):
    n = m.nb_dof_tot
    idof = m.ind_dof_free[:m.nb_dof_free] - 1
    Ks = sp.coo_matrix((K, (r, c)), shape=(n, n)).tocsc()
    Ms = sp.coo_matrix((M, (mr, mc)), shape=(n, n)).tocsc()
    vals, _ = spl.eigsh((Ks + Ks.T)[idof, :][:, idof], k=k, M=(Ms + Ms.T)[idof, :][:, idof], sigma=0.0)
    return np.sqrt(vals) / 2.0 / np.pi


def test_bench_platevib_oracle():
    m = platevib()
    K, r, c, _ = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())
    M, mr, mc = O.oracle_mass_coo(m.get_model_dict())
    assert np.allclose(frequencies(m, K, r, c, M, mr, mc), PLATEVIB_REF, rtol=1e-3)


def test_omp_inputs_match_serial():
    """benchs/omp/test_build_stifmatrix_omp.py's model: the keyword inputs of the OpenMP routine
    carry the same model as the positional ones (oracle, both layouts)."""
    from oracle.oracle import YETI_LAYOUT_PACKED, YETI_LAYOUT_SLOTTED
    m = IGAparametrization(filename=os.path.join(MODELS, "beam-dist"))
    m.refine(np.array([2, 2, 2]), np.array([2, 2, 2]))
    n = m.nb_dof_tot
    K, r, c, F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage(), layout=YETI_LAYOUT_PACKED)
    kw = m.get_inputs4system_elemStorage_OMP(num_threads=2)
    kw.pop("num_threads")
    K2, r2, c2, F2 = O.oracle_sys_linmat([], layout=YETI_LAYOUT_SLOTTED, kwargs=kw)
    A = sp.coo_matrix((K, (r, c)), shape=(n, n)).tocsc()
    B = sp.coo_matrix((K2, (r2, c2)), shape=(n, n)).tocsc()
    assert spl.norm(A - B) / spl.norm(A) < 1e-15
    assert np.linalg.norm(F - F2) / np.linalg.norm(F) < 1e-15


COMBI = os.path.join(MODELS, "load_combination")


def combination_rhs(assemble):
    out = {}
    for name in sorted(set(f[:-4] for f in os.listdir(COMBI) if f.endswith(".inp"))):
        m = IGAparametrization(filename=os.path.join(COMBI, name))
        nb_deg = np.zeros((3, 1), np.intp)
        nb_deg[:, 0] = (0, 0, 1)
        m.refine(np.zeros((3, 1), np.intp), nb_deg)
        out[name] = assemble(m.get_inputs4system_elemStorage())
    return out


def check_combinations(f):
    """The assertions of benchs/load_combination/test_beam_dload_triangle_centrif.py:128-178."""
    close = lambda a, b: np.allclose(a, b, rtol=1e-12)
    assert close(f["beam-dloadF2F4"], f["beam-dloadF2"] + f["beam-dloadF4"])
    assert close(f["beam-dloadF2F4"], f["beam-dloadF4F2"])
    assert close(f["beam-dloadF2-centrif"], f["beam-dloadF2"] + f["beam-centrif"])
    assert close(f["beam-dloadF2-centrif"], f["beam-centrif-dloadF2"])
    assert close(f["beam-dloadF4-centrif"], f["beam-dloadF4"] + f["beam-centrif"])
    assert close(f["beam-dloadF4-centrif"], f["beam-centrif-dloadF4"])
    all3 = f["beam-dloadF2-dloadF4-centrif"]
    assert close(all3, f["beam-dloadF2"] + f["beam-dloadF4"] + f["beam-centrif"])
    for other in ("beam-dloadF4-dloadF2-centrif", "beam-dloadF2-centrif-dloadF4", "beam-centrif-dloadF2-dloadF4",
                  "beam-centrif-dloadF4-dloadF2", "beam-dloadF4-centrif-dloadF2"):
        assert close(all3, f[other]), other


def test_bench_load_combination_oracle():
    check_combinations(combination_rhs(lambda args: O.oracle_sys_linmat(args)[3]))


def test_unused_distribution_builds():
    """benchs/dload/test_dload_beam_unused_distrib.py: a distribution declared before the one the
    Dload names. The load reads row `label` of nodal_dist, so the rows run up to that label
    (the reference's own stacking of the named distributions only would leave it out of range)."""
    m = IGAparametrization(filename=os.path.join(MODELS, "beam-dloadF2F4"))
    m.refine(np.array([2, 2, 5]), np.array([1, 1, 1]))
    d = m.get_model_dict()
    label = int(d["load_additionalInfos"][0])
    assert d["nodal_dist"].shape[0] >= label
    F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())[3]
    assert np.isfinite(F).all() and np.abs(F).max() > 0


CENTRIF2 = os.path.join(MODELS, "centrif_two_patches")
CENTRIF_REF = np.array([1.60785E-6, -1.94110E-8, 3.15129E-7])   # benchs/centrif/centrif_U1_U1_*/test_*.py


def centrif_two_patches(name):
    """benchs/centrif/centrif_U1_U1_{1load,2loads}: two single-element trilinear patches stacked
    in z, joined by the face of CPs 5-8, centrifugal body load; each patch degree +1 and two
    levels of mid-knot insertion per direction."""
    m = IGAparametrization(filename=os.path.join(CENTRIF2, name))
    nb_ref = np.full((3, m.nb_patch), 2, np.intp)
    nb_deg = np.full((3, m.nb_patch), 1, np.intp)
    m.refine(nb_ref, nb_deg)
    return m


def centrif_point_sol(m, sol):
    idx = np.nonzero(np.all(m.coords == np.array([[12.0], [1.5], [0.0]]), axis=0))[0]
    assert idx.size == 1
    return sol[idx[0]]


def test_refine_two_patches():
    """Two patches joined by a face: 6^3 refined CPs per patch, the 36 of the shared face counted
    once (396 CPs, all at distinct points); each refined patch is the coarse one (geometry at
    random parametric points within 1e-12); the reference's known answer (Abaqus, x within 2 %,
    the bench's own tolerance); the load given on each patch separately (2loads) gives the same F
    as on both at once."""
    coarse = IGAparametrization(filename=os.path.join(CENTRIF2, "centrif_U1_U1_1load"))
    m = centrif_two_patches("centrif_U1_U1_1load")
    assert m.nb_patch == 2 and m.nb_elem == 128 and m.nb_cp == 396
    assert np.unique(m.coords.T, axis=0).shape[0] == 396
    d, dc = m.get_model_dict(), coarse.get_model_dict()
    xi = np.random.default_rng(4).uniform(0.0, 1.0, (100, 3))
    for ip in (1, 2):
        args = lambda g: (ip, xi, np.asarray(g["COORDS3D"], np.float64), g["IEN"], g["nb_elem_patch"], g["Nkv"],
                          g["Ukv"], g["Nijk"], g["weight"], g["Jpqr"], g["NNODE"])
        xc, xf = O.oracle_evaldisp(*args(dc)), O.oracle_evaldisp(*args(d))
        assert np.abs(xf - xc).max() <= 1e-12 * np.abs(xc).max()
    K, r, c, F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())
    u = centrif_point_sol(m, full_sol(m, solve(m, K, r, c, F)))
    assert abs((u[0] - CENTRIF_REF[0]) / CENTRIF_REF[0]) < 0.02
    F2 = O.oracle_sys_linmat(centrif_two_patches("centrif_U1_U1_2loads").get_inputs4system_elemStorage())[3]
    assert np.abs(F2 - F).max() <= 1e-14 * np.abs(F).max()


@pytest.mark.parametrize("perm", ["flip xi", "swap xi eta"])
def test_refine_two_patches_any_orientation(perm):
    """The second patch re-parametrised (xi reversed, or xi and eta exchanged: the shared face is
    then met transposed or reversed): same number of CPs, every CP at a distinct point, and the
    displacement at the reference point within 1e-9 of the unchanged model's."""
    base = centrif_two_patches("centrif_U1_U1_1load")
    m = IGAparametrization(filename=os.path.join(CENTRIF2, "centrif_U1_U1_1load"))
    loc = m._local_nodes(np.array([1, 1, 1]))
    old = m._IEN[1][0].copy()
    for l, (i, j, k) in enumerate(loc):
        src = (1 - i, j, k) if perm == "flip xi" else (j, i, k)
        m._IEN[1][0][l] = old[loc.index(src)]
    m.refine(np.full((3, 2), 2, np.intp), np.full((3, 2), 1, np.intp))
    assert m.nb_cp == base.nb_cp == 396
    assert np.unique(m.coords.T, axis=0).shape[0] == 396
    sols = []
    for mm in (base, m):
        K, r, c, F = O.oracle_sys_linmat(mm.get_inputs4system_elemStorage())
        sols.append(centrif_point_sol(mm, full_sol(mm, solve(mm, K, r, c, F))))
    assert np.abs(sols[1] - sols[0]).max() <= 1e-9 * np.abs(sols[0]).max()


def test_refine_additional_knots():
    """additional_knots go in after the degree elevation and before the mid-knot levels
    (patchRefinement, IGA_refinementFcts.py:185-269): xi = 0.3 then two levels gives the breaks
    0, .075, .15, .225, .3, .475, .65, .825, 1 (mid-knots first would give .25/.5/.75 too); the
    geometry is unchanged (1e-12) and the benchs/dload beam-dist deflexion stays within its 2 %."""
    coarse = IGAparametrization(filename=os.path.join(MODELS, "beam-dist"))
    m = IGAparametrization(filename=os.path.join(MODELS, "beam-dist"))
    m.refine(np.array([2, 2, 5]), np.array([1, 1, 1]),
             additional_knots={"patches": np.array([0]), "1": np.array([0.3]), "2": np.array([]),
                               "3": np.array([0.55, 0.8])})
    brk = np.unique(m._Ukv[0][0])
    assert brk.size == 9 and np.abs(brk - [0.0, 0.075, 0.15, 0.225, 0.3, 0.475, 0.65, 0.825, 1.0]).max() <= 1e-15
    assert {0.55, 0.8} <= set(np.unique(m._Ukv[0][2]).tolist())
    xi = np.random.default_rng(5).uniform(0.0, 1.0, (200, 3))
    xc = geometry_at(coarse.get_model_dict(), xi)
    xf = geometry_at(m.get_model_dict(), xi)
    assert np.abs(xf - xc).max() <= 1e-12 * np.abs(xc).max()
    K, r, c, F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())
    assert np.allclose(tip_deflexion(m, full_sol(m, solve(m, K, r, c, F))), DLOAD["beam-dist"], rtol=2e-2)


def _boehm(U, p, u):
    """Single knot insertion (Boehm; Piegl & Tiller, The NURBS Book, eq. 5.15): Q = A P."""
    n = U.size - p - 1
    k = int(np.searchsorted(U, u, side="right")) - 1
    A = np.zeros((n + 1, n))
    for i in range(n + 1):
        if i <= k - p:
            A[i, i] = 1.0
        elif i <= k:
            a = (u - U[i]) / (U[i + p] - U[i])
            A[i, i] = a
            A[i, i - 1] = 1.0 - a
        else:
            A[i, i - 1] = 1.0
    return A


def _homogeneous_net(m, ip=0):
    """Homogeneous control points (x w, y w, z w, w) of patch ip, xi fastest, as rows."""
    grid = m._cp_grid(ip)
    p = m._Jpqr[:, ip]
    w = np.zeros(grid.shape)
    e0 = int(np.sum(m._elementsByPatch[:ip]))
    for e in range(e0, e0 + int(m._elementsByPatch[ip])):
        s = m._Nijk[:, e]
        for l, (i, j, k) in enumerate(m._local_nodes(p)):
            w[s[0] - 1 - i, s[1] - 1 - j, s[2] - 1 - k] = m._weight[e][l]
    ids = grid.ravel(order="F") - 1
    wf = w.ravel(order="F")
    return np.column_stack([m._COORDS[:, ids].T * wf[:, None], wf])


def test_refine_transformation_matrices():
    """refine_and_getTransformationMatrices returns per patch the whole-patch matrices
    [S_xi, S_eta, S_zeta] of the reference (IGA_refinementFcts.py:340-355): S_zeta S_eta S_xi maps
    the coarse homogeneous net onto the refined one (the product _updateRefinementMatHistory
    forms, IGA_parametrization.py:2456-2461); for one extra knot on a p = 2 patch S_xi is
    I (x) I (x) Boehm's insertion matrix and the other two are identities."""
    m = IGAparametrization(filename=os.path.join(MODELS, "beam-dist"))
    Hc = _homogeneous_net(m)
    nc = [m._Ukv[0][d].size - m._Jpqr[d, 0] - 1 for d in range(3)]
    T = m.refine_and_getTransformationMatrices(np.array([1, 1, 1]), np.array([1, 1, 1]))
    assert len(T) == 1 and len(T[0]) == 3
    Hf = _homogeneous_net(m)
    nf = [m._Ukv[0][d].size - m._Jpqr[d, 0] - 1 for d in range(3)]
    S_xi, S_eta, S_zeta = T[0]
    assert S_xi.shape == (nf[0] * nc[1] * nc[2], nc[0] * nc[1] * nc[2])
    assert S_eta.shape == (nf[0] * nf[1] * nc[2], nf[0] * nc[1] * nc[2])
    assert S_zeta.shape == (nf[0] * nf[1] * nf[2], nf[0] * nf[1] * nc[2])
    for t in T[0]:
        assert np.abs(np.asarray(t.sum(axis=1)).ravel() - 1.0).max() <= 1e-14
    assert np.abs(S_zeta @ (S_eta @ (S_xi @ Hc)) - Hf).max() <= 1e-12 * np.abs(Hf).max()
    # a single knot in xi of the refined (now p = 2) patch: Boehm's matrix on every (eta, zeta)
    # line, identities elsewhere
    assert m._Jpqr[0, 0] == 2
    U0 = m._Ukv[0][0].copy()
    nc = [m._Ukv[0][d].size - m._Jpqr[d, 0] - 1 for d in range(3)]
    T = m.refine_and_getTransformationMatrices(np.array([0, 0, 0]), np.array([0, 0, 0]),
                                               additional_knots={"patches": np.array([0]), "1": np.array([0.3]),
                                                                 "2": np.array([]), "3": np.array([])})
    B = _boehm(U0, 2, 0.3)
    assert np.abs(T[0][0].toarray() - np.kron(np.eye(nc[1] * nc[2]), B)).max() <= 1e-14
    for t in T[0][1:]:
        assert np.abs(t.toarray() - np.eye(t.shape[0])).max() == 0.0


def test_refine_rejects_nonconforming_faces():
    m = IGAparametrization(filename=os.path.join(CENTRIF2, "centrif_U1_U1_1load"))
    with pytest.raises(NotImplementedError):
        m.refine(np.array([[2, 1], [2, 1], [2, 1]]), np.array([[1, 1], [1, 1], [1, 1]]))


def test_refine_rejects_knots_that_meet_reversed():
    """The same additional knot (xi = 0.3) on both patches, with the second patch's xi reversed:
    the two sides of the shared face get knots 0.3 and 0.7, the same face grid size but a
    different refinement, which refine must reject rather than merge the CPs by index. With xi
    unchanged, or knots along directions not reversed, the patches stay joined."""
    def model(flip):
        m = IGAparametrization(filename=os.path.join(CENTRIF2, "centrif_U1_U1_1load"))
        if flip:
            loc = m._local_nodes(np.array([1, 1, 1]))
            old = m._IEN[1][0].copy()
            for l, (i, j, k) in enumerate(loc):
                m._IEN[1][0][l] = old[loc.index((1 - i, j, k))]
        return m

    def knots(d):
        ak = {"patches": np.array([0, 1]), "1": np.array([]), "2": np.array([]), "3": np.array([])}
        ak[d] = np.array([0.3])
        return ak

    zero = np.zeros((3, 2), np.intp)
    for d in "123":
        model(False).refine(zero, zero, additional_knots=knots(d))
    model(True).refine(zero, zero, additional_knots=knots("2"))
    with pytest.raises(NotImplementedError):
        model(True).refine(zero, zero, additional_knots=knots("1"))


# --------------------------------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(DLOAD))
def test_gpu_bench_dload(name):
    """benchs/dload/test_dload_beam_*.py on the drop-in stack."""
    from yeti_amd import reconstructionSOL as rsol
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static as build_stiffmatrix
    print("[case] bench dload %s" % name, flush=True)
    m = dload_model(name)
    data, row, col, rhs = build_stiffmatrix(*m.get_inputs4system_elemStorage())
    x = solve(m, data, row, col, rhs)
    sol, _ = rsol.reconstruction(**m.get_inputs4solution(x))
    assert np.allclose(tip_deflexion(m, np.asarray(sol)), DLOAD[name], rtol=2e-2)
    # and the same numbers as the oracle's assembly
    K, r, c, F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())
    assert np.array_equal(row, r) and np.array_equal(col, c)
    assert np.abs(data - K).max() <= 1e-12 * np.abs(K).max()
    assert np.abs(rhs - F).max() <= 1e-12 * np.abs(F).max()


@pytest.mark.gpu
def test_gpu_bench_volume_force_3d():
    """benchs/volume_force/test_volume_force_3d.py on the drop-in stack."""
    from yeti_amd import postproc as pp
    from yeti_amd import reconstructionSOL as rsol
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static as build_stiffmatrix
    print("[case] bench volume_force_3d", flush=True)
    m = volume_force_3d()
    data, row, col, rhs = build_stiffmatrix(*m.get_inputs4system_elemStorage())
    x = solve(m, data, row, col, rhs)
    sol, _ = rsol.reconstruction(**m.get_inputs4solution(x))
    disp = pp.evaldisp(*m.get_inputs4evaldisp(np.asarray(sol).transpose(), np.array([1.0, 1.0, 0.0]), numpatch=1))
    assert np.linalg.norm(disp - VF3D_REF) / np.linalg.norm(VF3D_REF) < 2e-3


@pytest.mark.gpu
def test_gpu_bench_platevib():
    """benchs/platevib/test_platevib.py on the drop-in stack (stiffness and mass on the GPU)."""
    from yeti_amd.massmtrx import build_cmassmatrix
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static as build_stiffmatrix
    print("[case] bench platevib", flush=True)
    m = platevib()
    data, row, col, _ = build_stiffmatrix(*m.get_inputs4system_elemStorage())
    mdata, mrow, mcol = build_cmassmatrix(*m.get_inputs4massmat())
    assert np.allclose(frequencies(m, data, row, col, mdata, mrow, mcol), PLATEVIB_REF, rtol=1e-3)


@pytest.mark.gpu
def test_gpu_bench_omp():
    """benchs/omp/test_build_stifmatrix_omp.py on the drop-ins: the serial (packed) and OpenMP
    (slotted) entry points give the same matrix and vector."""
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static as build_stiffmatrix
    from yeti_amd.stiffmtrx_elemstorage_omp import sys_linmat_lindef_static_omp as build_stiffmatrix_omp
    print("[case] bench omp", flush=True)
    m = IGAparametrization(filename=os.path.join(MODELS, "beam-dist"))
    m.refine(np.array([2, 2, 2]), np.array([2, 2, 2]))
    n = m.nb_dof_tot
    data, row, col, rhs_seq = build_stiffmatrix(*m.get_inputs4system_elemStorage())
    A = sp.coo_matrix((data, (row, col)), shape=(n, n)).tocsc()
    data, row, col, rhs_omp = build_stiffmatrix_omp(**m.get_inputs4system_elemStorage_OMP(num_threads=2))
    B = sp.coo_matrix((data, (row, col)), shape=(n, n)).tocsc()
    assert spl.norm(A - B) / spl.norm(A) < 1e-15
    assert np.linalg.norm(rhs_seq - rhs_omp) / np.linalg.norm(rhs_seq) < 1e-15


@pytest.mark.gpu
def test_gpu_bench_load_combination():
    """benchs/load_combination/test_beam_dload_triangle_centrif.py on the drop-in, and each
    right-hand side against the oracle's."""
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static as build_stiffmatrix
    print("[case] bench load_combination", flush=True)
    f = combination_rhs(lambda args: build_stiffmatrix(*args)[3])
    check_combinations(f)
    fo = combination_rhs(lambda args: O.oracle_sys_linmat(args)[3])
    for k in f:
        assert np.abs(f[k] - fo[k]).max() <= 1e-12 * np.abs(fo[k]).max(), k


@pytest.mark.gpu
def test_gpu_unused_distribution():
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static as build_stiffmatrix
    print("[case] bench dload unused distribution", flush=True)
    m = IGAparametrization(filename=os.path.join(MODELS, "beam-dloadF2F4"))
    m.refine(np.array([2, 2, 5]), np.array([1, 1, 1]))
    data, row, col, rhs = build_stiffmatrix(*m.get_inputs4system_elemStorage())
    F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())[3]
    assert np.abs(rhs - F).max() <= 1e-12 * np.abs(F).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["centrif_U1_U1_1load", "centrif_U1_U1_2loads"])
def test_gpu_bench_centrif_two_patches(name):
    """benchs/centrif/centrif_U1_U1_{1load,2loads}/test_*.py on the drop-in stack, with the
    oracle's assembly of the same refined model."""
    from yeti_amd import reconstructionSOL as rsol
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static as build_stiffmatrix
    print("[case] bench centrif, two patches:", name, flush=True)
    m = centrif_two_patches(name)
    data, row, col, rhs = build_stiffmatrix(*m.get_inputs4system_elemStorage())
    K, r, c, F = O.oracle_sys_linmat(m.get_inputs4system_elemStorage())
    assert np.array_equal(row, r) and np.array_equal(col, c)
    assert np.abs(data - K).max() <= 1e-12 * np.abs(K).max()
    assert np.abs(rhs - F).max() <= 1e-12 * np.abs(F).max()
    x = solve(m, data, row, col, rhs)
    sol, _ = rsol.reconstruction(**m.get_inputs4solution(x))
    u = centrif_point_sol(m, np.asarray(sol))
    assert abs((u[0] - CENTRIF_REF[0]) / CENTRIF_REF[0]) < 0.02


@pytest.mark.gpu
def test_gpu_volume_dropin():
    """volume.computevolume / computegradvolume on get_inputs4area: the two stacked boxes of the
    centrifugal bench hold 2 x 3 x 3 = 18 and 2 x 3 x 2 = 12; listpatch selects them; the
    gradient of both patches equals Assembler.volume_grad, that of the unit cube is +-1/4."""
    from yeti_amd import volume
    from yeti_amd._abi import args_to_dict
    from yeti_amd.assembly import Assembler
    print("[case] volume drop-in", flush=True)
    m = centrif_two_patches("centrif_U1_U1_1load")
    for flags, ref in (((1, 0), 18.0), ((0, 1), 12.0), ((1, 1), 30.0)):
        V = volume.computevolume(*m.get_inputs4area(activepatch=np.array(flags)))
        assert abs(V - ref) <= 1e-12 * ref, (flags, V)
    g = volume.computegradvolume(*m.get_inputs4area())
    assert g.shape == (3, m.nb_cp) and g.flags.f_contiguous
    asm = Assembler(args_to_dict(m.get_inputs4system_elemStorage()))
    try:
        g0 = asm.volume_grad()[1].cpu().numpy().reshape(-1, 3).T
    finally:
        asm.close()
    assert np.abs(g - g0).max() <= 1e-13 * np.abs(g0).max()
    cube = IGAparametrization(filename=os.path.join(MODELS, "grad_solid", "unit_cube_U1"))
    gc = volume.computegradvolume(*cube.get_inputs4area())
    assert np.abs(np.abs(gc) - 0.25).max() <= 1e-14


@pytest.mark.gpu
def test_gpu_rejects_distribution_out_of_range():
    """A nodal-distribution load whose label has no row in nodal_dist is refused on the host
    (YETI_ERR_INVALID_ARGUMENT) instead of being read past the array on the device."""
    from yeti_amd._lib import YetiError
    from yeti_amd.stiffmtrx_elemstorage import sys_linmat_lindef_static as build_stiffmatrix
    print("[case] distribution label out of range", flush=True)
    m = dload_model("beam-triangle-1")
    d = m.get_model_dict()
    d["load_additionalInfos"] = np.array([2.0])          # nodal_dist has one row
    with pytest.raises(YetiError) as ei:
        build_stiffmatrix(**d)
    assert ei.value.code == 1
