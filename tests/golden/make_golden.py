"""Writes the golden fixtures of tests/golden/ from YETI's own Fortran, compiled from its sources
into oracle/_ref/libyetiref.so by oracle/build_ref.sh. Run in a tree where that library exists:

    python tests/golden/make_golden.py

Each fixture is data only (inputs and expected outputs):

* unit_cube_p1.npz: YETI's one-element p=1 cube of benchs/grad_solid/unit_cube_U1.{NB,inp}
  (8 control points, centrifugal load U101, E=210000, nu=0.3, rho=1), full COO + F from the
  serial routine, and gradLinElastWork_AN for a seeded state / adjoint;
* cantilever_16x4x4_p2.npz: the 10x1x1 cantilever (16x4x4 elements, p=2, seed 20261202):
  sha256 of the int32 Krow/Kcol arrays, |Kdata|_F, 2000 seeded sampled entries, per-element
  Frobenius norms, full F, the geometry's sha256 (pins the generator) and the displacement of
  a direct solve of (K + K^T)[free][:, free] u = F[free];
* box_1x1x1_p3.npz: one p=3 element (full COO + F);
* all_loads_3x2x2_p2.npz: every surface-load code plus U101/U102 body loads on a 3x2x2 p=2
  patch: COO hashes, |Kdata|_F, full F.
* grad_4x2x2_p2.npz: a 4x2x2 p=2 cantilever (seed 11, traction U23 at x=10): the state u of
  the direct solve, gradLinElastWork_AN for the adjoints (1/2 u, a seeded random field), the
  compliance J = F . u, and central finite differences (steps 1e-6 and 1e-4) of J at 24
  seeded (control point, direction) pairs, each J evaluated with the compiled serial assembly
  and a direct solve.
* grad_two_patches_p2.npz: benchs/centrif/centrif_U1_U1_1load (two patches joined by a face,
  centrifugal load) refined by yeti_amd.model (degree +1, one mid-knot level per direction per
  patch; the sha256 of its COORDS3D, weights and IEN pin that refinement): gradLinElastWork_AN
  (internal and external terms) for a seeded state and two seeded adjoints.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.helpers import all_face_loads  # noqa: E402
from yeti_amd.geometry import make_box_patch  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unit_cube_args():
    """benchs/grad_solid/unit_cube_U1.{NB,inp} as the f2py argument dict (no refinement)."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                  [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], np.float64).T
    free = np.array([d for cp in range(1, 9) for d in range(3 * cp - 2, 3 * cp + 1)
                     if cp not in (1, 3, 5, 7)], np.int32)
    idf = np.zeros(24, np.int32)
    idf[: free.size] = free
    clamp = np.array([1, 3, 5, 7], np.int32)
    return dict(
        activeElement=np.ones(1, np.int32), nb_data=300.0, COORDS3D=np.asfortranarray(X),
        IEN=np.array([8, 7, 6, 5, 4, 3, 2, 1], np.int32), nb_elem_patch=np.array([1], np.int32),
        Nkv=np.array([[4], [4], [4]], np.int32),
        Ukv=np.array([0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1], np.float64),
        Nijk=np.array([[2], [2], [2]], np.int32), weight=np.ones(8),
        Jpqr=np.array([[1], [1], [1]], np.int32), ELT_TYPE="U1", PROPS=np.array([1.0]),
        JPROPS=np.array([1], np.int32), MATERIAL_PROPERTIES=np.array([[210000.0], [0.3]]),
        n_mat_props=np.array([2], np.int32), RHO=np.array([1.0]), TENSOR="/THREED",
        indDLoad=np.array([1], np.int32), JDLType=np.array([101], np.int32),
        ADLMAG=np.array([100.0]), load_target_nbelem=np.array([1], np.int32),
        load_additionalInfos=np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0]),
        nb_load_additionalInfos=np.array([6], np.int32),
        bc_values=np.array([[1.0, 2.0, 3.0], [0.0, 0.0, 0.0]]),
        bc_target=np.concatenate([clamp, clamp, clamp]), bc_target_nbelem=np.array([4, 4, 4], np.int32),
        ind_dof_free=idf, nb_dof_free=int(free.size), MCRD=3, NBINT=np.array([8], np.int32),
        NNODE=np.array([8], np.int32), nodal_dist=np.zeros((1, 8), order="F"),
        nb_bc=3, nb_load=1, nb_patch=1, nb_elem=1, nb_cp=8, nb_dof_tot=24)


def element_fro(K, nel, nde):
    return np.sqrt((K.reshape(nel, nde) ** 2).sum(axis=1))


def solve(K, r, c, F, free0, ndof):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    A = sp.coo_matrix((K, (r, c)), shape=(ndof, ndof)).tocsc()
    A = (A + A.T).tocsc()[free0][:, free0]
    return spla.spsolve(A, F[free0])


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libyetiref.so is not built")
    rng = np.random.default_rng(7)

    a = unit_cube_args()
    K, r, c, F = O.ref_sys_linmat_serial((), kwargs=a)
    SOL = rng.uniform(-1e-3, 1e-3, size=(3, 8))
    ADJ = rng.uniform(-1e-3, 1e-3, size=(1, 3, 8))
    from yeti_amd._abi import F2PY_ARGS
    args = [a[n] for n in F2PY_ARGS] + [a["nb_bc"], a["nb_load"], a["nb_patch"], a["nb_elem"],
                                         a["nb_cp"], a["nb_dof_tot"]]
    gi, ge = O.ref_grad_linelast_work(args, SOL, ADJ)
    np.savez_compressed(os.path.join(HERE, "unit_cube_p1.npz"), Kdata=K, Krow=r, Kcol=c, F=F,
                        SOL=SOL, ADJ=ADJ, gradWint=gi, gradWext=ge)

    bp = make_box_patch(nel=(16, 4, 4), p=2)
    args = bp.f2py_args()
    K, r, c, F = O.ref_sys_linmat_serial(args)
    nde = 3321
    idx = np.sort(rng.choice(K.size, 2000, replace=False))
    u = solve(K, r, c, F, bp.ind_dof_free0, bp.nb_dof_tot)
    geo = sha(np.concatenate([bp.args["COORDS3D"].ravel(order="F"), bp.args["weight"]])) + \
        sha(bp.args["IEN"])
    np.savez_compressed(os.path.join(HERE, "cantilever_16x4x4_p2.npz"),
                        krow_sha=sha(r), kcol_sha=sha(c), kfro=np.linalg.norm(K),
                        sample_idx=idx, sample_val=K[idx], elem_fro=element_fro(K, 256, nde),
                        F=F, u=u, free0=bp.ind_dof_free0, geometry_sha=geo)

    bp = make_box_patch(nel=(1, 1, 1), p=3, seed=3)
    K, r, c, F = O.ref_sys_linmat_serial(bp.f2py_args())
    np.savez_compressed(os.path.join(HERE, "box_1x1x1_p3.npz"), Kdata=K, Krow=r, Kcol=c, F=F)

    bp = make_box_patch(nel=(3, 2, 2), p=2, seed=5)
    al = all_face_loads(bp)
    K, r, c, F = O.ref_sys_linmat_serial((), kwargs=al)
    np.savez_compressed(os.path.join(HERE, "all_loads_3x2x2_p2.npz"), krow_sha=sha(r),
                        kcol_sha=sha(c), kfro=np.linalg.norm(K), F=F)
    grad_fixture()
    print("wrote fixtures to", HERE)


def compliance_ref(bp, X):
    """J = F . u with the compiled serial assembly at control points X (3, nb_cp)."""
    a = dict(bp.args)
    a["COORDS3D"] = np.asfortranarray(X)
    args = [a[n] for n in __import__("yeti_amd._abi", fromlist=["F2PY_ARGS"]).F2PY_ARGS]
    args += [a["nb_bc"], a["nb_load"], a["nb_patch"], a["nb_elem"], a["nb_cp"], a["nb_dof_tot"]]
    K, r, c, F = O.ref_sys_linmat_serial(args)
    free = bp.ind_dof_free0
    uf = solve(K, r, c, F, free, bp.nb_dof_tot)
    u = np.zeros(bp.nb_dof_tot)
    u[free] = uf
    return float(F @ u), u, F


def grad_fixture():
    bp = make_box_patch(nel=(4, 2, 2), p=2, seed=11)
    X0 = np.array(bp.args["COORDS3D"], dtype=np.float64)
    J, u, F = compliance_ref(bp, X0)
    SOL = u.reshape(-1, 3).T.copy()
    rng = np.random.default_rng(12)
    ADJ = np.stack([0.5 * SOL, rng.uniform(-1e-3, 1e-3, size=SOL.shape)])
    gi, ge = O.ref_grad_linelast_work(bp.f2py_args(), SOL, ADJ)
    n = bp.nb_cp
    cps = rng.choice(n, 24, replace=False)
    dirs = rng.integers(0, 3, 24)
    # make sure the loaded face x = 10 and the clamped face x = 0 are both represented
    xs = X0[0]
    cps[:4] = np.flatnonzero(np.isclose(xs, 10.0))[rng.choice(int(np.isclose(xs, 10.0).sum()), 4, replace=False)]
    cps[4:6] = np.flatnonzero(np.isclose(xs, 0.0))[:2]
    # step 1e-6 as the reference's _sensitivityFD default (OPT_modelling.py:411-437); at that step
    # the ~1e-14 noise of J (direct solve) gives ~1e-8 absolute error, so 1e-4 is kept as well
    fds = {}
    for h in (1e-6, 1e-4):
        fd = np.zeros(24)
        for t in range(24):
            Xp = X0.copy(); Xp[dirs[t], cps[t]] += h
            Xm = X0.copy(); Xm[dirs[t], cps[t]] -= h
            fd[t] = (compliance_ref(bp, Xp)[0] - compliance_ref(bp, Xm)[0]) / (2 * h)
        fds[h] = fd
    geo = sha(np.concatenate([bp.args["COORDS3D"].ravel(order="F"), bp.args["weight"]])) + sha(bp.args["IEN"])
    np.savez_compressed(os.path.join(HERE, "grad_4x2x2_p2.npz"), SOL=SOL, ADJ=ADJ, gradWint=gi, gradWext=ge,
                        J=J, F=F, fd_cp=cps, fd_dir=dirs, fd_val=fds[1e-6], fd_val_h1e4=fds[1e-4],
                        geometry_sha=geo)


def two_patch_grad_fixture():
    from yeti_amd.model import IGAparametrization
    m = IGAparametrization(filename=os.path.join(HERE, "models", "centrif_two_patches", "centrif_U1_U1_1load"))
    m.refine(np.full((3, 2), 1, np.intp), np.full((3, 2), 1, np.intp))
    args = m.get_inputs4system_elemStorage()
    d = m.get_model_dict()
    rng = np.random.default_rng(21)
    SOL = rng.uniform(-1e-3, 1e-3, size=(3, m.nb_cp))
    ADJ = rng.uniform(-1e-3, 1e-3, size=(2, 3, m.nb_cp))
    gi, ge = O.ref_grad_linelast_work(args, SOL, ADJ)
    geo = sha(np.concatenate([np.asarray(d["COORDS3D"]).ravel(order="F"), np.asarray(d["weight"], np.float64)])) + \
        sha(np.asarray(d["IEN"], np.int64))
    np.savez_compressed(os.path.join(HERE, "grad_two_patches_p2.npz"), SOL=SOL, ADJ=ADJ, gradWint=gi,
                        gradWext=ge, geometry_sha=geo)


def load_combination_grad_fixture():
    """gradLinElastWork_AN on two refined beams of benchs/load_combination carrying a nodal
    pressure field on face 2 (U24, one additional value) and a centrifugal load (U101, six
    values), in both orders, with half of the elements active: the reference reads the
    centrifugal axis points from the start of load_additionalInfos whatever precedes them."""
    from yeti_amd.model import IGAparametrization
    out = {}
    rng = np.random.default_rng(31)
    for tag, name in (("pf", "beam-dloadF2-centrif"), ("fp", "beam-centrif-dloadF2")):
        m = IGAparametrization(filename=os.path.join(HERE, "models", "load_combination", name))
        m.refine(np.array([1, 1, 1]), np.array([1, 1, 1]))
        args = list(m.get_inputs4system_elemStorage())
        act = (rng.uniform(size=m.nb_elem) < 0.5).astype(np.int64)
        act[0] = 1
        args[0] = act                                     # activeElement
        SOL = rng.uniform(-1e-3, 1e-3, size=(3, m.nb_cp))
        ADJ = rng.uniform(-1e-3, 1e-3, size=(2, 3, m.nb_cp))
        gi, ge = O.ref_grad_linelast_work(args, SOL, ADJ)
        out.update({tag + "_active": act, tag + "_SOL": SOL, tag + "_ADJ": ADJ, tag + "_gradWint": gi,
                    tag + "_gradWext": ge})
    np.savez_compressed(os.path.join(HERE, "grad_load_combination_p2.npz"), **out)


def c5_fields(X):
    """The state and adjoint fields of the 32^3 p=2 gradient fixture: closed-form functions of the
    control-point coordinates X (3, nb_cp), so that the fixture needs no stored inputs.
    SOL (3, nb_cp); ADJ (2, 3, nb_cp) = (SOL / 2, a second field)."""
    ph = np.arange(3)[:, None]
    SOL = 1e-3 * np.sin(0.7 * X[0][None] + 1.3 * X[1][None] + 0.4 * X[2][None] + ph)
    A1 = 1e-3 * np.cos(0.3 * X[0][None] - 2.1 * X[1][None] + 1.7 * X[2][None] + 0.5 * ph)
    return SOL, np.stack([0.5 * SOL, A1])


def c5_grad_fixture():
    """gradLinElastWork_AN on the 32^3 p=2 cantilever (39,304 control points): gradWint for the
    closed-form fields of c5_fields at 8,192 seeded control points (all three directions, both
    adjoints) plus the Frobenius norm and the sum of each adjoint's full array. With ADJ_0 =
    SOL / 2 the first adjoint is the compliance's (gradWint_0 = -1/2 u^T dK/dX u)."""
    bp = make_box_patch(nel=(32, 32, 32), p=2)
    X = np.asarray(bp.args["COORDS3D"], np.float64)
    SOL, ADJ = c5_fields(X)
    gi, ge = O.ref_grad_linelast_work(bp.f2py_args(), SOL, ADJ)
    rng = np.random.default_rng(32)
    cps = np.sort(rng.choice(bp.nb_cp, 8192, replace=False))
    geo = sha(np.concatenate([bp.args["COORDS3D"].ravel(order="F"), bp.args["weight"]])) + sha(bp.args["IEN"])
    np.savez_compressed(os.path.join(HERE, "grad_32cube_p2.npz"), sample_cp=cps, gradWint_sample=gi[:, :, cps],
                        fro=np.sqrt((gi ** 2).sum(axis=(1, 2))), total=gi.sum(axis=(1, 2)),
                        max_abs=np.abs(gi).max(axis=(1, 2)), gradWext_max=np.abs(ge).max(), geometry_sha=geo)


if __name__ == "__main__":
    main()
    two_patch_grad_fixture()
    load_combination_grad_fixture()
    c5_grad_fixture()
