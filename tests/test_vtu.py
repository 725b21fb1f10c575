"""VTU output of solid patches (postproc_vtu.f:22-297 with compute_svars_Q1 and writePatch):
oracle pins on the CPU, the native writer on the CPU, the GPU vertex fields against the oracle
and the postproc.generatevtu drop-in end to end.

Pins of the oracle's vertex fields (oracle_svars_solid):
* YETI's generateVTU compiled from its sources (oracle/_ref) writes the file; its points and
  displacements must equal the oracle's bit for bit, stress and strain within 1e-14 of their
  largest entry, the cell connectivity, offsets and types exactly;
* a displacement linear in the physical coordinates, u = A x + b, is reproduced exactly by the
  rational basis (partition of unity + the point being sum R x), so at every vertex u = A x + b,
  the strain is (A11, A22, A33, A12 + A21, A13 + A31, A23 + A32) and the stress is D times it: a
  wrong shear pairing, a transposed inverse Jacobian or a dropped quotient-rule term fails it;
* on an undistorted box with unit weights the geometry map is affine, so the vertex points are
  L * u with u the vertex's parametric coordinate.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from yeti_amd.geometry import make_box_patch

CASES = [((1, 1, 1), (True, True, True)), ((3, 2, 1), (True, True, True)), ((2, 2, 2), (False, True, False)),
         ((2, 1, 3), (True, False, True))]


def model(nel=(2, 2, 1), p=2, seed=1, jitter=0.05, unit_weights=False):
    bp = make_box_patch(nel=nel, p=p, seed=seed, jitter=jitter)
    d = dict(bp.args)
    if unit_weights:
        d["weight"] = np.ones_like(np.asarray(d["weight"], np.float64))
    return d


def random_sol(d, seed=0, scale=1e-3):
    return np.random.default_rng(seed).standard_normal((int(d["nb_cp"]), 3)) * scale


def split(sv, flags):
    out = {"points": sv[:, :, 0:3]}
    o = 3
    for name, on, n in (("disp", flags[0], 3), ("stress", flags[1], 6), ("strain", flags[2], 6)):
        if on:
            out[name] = sv[:, :, o:o + n]
            o += n
    return out


def von_mises(s):
    return np.sqrt(0.5 * ((s[..., 0] - s[..., 1]) ** 2 + (s[..., 0] - s[..., 2]) ** 2 + (s[..., 1] - s[..., 2]) ** 2
                          + 6.0 * (s[..., 3] ** 2 + s[..., 4] ** 2 + s[..., 5] ** 2)))


def expected_cells(nelem, nb_ref):
    nb = [2 ** max(r - 1, 0) + 1 for r in nb_ref]
    nv = nb[0] * nb[1] * nb[2]
    conn = []
    for e in range(nelem):
        for k in range(nb[2] - 1):
            for j in range(nb[1] - 1):
                for i in range(nb[0] - 1):
                    c = e * nv + k * nb[0] * nb[1] + j * nb[0] + i
                    t = c + nb[0] * nb[1]
                    conn += [c, c + 1, c + nb[0] + 1, c + nb[0], t, t + 1, t + nb[0] + 1, t + nb[0]]
    return np.array(conn, np.float64)


def check_piece(pc, sv, flags, nb_ref, tol_pts=0.0, tol_f=1e-14):
    nelem, nv, _ = sv.shape
    f = split(sv, flags)
    assert pc["NumberOfPoints"] == nelem * nv
    ncell = len(expected_cells(nelem, nb_ref)) // 8
    assert pc["NumberOfCells"] == ncell
    assert np.array_equal(pc["connectivity"], expected_cells(nelem, nb_ref))
    assert np.array_equal(pc["offsets"], 8.0 * np.arange(1, ncell + 1))
    assert np.all(pc["types"] == 12.0)
    assert np.abs(pc["points"] - f["points"].reshape(-1)).max() <= tol_pts * np.abs(f["points"]).max()
    for name in ("disp", "stress", "strain"):
        if name in f:
            ref = f[name].reshape(-1)
            assert np.abs(pc[name] - ref).max() <= tol_f * np.abs(ref).max(), name
        else:
            assert name not in pc
    if "stress" in f:
        vm = von_mises(f["stress"]).reshape(-1)
        assert np.abs(pc["VM"] - vm).max() <= max(tol_f, 1e-15) * 4 * vm.max()
    else:
        assert "VM" not in pc


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_svars_matches_reference(p, case, tmp_path):
    nb_ref, flags = CASES[case]
    d = model(p=p, seed=p)
    sol = random_sol(d, seed=case)
    path = O.ref_generate_vtu(d, sol, tmp_path, "ref", nb_ref, flags)
    pcs = O.parse_vtu(path)
    assert len(pcs) == 1
    sv = O.oracle_svars_solid(d, 0, sol, nb_ref, flags)
    f = split(sv, flags)
    pc = pcs[0]
    assert np.array_equal(pc["points"], f["points"].reshape(-1))
    if flags[0]:
        assert np.array_equal(pc["disp"], f["disp"].reshape(-1))
    check_piece(pc, sv, flags, nb_ref, tol_pts=0.0, tol_f=1e-14)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_oracle_svars_linear_field(p):
    d = model(nel=(3, 2, 2), p=p, seed=5)
    X = np.asarray(d["COORDS3D"], np.float64)            # (3, ncp)
    A = np.array([[1.0, 0.3, -0.2], [0.5, -0.7, 0.1], [0.25, 0.4, 0.9]]) * 1e-3
    b = np.array([0.01, -0.02, 0.03])
    sol = (A @ X).T + b
    sv = O.oracle_svars_solid(d, 0, sol, (2, 2, 2), (True, True, True))
    f = split(sv, (True, True, True))
    x = f["points"].reshape(-1, 3)
    assert np.abs(f["disp"].reshape(-1, 3) - (x @ A.T + b)).max() <= 1e-14
    eps = np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[0, 2] + A[2, 0], A[1, 2] + A[2, 1]])
    assert np.abs(f["strain"].reshape(-1, 6) - eps).max() <= 1e-15
    E, nu = 210000.0, 0.3
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    sig = np.concatenate([2 * mu * eps[:3] + lam * eps[:3].sum(), mu * eps[3:]])
    assert np.abs(f["stress"].reshape(-1, 6) - sig).max() <= 1e-12 * np.abs(sig).max()


def test_oracle_svars_affine_points():
    p, nel, L = 2, (3, 2, 2), np.array([10.0, 1.0, 1.0])
    d = model(nel=nel, p=p, jitter=0.0, unit_weights=True)
    sv = O.oracle_svars_solid(d, 0, np.zeros((int(d["nb_cp"]), 3)), (2, 2, 2), (False, False, False))
    nijk = np.asarray(d["Nijk"]).reshape(3, -1, order="F")
    Uk = [np.concatenate([np.zeros(p), np.linspace(0, 1, n + 1), np.ones(p)]) for n in nel]
    t = np.array([-1.0, 0.0, 1.0])
    for e in range(sv.shape[0]):
        u = []
        for dd in range(3):
            s1 = nijk[dd, e]
            u0, u1 = Uk[dd][s1 - 1], Uk[dd][s1]
            u.append(((u1 - u0) * t + (u1 + u0)) * 0.5)
        zz, yy, xx = np.meshgrid(u[2], u[1], u[0], indexing="ij")
        exp = np.stack([xx.ravel() * L[0], yy.ravel() * L[1], zz.ravel() * L[2]], axis=1)
        assert np.abs(sv[e] - exp).max() <= 1e-13


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("case", range(len(CASES)))
def test_writer_matches_reference_file(case, tmp_path):
    """The native writer (host code of the library, no GPU call) on the oracle's fields against
    the file YETI writes: same arrays, integers exact, numbers read back to the written doubles."""
    import ctypes as C

    from yeti_amd import _lib
    nb_ref, flags = CASES[case]
    d = model(p=2, seed=7)
    sol = random_sol(d, seed=case)
    sv = np.ascontiguousarray(O.oracle_svars_solid(d, 0, sol, nb_ref, flags))
    path = str(tmp_path / "ours.vtu")
    ref_ = np.asarray(nb_ref, np.int32)
    fl = np.asarray([int(f) for f in flags], np.int32)
    with open(path, "w") as fh:
        fh.write(' <VTKFile type="UnstructuredGrid"  version="0.1"   >\n <UnstructuredGrid>\n')
    _lib.check(_lib.lib().yeti_vtu_write_piece(path.encode(), sv.ctypes.data_as(C.c_void_p), sv.shape[0],
                                               ref_.ctypes.data_as(C.c_void_p), fl.ctypes.data_as(C.c_void_p)))
    with open(path, "a") as fh:
        fh.write(" </UnstructuredGrid>\n </VTKFile>\n")
    ours = O.parse_vtu(path)[0]
    check_piece(ours, sv, flags, nb_ref, tol_pts=0.0, tol_f=0.0)
    theirs = O.parse_vtu(O.ref_generate_vtu(d, sol, tmp_path, "ref", nb_ref, flags))[0]
    assert sorted(ours) == sorted(theirs)
    for k in ("connectivity", "offsets", "types", "points"):
        assert np.array_equal(ours[k], theirs[k]), k
    for k in ("disp", "stress", "strain", "VM"):
        if k in ours:
            assert np.abs(ours[k] - theirs[k]).max() <= 1e-14 * np.abs(theirs[k]).max(), k


def test_writer_rejects_bad_refinement(tmp_path):
    import ctypes as C

    from yeti_amd import _lib
    sv = np.zeros((1, 8, 3))
    ref_ = np.array([0, 1, 1], np.int32)
    fl = np.zeros(3, np.int32)
    rc = _lib.lib().yeti_vtu_write_piece(str(tmp_path / "x.vtu").encode(), sv.ctypes.data_as(C.c_void_p), 1,
                                         ref_.ctypes.data_as(C.c_void_p), fl.ctypes.data_as(C.c_void_p))
    assert rc != 0


# --------------------------------------------------------------------------------------- GPU


def _svars_gpu(d, sol, nb_ref, flags):
    from yeti_amd.postproc import svars_solid
    return svars_solid(1, nb_ref, flags, np.asarray(sol).T, d["COORDS3D"], d["IEN"], d["nb_elem_patch"], d["Nkv"],
                       d["Ukv"], d["Nijk"], d["weight"], d["Jpqr"], d["NNODE"], d["MATERIAL_PROPERTIES"])


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_gpu_svars_matches_oracle(p, case):
    nb_ref, flags = CASES[case]
    print("[case] svars p=%d nb_ref=%s flags=%s" % (p, nb_ref, flags), flush=True)
    d = model(nel=(3, 2, 2), p=p, seed=20 + p)
    sol = random_sol(d, seed=case)
    sv = _svars_gpu(d, sol, nb_ref, flags)
    so = O.oracle_svars_solid(d, 0, sol, nb_ref, flags)
    assert sv.shape == so.shape
    f, fo = split(sv, flags), split(so, flags)
    for name in fo:
        tol = 1e-13 if name in ("points", "disp") else 1e-11
        assert np.abs(f[name] - fo[name]).max() <= tol * np.abs(fo[name]).max(), name


@pytest.mark.gpu
def test_gpu_svars_linear_field():
    print("[case] svars linear field", flush=True)
    d = model(nel=(3, 3, 2), p=3, seed=9)
    X = np.asarray(d["COORDS3D"], np.float64)
    A = np.array([[1.0, 0.3, -0.2], [0.5, -0.7, 0.1], [0.25, 0.4, 0.9]]) * 1e-3
    b = np.array([0.01, -0.02, 0.03])
    sol = (A @ X).T + b
    f = split(_svars_gpu(d, sol, (3, 2, 2), (True, True, True)), (True, True, True))
    x = f["points"].reshape(-1, 3)
    assert np.abs(f["disp"].reshape(-1, 3) - (x @ A.T + b)).max() <= 1e-14
    eps = np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[0, 2] + A[2, 0], A[1, 2] + A[2, 1]])
    assert np.abs(f["strain"].reshape(-1, 6) - eps).max() <= 1e-14


@pytest.mark.gpu
def test_gpu_generatevtu_dropin(tmp_path):
    """postproc.generatevtu with the argument list of get_inputs4postprocVTU
    (IGA_parametrization.py:744-750); sol passed as SOL.transpose() as its callers do."""
    from yeti_amd.postproc import generatevtu
    print("[case] generatevtu", flush=True)
    d = model(nel=(4, 2, 2), p=2, seed=3)
    sol = random_sol(d, seed=4)
    nb_ref, flags = np.array([3, 2, 2]), np.array([True, True, True])
    args = ["out", str(tmp_path), flags, nb_ref, sol.transpose(), d["COORDS3D"], d["IEN"], d["nb_elem_patch"],
            d["Nkv"], d["Ukv"], d["Nijk"], d["weight"], d["Jpqr"], d["ELT_TYPE"], d["MATERIAL_PROPERTIES"][:2],
            d["TENSOR"], d["PROPS"], d["JPROPS"], d["NNODE"], 1, int(d["nb_elem"]), int(d["nb_cp"]), 3]
    path = generatevtu(*args)
    assert os.path.basename(path) == "out.vtu"
    pcs = O.parse_vtu(path)
    assert len(pcs) == 1
    so = O.oracle_svars_solid(d, 0, sol, tuple(nb_ref), tuple(flags))
    check_piece(pcs[0], so, tuple(flags), tuple(nb_ref), tol_pts=1e-13, tol_f=1e-11)


@pytest.mark.gpu
def test_gpu_generatevtu_two_patches(tmp_path):
    """generatevtu on the refined two-patch centrifugal bench (get_inputs4postprocVTU): one piece
    per patch, each equal to the oracle's vertex fields of that patch; the vertices of the shared
    face z = 3 carry the same displacement in both pieces."""
    from yeti_amd.model import IGAparametrization
    from yeti_amd.postproc import generatevtu
    print("[case] generatevtu two patches", flush=True)
    m = IGAparametrization(filename=os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models",
                                                 "centrif_two_patches", "centrif_U1_U1_1load"))
    m.refine(np.full((3, 2), 1, np.intp), np.full((3, 2), 1, np.intp))
    d = m.get_model_dict()
    sol = random_sol(d, seed=6)
    nb_ref, flags = np.array([2, 2, 2]), np.array([True, True, True])
    path = generatevtu(*m.get_inputs4postprocVTU("two", sol.transpose(), nb_ref=nb_ref, Flag=flags,
                                                 output_path=str(tmp_path)))
    pcs = O.parse_vtu(path)
    assert len(pcs) == 2
    for ip in range(2):
        so = O.oracle_svars_solid(d, ip, sol, tuple(nb_ref), tuple(flags))
        check_piece(pcs[ip], so, tuple(flags), tuple(nb_ref), tol_pts=1e-13, tol_f=1e-11)
    face = []
    for pc in pcs:
        x = pc["points"].reshape(-1, 3)
        u = pc["disp"].reshape(-1, 3)
        on = np.abs(x[:, 2] - 3.0) <= 1e-12
        key = np.round(x[on], 9)
        order = np.lexsort(key.T)
        face.append((key[order], u[on][order]))
    (xa, ua), (xb, ub) = face
    uniq_a, ia = np.unique(xa, axis=0, return_index=True)
    uniq_b, ib = np.unique(xb, axis=0, return_index=True)
    assert uniq_a.shape[0] > 0 and np.array_equal(uniq_a, uniq_b)
    assert np.abs(ua[ia] - ub[ib]).max() <= 1e-12 * np.abs(sol).max()
