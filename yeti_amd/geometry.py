"""Single-patch 3D NURBS box generator emitting YETI's exact input arrays.

The arrays follow the conventions YETI's refinement code produces, so the generated patch
is what ``IGAparametrization.get_inputs4system_elemStorage`` would marshal for the same
geometry:

* control points numbered with xi fastest, then eta, then zeta (1-based);
* elements numbered with xi fastest; ``Nijk`` holds the 1-based knot-span index per
  direction (IGA_refinementFcts.py:472-476);
* ``IEN`` lists an element's control points in reversed local order, local node 1 being the
  highest (i, j, k) corner (IGA_refinementFcts.py:554-597, the ``[:, ::-1]`` at :596;
  docs/source/usage/connectivity_convention.rst);
* ``weight`` repeats the control-point weights per element in that same local order;
* ``ind_dof_free`` is the ascending 1-based list of free dofs, zero padded (indDOF.f:18-110).

Geometry: open uniform knot vectors of degree p, control points at the Greville abscissae
scaled to the box (so with no jitter the patch is exactly the box, for any weights: the open
knots pin every boundary face to its plane), interior control points jittered by
U[-j, j] * h per coordinate (h = edge length / elements), weights U[w0, w1] per control point.
Face x = 0 is clamped (all three dofs), and a surface load of type ``load_code`` (default
23: face 2 = xi max, direction z) of magnitude ``traction`` acts on the elements at x = L_x.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._abi import F2PY_ARGS


def open_uniform_knots(nel: int, p: int) -> np.ndarray:
    inner = np.arange(1, nel, dtype=np.float64) / nel
    return np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)])


def greville(U: np.ndarray, p: int) -> np.ndarray:
    ncp = U.size - p - 1
    return np.array([U[i + 1:i + p + 1].mean() for i in range(ncp)])


@dataclass
class BoxPatch:
    """A generated patch: the 38 f2py arguments plus the bookkeeping a solver needs."""
    nel: tuple
    p: int
    L: tuple
    ngauss: int
    args: dict = field(default_factory=dict)
    cp_weights: np.ndarray = None
    n_cp_dir: tuple = None

    # --- views used by callers -------------------------------------------------------
    @property
    def nb_cp(self) -> int:
        return int(self.args["COORDS3D"].shape[1])

    @property
    def nb_elem(self) -> int:
        return int(np.prod(self.nel))

    @property
    def nnode(self) -> int:
        return (self.p + 1) ** 3

    @property
    def nb_dof_tot(self) -> int:
        return 3 * self.nb_cp

    @property
    def ind_dof_free0(self) -> np.ndarray:
        """0-based free dofs (ascending)."""
        idf = self.args["ind_dof_free"]
        return idf[: self.args["nb_dof_free"]].astype(np.int64) - 1

    @property
    def nb_data(self) -> int:
        return int(self.args["nb_data"])

    def f2py_args(self) -> list:
        """Positional argument list of sys_linmat_lindef_static, in the order
        IGA_parametrization.py:1705-1716 builds it (the trailing optional sizes included)."""
        a = self.args
        lst = [a[n] for n in F2PY_ARGS]
        lst += [a["nb_bc"], a["nb_load"], a["nb_patch"], a["nb_elem"], a["nb_cp"], a["nb_dof_tot"]]
        return lst

    def omp_kwargs(self, num_threads: int = -1) -> dict:
        """Keyword arguments of sys_linmat_lindef_static_omp (build_system_elemStorage_OMP.f90:26-38)."""
        a = self.args
        kw = {k.lower(): v for k, v in a.items() if k in F2PY_ARGS or k in (
            "nb_bc", "nb_load", "nb_patch", "nb_elem", "nb_cp", "nb_dof_tot")}
        kw["num_threads"] = num_threads
        return kw


def make_box_patch(nel=(16, 4, 4), p=2, L=(10.0, 1.0, 1.0), jitter=0.05,
                   weight_range=(0.8, 1.2), seed=20261202, E=210000.0, nu=0.3, rho=0.0,
                   traction=-1.0, load_code=23, ngauss=None, clamp_face_x0=True) -> BoxPatch:
    nel = tuple(int(n) for n in nel)
    ng = p + 1 if ngauss is None else int(ngauss)
    U = [open_uniform_knots(nel[d], p) for d in range(3)]
    ncpd = tuple(nel[d] + p for d in range(3))
    g = [greville(U[d], p) * L[d] for d in range(3)]
    n1, n2, n3 = ncpd
    ncp = n1 * n2 * n3
    # control points, xi fastest
    i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    i1 = i1.ravel(); i2 = i2.ravel(); i3 = i3.ravel()
    X = np.stack([g[0][i1], g[1][i2], g[2][i3]], axis=0)       # (3, ncp)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-1.0, 1.0, size=(ncp, 3))
    w_cp = rng.uniform(weight_range[0], weight_range[1], size=ncp)
    interior = ((i1 > 0) & (i1 < n1 - 1) & (i2 > 0) & (i2 < n2 - 1) & (i3 > 0) & (i3 < n3 - 1))
    h = np.array([L[d] / nel[d] for d in range(3)])
    X[:, interior] += (jit[interior] * jitter * h).T
    X = np.asfortranarray(X)

    # elements, xi fastest; local node l = i + (p+1)(j + (p+1)k) <-> cp (e1+p-i, e2+p-j, e3+p-k)
    ne = nel[0] * nel[1] * nel[2]
    e3, e2, e1 = np.meshgrid(np.arange(nel[2]), np.arange(nel[1]), np.arange(nel[0]), indexing="ij")
    e1 = e1.ravel(); e2 = e2.ravel(); e3 = e3.ravel()
    k3, k2, k1 = np.meshgrid(np.arange(p + 1), np.arange(p + 1), np.arange(p + 1), indexing="ij")
    k1 = k1.ravel(); k2 = k2.ravel(); k3 = k3.ravel()               # local i, j, k per node
    c1 = e1[:, None] + p - k1[None, :]
    c2 = e2[:, None] + p - k2[None, :]
    c3 = e3[:, None] + p - k3[None, :]
    ien0 = c1 + n1 * (c2 + n2 * c3)                                   # (ne, nnode) 0-based
    IEN = (ien0 + 1).astype(np.int32).ravel()
    weight = w_cp[ien0].ravel()
    Nijk = np.asfortranarray(np.stack([e1 + p + 1, e2 + p + 1, e3 + p + 1]).astype(np.int32))
    nnode = (p + 1) ** 3
    ndofel = 3 * nnode
    nb_data = float(ne * ndofel * (ndofel + 1) / 2)   # float, as the marshaller computes it

    # boundary conditions: clamp face x = 0
    if clamp_face_x0:
        face0 = np.where(i1 == 0)[0] + 1
        bc_target = np.concatenate([face0, face0, face0]).astype(np.int32)
        bc_target_nbelem = np.array([face0.size] * 3, dtype=np.int32)
        bc_values = np.asfortranarray(np.array([[1.0, 2.0, 3.0], [0.0, 0.0, 0.0]]))
        nb_bc = 3
    else:
        bc_target = np.zeros(1, np.int32)
        bc_target_nbelem = np.zeros(1, np.int32)
        bc_values = np.zeros((2, 1), order="F")
        nb_bc = 1
    nb_dof_tot = 3 * ncp
    fixed = np.zeros(nb_dof_tot, bool)
    if clamp_face_x0:
        for d in range(3):
            fixed[(face0 - 1) * 3 + d] = True
    free = np.where(~fixed)[0] + 1
    ind_dof_free = np.zeros(nb_dof_tot, np.int32)
    ind_dof_free[: free.size] = free

    # load on the elements at x = L (face 2 of those elements)
    if load_code is not None:
        loaded = np.where(e1 == nel[0] - 1)[0] + 1
        indDLoad = loaded.astype(np.int32)
        JDLType = np.array([load_code], np.int32)
        ADLMAG = np.array([traction], np.float64)
        load_target_nbelem = np.array([loaded.size], np.int32)
        nb_load = 1
    else:
        indDLoad = np.zeros(1, np.int32)
        JDLType = np.zeros(1, np.int32)
        ADLMAG = np.zeros(1, np.float64)
        load_target_nbelem = np.zeros(1, np.int32)
        nb_load = 1

    args = dict(
        activeElement=np.ones(ne, np.int32),
        nb_data=nb_data,
        COORDS3D=X,
        IEN=IEN,
        nb_elem_patch=np.array([ne], np.int32),
        Nkv=np.asfortranarray(np.array([[U[0].size], [U[1].size], [U[2].size]], np.int32)),
        Ukv=np.concatenate(U),
        Nijk=Nijk,
        weight=weight,
        Jpqr=np.asfortranarray(np.array([[p], [p], [p]], np.int32)),
        ELT_TYPE="U1",
        PROPS=np.zeros(1, np.float64),
        JPROPS=np.array([1], np.int32),
        MATERIAL_PROPERTIES=np.asfortranarray(np.array([[E], [nu]], np.float64)),
        n_mat_props=np.array([2], np.int32),
        RHO=np.array([rho], np.float64),
        TENSOR="/THREED",
        indDLoad=indDLoad,
        JDLType=JDLType,
        ADLMAG=ADLMAG,
        load_target_nbelem=load_target_nbelem,
        load_additionalInfos=np.array([], np.float64),
        nb_load_additionalInfos=np.zeros(1, np.int32),
        bc_values=bc_values,
        bc_target=bc_target,
        bc_target_nbelem=bc_target_nbelem,
        ind_dof_free=ind_dof_free,
        nb_dof_free=int(free.size),
        MCRD=3,
        NBINT=np.array([ng ** 3], np.int32),
        NNODE=np.array([nnode], np.int32),
        nodal_dist=np.zeros((1, ncp), order="F"),
        nb_bc=nb_bc,
        nb_load=nb_load,
        nb_patch=1,
        nb_elem=ne,
        nb_cp=ncp,
        nb_dof_tot=nb_dof_tot,
    )
    return BoxPatch(nel=nel, p=p, L=tuple(L), ngauss=ng, args=args, cp_weights=w_cp, n_cp_dir=ncpd)


# The workloads the path is measured and checked on.
def cantilever_16x4x4_p2(seed=20261202):
    """Reference CPU-runnable case: 10x1x1 beam, 16x4x4 elements, p=2 (1944 dofs)."""
    return make_box_patch(nel=(16, 4, 4), p=2, seed=seed)


def cube_64_p3(seed=20261202):
    """64x64x64 elements, p=3 (902,289 dofs, 262,144 elements)."""
    return make_box_patch(nel=(64, 64, 64), p=3, seed=seed)


def cube_128_p2(seed=20261202):
    """128x128x128 elements, p=2 (6,591,000 dofs)."""
    return make_box_patch(nel=(128, 128, 128), p=2, seed=seed)


def cube_32_p2(seed=20261202):
    """32x32x32 elements, p=2 (117,912 dofs)."""
    return make_box_patch(nel=(32, 32, 32), p=2, seed=seed)
