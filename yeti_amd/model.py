"""Model files and refinement for solid (U1) models: the part of YETI's pre-processing that turns
a ``.inp`` + ``.NB`` pair into the flattened f2py argument lists
(``IGAparametrization(filename=...)``, IGA_parametrization.py; readers
preprocessing/geometricmodel/nurbsmodel.py and preprocessing/mechanicalmodel/read/abaqus.py) and
the k-refinement ``IGAparametrization.refine`` (IGA_parametrization.py:2294-2437,
IGA_refinementFcts.py).

Scope: solid patches; the keywords the solid models use (``*USER ELEMENT``, ``*Node``,
``*Element``, ``*ELSET`` / ``*NSET`` (lists or ``generate``), ``*UEL PROPERTY``, ``*MATERIAL``
with ``*Elastic`` / ``*Density``, ``*Boundary`` (first/last dof, value), ``*Dload`` (``U<code>``,
magnitude, additional values, or the name of a nodal ``*DISTRIBUTION``), ``*Cload`` (node or
node set, dof, magnitude); one instance. Refinement is implemented for
single-patch models: degree elevation, then levels of mid-knot insertion on every non-empty
interval, both exact. The refined control net is obtained by interpolating the coarse NURBS (in
homogeneous coordinates) at the Greville abscissae of the refined space, which contains it.
Node sets that are a whole patch, faces, edges or vertices of the coarse patch are mapped to the
same entities of the refined one (updateBoundaryConditions, IGA_refinementFcts.py:610-690);
loaded elements to the refined elements inside them, and for a face load only to those on that
face; concentrated loads keep the patch vertices among their CPs (updateLoads,
IGA_refinementFcts.py:700-878).
"""
from __future__ import annotations

import re

import numpy as np
from scipy import sparse

from ._abi import F2PY_ARGS, F2PY_OPTIONAL

# --------------------------------------------------------------------------------------- NB


def _clean_lines(path):
    """Lines without comments, a definition continued on the next line when it ends with ','
    (nurbsmodel.py:38-61)."""
    out, cur = [], ""
    with open(path) as fh:
        for raw in fh:
            line = raw.rstrip()
            if line.startswith("**"):
                continue
            cur += line
            if line.rstrip().endswith(","):
                continue
            if cur.strip():
                out.append(cur.strip())
            cur = ""
    if cur.strip():
        out.append(cur.strip())
    return out


def _nums(line, dtype=float):
    return np.array([dtype(float(t)) for t in re.split(r"[,\s]+", line.strip()) if t], dtype=dtype)


def read_nb(path):
    """Geometry of a ``.NB`` file (nurbsmodel.py:24-165)."""
    lines = _clean_lines(path)
    low = [l.lower() for l in lines]

    def at(key):
        for i, l in enumerate(low):
            if l.startswith(key):
                return i
        raise ValueError("%s: keyword %r missing" % (path, key))

    nb_patch = int(_nums(lines[at("*number of patch") + 1], int)[0])
    dim = _nums(lines[at("*dimension") + 1], int)
    dim = np.resize(dim, nb_patch) if dim.size == 1 else dim
    nnode = _nums(lines[at("*number of cp by element") + 1], int)
    nnode = np.resize(nnode, nb_patch) if nnode.size == 1 else nnode
    nb_elem = int(_nums(lines[at("*total number of element") + 1], int)[0])
    nb_elem_patch = _nums(lines[at("*number of element by patch") + 1], int)
    Ukv, Nkv = [], np.zeros((3, nb_patch), np.int64)
    for ip in range(nb_patch):
        i = at("*patch(%d)" % (ip + 1)) + 1
        kv = []
        for d in range(dim[ip]):
            Nkv[d, ip] = int(float(lines[i]))
            kv.append(_nums(lines[i + 1]))
            i += 2
        Ukv.append(kv)
    Jpqr = np.zeros((3, nb_patch), np.int64)
    i = at("*jpqr")
    Jpqr[: dim[0], 0] = _nums(lines[i + 1], int)
    if nb_patch > 1:
        if i + 2 >= len(lines) or lines[i + 2].startswith("*"):
            Jpqr[:, 1:] = Jpqr[:, :1]
        else:
            for j in range(1, nb_patch):
                Jpqr[: dim[j], j] = _nums(lines[i + 1 + j], int)
    Nijk = np.zeros((3, nb_elem), np.int64)
    i = at("*nijk")
    for e in range(nb_elem):
        v = _nums(lines[i + 1 + e], int)
        Nijk[: v.size - 1, e] = v[1:]
    weight = []
    i = at("*weight")
    k = 0
    for ip in range(nb_patch):
        for e in range(nb_elem_patch[ip]):
            weight.append(_nums(lines[i + 1 + k])[1:])
            k += 1
    return dict(dim=dim, nnode=nnode, nb_patch=nb_patch, nb_elem=nb_elem, nb_elem_patch=nb_elem_patch, Ukv=Ukv,
                Nkv=Nkv, Jpqr=Jpqr, Nijk=Nijk, weight=weight)


# -------------------------------------------------------------------------------------- INP


def _kw_opts(line):
    opts = {}
    for part in line.split(",")[1:]:
        if "=" in part:
            k, v = part.split("=", 1)
            opts[k.strip().upper()] = v.strip()
        elif part.strip():
            opts[part.strip().upper()] = True
    return line.split(",")[0].strip().upper(), opts


def read_inp(path):
    """Mechanical settings of a ``.inp`` file (abaqus.py: user elements, nodes, elements, sets,
    UEL properties, materials, boundary conditions, distributed and concentrated loads)."""
    with open(path) as fh:
        raw = [l.rstrip("\n") for l in fh]
    blocks = []   # (keyword, opts, data lines)
    for line in raw:
        s = line.strip()
        if not s or s.startswith("**"):
            continue
        if s.startswith("*"):
            kw, opts = _kw_opts(s)
            blocks.append([kw, opts, []])
        elif blocks:
            blocks[-1][2].append(s)
    m = dict(uel={}, nodes={}, elements={}, elsets={}, nsets={}, props=[], materials={}, bcs=[], dloads=[],
             cloads=[], dists=[], load_order=[])
    mat = None
    for kw, opts, data in blocks:
        if kw == "*USER ELEMENT":
            m["uel"][opts.get("TYPE", "U1").upper()] = opts
        elif kw == "*NODE":
            for d in data:
                v = _nums(d)
                m["nodes"][int(v[0])] = v[1:]
            if "NSET" in opts:
                m["nsets"][opts["NSET"]] = sorted(int(_nums(d)[0]) for d in data)
        elif kw == "*ELEMENT":
            ids = []
            for d in data:
                v = _nums(d, int)
                m["elements"][int(v[0])] = (opts.get("TYPE", "U1").upper(), v[1:])
                ids.append(int(v[0]))
            if "ELSET" in opts:
                m["elsets"][opts["ELSET"]] = ids
        elif kw in ("*ELSET", "*NSET"):
            name = opts.get("ELSET" if kw == "*ELSET" else "NSET")
            vals = []
            for d in data:
                vals.extend(int(x) for x in _nums(d, int))
            if opts.get("GENERATE"):
                a, b, st = (vals + [1])[:3]
                vals = list(range(a, b + 1, st))
            (m["elsets"] if kw == "*ELSET" else m["nsets"])[name] = vals
        elif kw == "*UEL PROPERTY":
            vals = []
            for d in data:
                vals.extend(_nums(d).tolist())
            m["props"].append((opts["ELSET"], opts.get("MATERIAL"), np.array(vals, np.float64)))
        elif kw == "*MATERIAL":
            mat = opts["NAME"]
            m["materials"][mat] = {}
        elif kw == "*ELASTIC":
            v = _nums(data[0])
            m["materials"][mat]["E"], m["materials"][mat]["nu"] = float(v[0]), float(v[1])
        elif kw == "*DENSITY":
            m["materials"][mat]["rho"] = float(_nums(data[0])[0])
        elif kw == "*BOUNDARY":
            for d in data:
                parts = [x.strip() for x in d.split(",")]
                first, last = int(parts[1]), int(parts[2])
                value = float(parts[3]) if len(parts) > 3 and parts[3] else 0.0
                m["bcs"].append((parts[0], first, last, value))
        elif kw == "*DISTRIBUTION":
            # (name, default values, {cp: values}); labels are 1-based in file order
            if str(opts.get("LOCATION", "")).upper() != "NODE":
                raise NotImplementedError("*DISTRIBUTION: only LOCATION=NODE is supported")
            default, table = None, {}
            for d in data:
                parts = [x.strip() for x in d.split(",")]
                if parts[0] == "":
                    default = np.array([float(x) for x in parts[1:]])
                else:
                    table[int(parts[0])] = np.array([float(x) for x in parts[1:]])
            m["dists"].append((opts["NAME"], default, table))
        elif kw == "*DLOAD":
            for d in data:
                parts = [x.strip() for x in d.split(",")]
                code = int(parts[1].upper().split("U")[1])
                try:
                    m["dloads"].append((parts[0], code, float(parts[2]),
                                        np.array([float(x) for x in parts[3:] if x])))
                except ValueError:   # magnitude given by a nodal distribution: its label
                    names = [n for n, _, _ in m["dists"]]
                    m["dloads"].append((parts[0], code, 0.0, np.array([float(names.index(parts[2]) + 1)])))
                m["load_order"].append(("d", len(m["dloads"]) - 1))
        elif kw == "*CLOAD":
            for d in data:
                parts = [x.strip() for x in d.split(",")]
                m["cloads"].append((parts[0], int(parts[1]), float(parts[2])))
                m["load_order"].append(("c", len(m["cloads"]) - 1))
    return m


def _set_name(target):
    """'I1.CPonX0' -> 'CPonX0' (one instance)."""
    return target.split(".", 1)[1] if "." in target else target


# ---------------------------------------------------------------------------------- B-splines


def _basis_row(U, p, u):
    """All n basis functions of degree p on knot vector U at u (Cox-de Boor; the last basis
    function is 1 at the end of the last non-empty span)."""
    U = np.asarray(U, np.float64)
    n = U.size - p - 1
    N = np.zeros(U.size - 1)
    last = np.max(np.nonzero(U[1:] > U[:-1])[0])
    for i in range(U.size - 1):
        if (U[i] <= u < U[i + 1]) or (i == last and u == U[i + 1]):
            N[i] = 1.0
    for k in range(1, p + 1):
        Nk = np.zeros(U.size - 1 - k)
        for i in range(U.size - 1 - k):
            a = (u - U[i]) / (U[i + k] - U[i]) * N[i] if U[i + k] > U[i] else 0.0
            b = (U[i + k + 1] - u) / (U[i + k + 1] - U[i + 1]) * N[i + 1] if U[i + k + 1] > U[i + 1] else 0.0
            Nk[i] = a + b
        N = Nk
    return N[:n]


def _elevated_knots(U, p, t):
    """Knot vector after degree elevation by t: every distinct knot's multiplicity grows by t."""
    vals, counts = np.unique(np.asarray(U, np.float64), return_counts=True)
    return np.repeat(vals, counts + t)


def _midpoint_knots(U, levels):
    """`levels` times: insert the middle of every non-empty interval (averageKnot_insertions)."""
    U = np.asarray(U, np.float64)
    for _ in range(levels):
        mids = [(a + b) * 0.5 for a, b in zip(U[:-1], U[1:]) if b > a]
        U = np.sort(np.concatenate([U, mids]))
    return U


def _face_slices(shape):
    """Index tuples of the six faces of a 3D grid: xi-, xi+, eta-, eta+, zeta-, zeta+."""
    out = []
    for d in range(3):
        for side in (0, 1):
            sl = [slice(None)] * 3
            sl[d] = 0 if side == 0 else shape[d] - 1
            out.append(tuple(sl))
    return out


def _face_orientation(A, B):
    """The map t (transpose and/or reversal of the two axes) with t(B) == A for two 2D arrays of
    CP ids, or None when B is no reorientation of A."""
    for tr in (False, True):
        B1 = B.T if tr else B
        for f0 in (1, -1):
            for f1 in (1, -1):
                B2 = B1[::f0, ::f1]
                if B2.shape == A.shape and np.array_equal(B2, A):
                    return lambda Y, tr=tr, f0=f0, f1=f1: (Y.T if tr else Y)[::f0, ::f1]
    return None


def _patch_refinement_matrices(T):
    """[S_xi, S_eta, S_zeta] of a patch from its 1D refinement matrices T[d] (n_fine x n_coarse):
    S_xi = I_c2 (x) I_c1 (x) T0, S_eta = I_c2 (x) T1 (x) I_f0, S_zeta = T2 (x) I_f1 (x) I_f0 on
    control-point vectors ordered xi fastest."""
    (f0, c0), (f1, c1), (f2, c2) = (t.shape for t in T)
    Ts = [sparse.csc_matrix(t) for t in T]
    eye = lambda n: sparse.identity(n, format="csc")
    S_xi = sparse.kron(eye(c2), sparse.kron(eye(c1), Ts[0]), format="csc")
    S_eta = sparse.kron(eye(c2), sparse.kron(Ts[1], eye(f0)), format="csc")
    S_zeta = sparse.kron(Ts[2], sparse.kron(eye(f1), eye(f0)), format="csc")
    return [S_xi, S_eta, S_zeta]


def _same_face_net(q, fa, o, fb, t):
    """Whether two refined patches give the same control points on a shared face: q's face fa
    and o's face fb seen through the orientation t. Refining the two sides with different knots
    (say 0.3 on one side and 0.7 on the other) moves the interior face points apart."""
    A = q["Hf"][(slice(None),) + fa]
    B = o["Hf"][(slice(None),) + fb]
    B = np.stack([t(B[c]) for c in range(B.shape[0])])
    if A.shape != B.shape:
        return False
    xa, xb = A[:3] / A[3], B[:3] / B[3]
    scale = max(1.0, float(np.abs(xa).max()))
    return bool(np.abs(xa - xb).max() <= 1e-10 * scale)


def refinement_matrix(Uc, pc, Uf, pf):
    """T (n_fine x n_coarse) with N_coarse_j = sum_i T_ij N_fine_i: the coarse space lies in the
    fine one, so interpolation at the fine Greville abscissae is exact."""
    nf = Uf.size - pf - 1
    g = np.array([Uf[i + 1: i + pf + 1].mean() if pf > 0 else 0.5 * (Uf[i] + Uf[i + 1]) for i in range(nf)])
    Bf = np.array([_basis_row(Uf, pf, u) for u in g])
    Bc = np.array([_basis_row(Uc, pc, u) for u in g])
    T = np.linalg.solve(Bf, Bc)
    T[np.abs(T) < 1e-14 * np.abs(T).max()] = 0.0
    return T


# ------------------------------------------------------------------------------------ model


class IGAparametrization:
    """The flattened model of a ``.inp`` + ``.NB`` pair, with the f2py argument-list builders of
    the reference's IGAparametrization for the solid path."""

    def __init__(self, filename=None, inp=None, nb=None):
        if filename is not None:
            inp, nb = filename + ".inp", filename + ".NB"
        self._geo = read_nb(nb)
        self._mech = read_inp(inp)
        self._build_from_files()

    # -- construction ------------------------------------------------------------------------
    def _build_from_files(self):
        g, m = self._geo, self._mech
        self._nb_patch = int(g["nb_patch"])
        self._dim = np.asarray(g["dim"])
        ncp = max(m["nodes"])
        self._COORDS = np.zeros((3, ncp), order="F")
        for k, v in m["nodes"].items():
            self._COORDS[: v.size, k - 1] = v
        self._mcrd = int(next(iter(m["uel"].values())).get("COORDINATES", 3)) if m["uel"] else 3
        # patches: the element sets of the UEL PROPERTY lines, in order
        self._patch_elems = [list(m["elsets"][es]) for es, _, _ in m["props"]]
        self._elt_type = [m["elements"][els[0]][0] for els in self._patch_elems]
        self._IEN = [np.array([m["elements"][e][1] for e in els], np.int64) for els in self._patch_elems]
        self._Nkv = np.asarray(g["Nkv"], np.int64)
        self._Ukv = [[np.asarray(u, np.float64) for u in kv] for kv in g["Ukv"]]
        self._Jpqr = np.asarray(g["Jpqr"], np.int64)
        self._Nijk = np.asarray(g["Nijk"], np.int64)
        self._weight = [np.asarray(w, np.float64) for w in g["weight"]]
        self._nnode = np.asarray(g["nnode"], np.int64)
        self._elementsByPatch = np.asarray(g["nb_elem_patch"], np.int64)
        self._props = [p for _, _, p in m["props"]]
        mats = [m["materials"][mn] for _, mn, _ in m["props"]]
        self._MATERIAL_PROPERTIES = np.array([[mt["E"] for mt in mats], [mt["nu"] for mt in mats],
                                              [mt.get("rho", 0.0) for mt in mats]])
        uel = [m["uel"].get(t, {}) for t in self._elt_type]
        self._TENSOR = [str(u.get("TENSOR", "THREED")) for u in uel]
        self._NBPINT = np.array([int(u.get("INTEGRATION", 8)) for u in uel], np.int64)
        # boundary conditions: one per (target set, dof), dof <= 3 (displacements)
        self._bcs = []
        for tgt, first, last, value in m["bcs"]:
            ids = self._node_set(tgt)
            for dof in range(first, last + 1):
                if dof <= self._mcrd:
                    self._bcs.append((np.asarray(ids, np.int64), dof, value))
        self._dloads = [(np.asarray(self._elem_set(t), np.int64), code, mag, extra) for t, code, mag, extra in
                        m["dloads"]]
        self._cloads = [(np.asarray(self._node_set(t), np.int64), dof, mag) for t, dof, mag in m["cloads"]]
        self._load_order = list(m["load_order"])
        # nodal distributions (IGA_parametrization.py:133-155): (nb_cp, n) fields, unlisted CPs at
        # the default values; those a Dload names form nodal_dist (abaqus.py:1209-1214)
        self._dists = []
        used = set(int(x[0]) for _, code, _, x in m["dloads"] if code % 10 == 4 and x.size)
        for lab, (name, default, table) in enumerate(m["dists"], start=1):
            ncol = len(next(iter(table.values()))) if table else (default.size if default is not None else 1)
            f = np.zeros((ncp, ncol))
            if default is not None:
                f[:] = default
            for cp, v in table.items():
                f[cp - 1] = v
            self._dists.append((name, f, lab in used))
        self._update_dofs()

    def _node_set(self, target):
        try:
            return [int(target)]
        except ValueError:
            return list(self._mech["nsets"][_set_name(target)])

    def _elem_set(self, target):
        try:
            return [int(target)]
        except ValueError:
            return list(self._mech["elsets"][_set_name(target)])

    def _update_dofs(self):
        ncp = self._COORDS.shape[1]
        self._nb_cp = ncp
        self._nb_dof_tot = self._mcrd * ncp
        fixed = np.zeros(self._nb_dof_tot, bool)
        for ids, dof, _ in self._bcs:
            fixed[(ids - 1) * self._mcrd + dof - 1] = True
        free = np.nonzero(~fixed)[0] + 1
        self._nb_dof_free = int(free.size)
        self._ind_dof_free = np.zeros(self._nb_dof_tot, np.int64)
        self._ind_dof_free[: free.size] = free

    # -- properties of the reference's object ------------------------------------------------
    nb_patch = property(lambda self: self._nb_patch)
    nb_elem = property(lambda self: int(self._elementsByPatch.sum()))
    nb_cp = property(lambda self: self._nb_cp)
    nb_dof_tot = property(lambda self: self._nb_dof_tot)
    nb_dof_free = property(lambda self: self._nb_dof_free)
    ind_dof_free = property(lambda self: self._ind_dof_free)
    coords = property(lambda self: self._COORDS)
    mcrd = property(lambda self: self._mcrd)

    # -- refinement ----------------------------------------------------------------------------
    def refine(self, nb_refinementByDirection, nb_degreeElevationByDirection=np.array([0, 0, 0]),
               additional_knots=None):
        """k-refinement of 3D solid patches (IGA_parametrization.py:2294-2437, iga_refinement in
        IGA_refinementFcts.py:880-1015): per patch, degree elevation by
        nb_degreeElevationByDirection, then nb_refinementByDirection levels of mid-knot insertion;
        arrays of shape (3,) (every patch alike) or (3, k) (patch by patch; patches beyond k are
        not refined). additional_knots = {"patches": 0-based patch indices, "1": xi knots, "2":
        eta knots, "3": zeta knots} are inserted into those patches after the degree elevation
        and before the mid-knot levels (patchRefinement, IGA_refinementFcts.py:185-269).

        Patches that share a whole face of control points stay joined: the refined CPs of that
        face keep the numbers the earlier patch gave them and the other CPs are numbered patch by
        patch, xi fastest (build_coordsFine / updateIndCPbyPatch, IGA_refinementFcts.py:385-552).
        Sharing only edges or vertices, or faces refined differently on the two sides, raises
        NotImplementedError."""
        if any(t != "U1" for t in self._elt_type) or np.any(np.asarray(self._dim).reshape(-1) != 3):
            raise NotImplementedError("refine: 3D solid (U1) patches only")
        npch = self._nb_patch

        def per_patch(a):
            a = np.asarray(a, np.int64)
            if a.ndim == 1:
                a = np.concatenate([a, np.zeros(3 - a.size, np.int64)])
                return np.tile(a.reshape(3, 1), (1, npch))
            if a.shape[1] < npch:
                a = np.concatenate([a, np.zeros((3, npch - a.shape[1]), np.int64)], axis=1)
            return a[:, :npch]

        nref, ndeg = per_patch(nb_refinementByDirection), per_patch(nb_degreeElevationByDirection)
        eoff = np.concatenate([[0], np.cumsum(self._elementsByPatch)]).astype(np.int64)
        pts = []
        for ip in range(npch):
            p0 = self._Jpqr[:, ip].copy()
            Uc = self._Ukv[ip]
            nc = [Uc[d].size - p0[d] - 1 for d in range(3)]
            grid = self._cp_grid(ip)                               # (nc0, nc1, nc2) -> CP id
            w = np.zeros(nc)
            for e in range(eoff[ip], eoff[ip + 1]):
                s = self._Nijk[:, e]
                for l, (i, j, k) in enumerate(self._local_nodes(p0)):
                    w[s[0] - 1 - i, s[1] - 1 - j, s[2] - 1 - k] = self._weight[e][l]
            X = self._COORDS[:, grid - 1]
            # nodal fields ride along as further homogeneous coordinates (f w), so each stays the
            # same rational function (_updateNodalField, IGA_parametrization.py:2474-2497)
            extra = [np.moveaxis(f[grid - 1], -1, 0) * w[None] for _, f, _ in self._dists]
            H = np.concatenate([X * w[None], w[None]] + extra, axis=0)
            pf = p0 + ndeg[:, ip]
            Uf = []
            for d in range(3):
                Ue = _elevated_knots(Uc[d], p0[d], ndeg[d, ip])
                extra_k = np.asarray(additional_knots.get(str(d + 1), []), np.float64).reshape(-1) \
                    if additional_knots is not None and ip in np.asarray(additional_knots.get("patches", [])) \
                    else np.zeros(0)
                if extra_k.size:
                    if np.any(extra_k <= Ue[0]) or np.any(extra_k >= Ue[-1]):
                        raise ValueError("refine: additional knots must lie inside the knot vector")
                    Ue = np.sort(np.concatenate([Ue, extra_k]))
                Uf.append(_midpoint_knots(Ue, nref[d, ip]))
            T = [refinement_matrix(Uc[d], p0[d], Uf[d], pf[d]) for d in range(3)]
            Hf = np.einsum("ai,bj,ck,xijk->xabc", T[0], T[1], T[2], H, optimize=True)
            spans = [np.nonzero(Uf[d][pf[d]:-pf[d] - 1] < Uf[d][pf[d] + 1:Uf[d].size - pf[d]])[0] + pf[d] + 1
                     for d in range(3)]                            # 1-based span indices
            nijk = [(s0, s1, s2) for s2 in spans[2] for s1 in spans[1] for s0 in spans[0]]
            pts.append(dict(p0=p0, Uc=Uc, nc=nc, grid=grid, pf=pf, Uf=Uf, Hf=Hf, nf=Hf.shape[1:], nijk=nijk, T=T))

        # numbering of the refined CPs
        count = 0
        for ip, q in enumerate(pts):
            gridf = np.zeros(q["nf"], np.int64)
            for jp in range(ip):
                o = pts[jp]
                common = set(np.intersect1d(q["grid"], o["grid"]).tolist())
                if not common:
                    continue
                covered = set()
                for fa in _face_slices(q["grid"].shape):
                    A = q["grid"][fa]
                    for fb in _face_slices(o["grid"].shape):
                        B = o["grid"][fb]
                        t = _face_orientation(A, B)
                        if t is None:
                            continue
                        fa_f = _face_slices(q["nf"])[_face_slices(q["grid"].shape).index(fa)]
                        fb_f = _face_slices(o["nf"])[_face_slices(o["grid"].shape).index(fb)]
                        Bf = t(o["gridf"][fb_f])
                        if Bf.shape != gridf[fa_f].shape or not _same_face_net(q, fa_f, o, fb_f, t):
                            raise NotImplementedError("refine: patches %d and %d are refined differently on "
                                                      "their shared face" % (jp + 1, ip + 1))
                        gridf[fa_f] = Bf
                        covered |= set(A.ravel().tolist())
                if common - covered:
                    raise NotImplementedError("refine: patches %d and %d share control points outside a whole "
                                              "face" % (jp + 1, ip + 1))
            flat = gridf.ravel(order="F")
            new = np.flatnonzero(flat == 0)
            flat[new] = count + 1 + np.arange(new.size)
            q["gridf"] = flat.reshape(q["nf"], order="F")
            q["new"] = new
            count = max(count, int(flat.max()))
        nb_cp = count

        # coordinates from the patch that numbered the CP; weights and nodal fields patch by
        # patch, a later patch overwriting a shared CP (as the reference's vectWeight_fine)
        COORDS = np.zeros((3, nb_cp), order="F")
        vw = np.zeros(nb_cp)
        nfield = [f.shape[1] for _, f, _ in self._dists]
        hfields = [np.zeros((nb_cp, n)) for n in nfield]
        for q in pts:
            ids = q["gridf"].ravel(order="F") - 1
            Hf = q["Hf"].reshape(q["Hf"].shape[0], -1, order="F")
            wf = Hf[3]
            COORDS[:, ids[q["new"]]] = Hf[:3, q["new"]] / wf[q["new"]][None]
            vw[ids] = wf
            k0 = 4
            for hfd, n in zip(hfields, nfield):
                hfd[ids] = Hf[k0:k0 + n].T
                k0 += n
        self._dists = [(name, hfd / vw[:, None], used_) for (name, _, used_), hfd in zip(self._dists, hfields)]
        self._COORDS = COORDS

        # elements: every non-empty span of each patch, xi fastest, patch by patch
        old_nijk = self._Nijk.copy()
        Nijk, IEN, W, feoff = [], [], [], [0]
        for q in pts:
            loc = self._local_nodes(q["pf"])
            rows = []
            for (s0, s1, s2) in q["nijk"]:
                Nijk.append((s0, s1, s2))
                ids = [q["gridf"][s0 - 1 - i, s1 - 1 - j, s2 - 1 - k] for (i, j, k) in loc]
                rows.append(ids)
                W.append(vw[np.asarray(ids) - 1])
            IEN.append(np.asarray(rows, np.int64))
            feoff.append(feoff[-1] + len(rows))
        self._Nijk = np.asarray(Nijk, np.int64).T.copy()
        self._IEN = IEN
        self._weight = [np.asarray(x) for x in W]
        self._Ukv = [q["Uf"] for q in pts]
        self._Nkv = np.array([[q["Uf"][d].size for q in pts] for d in range(3)], np.int64)
        self._Jpqr = np.array([[q["pf"][d] for q in pts] for d in range(3)], np.int64)
        self._nnode = np.array([int(np.prod(q["pf"] + 1)) for q in pts], np.int64)
        self._elementsByPatch = np.array([len(q["nijk"]) for q in pts], np.int64)
        self._NBPINT = np.array([int(np.max(q["pf"]) + 1) ** 3 for q in pts], np.int64)

        # loaded elements: the refined elements inside a loaded coarse element; for a face load
        # (code 10 f + t, f = 1..6: xi-, xi+, eta-, eta+, zeta-, zeta+) only those on that face
        # of it (updateLoads, IGA_refinementFcts.py:700-878)
        def children(coarse_ids, code):
            face = 0 if code > 99 else code // 10
            out = []
            for ip, q in enumerate(pts):
                boxes = []
                for ce in coarse_ids:
                    if eoff[ip] < ce <= eoff[ip + 1]:
                        s = old_nijk[:, ce - 1]
                        boxes.append([(q["Uc"][d][s[d] - 1], q["Uc"][d][s[d]]) for d in range(3)])
                if not boxes:
                    continue
                Uf = q["Uf"]
                for fe, s in enumerate(q["nijk"]):
                    lo = [Uf[d][s[d] - 1] for d in range(3)]
                    hi = [Uf[d][s[d]] for d in range(3)]
                    for b in boxes:
                        if not all(b[d][0] <= lo[d] and hi[d] <= b[d][1] for d in range(3)):
                            continue
                        if face:
                            d, side = (face - 1) // 2, (face - 1) % 2
                            if (side == 0 and lo[d] != b[d][0]) or (side == 1 and hi[d] != b[d][1]):
                                continue
                        out.append(feoff[ip] + fe + 1)
                        break
            return np.asarray(out, np.int64)
        self._dloads = [(children(ids, code), code, mag, extra) for ids, code, mag, extra in self._dloads]

        # node sets: per patch, the whole patch, a face, edges or vertices of its coarse grid ->
        # the same on its refined grid; the union over the patches
        def map_patch(q, mask):
            grid, gridf, nc, nf = q["grid"], q["gridf"], q["nc"], q["nf"]
            if mask.all():
                return gridf.ravel(order="F").tolist()
            nce = [np.array([0, n - 1]) for n in nc]
            for d in range(3):          # faces
                for side in (0, 1):
                    sl = [slice(None)] * 3
                    sl[d] = nce[d][side]
                    if mask[tuple(sl)].all():
                        slf = [slice(None)] * 3
                        slf[d] = 0 if side == 0 else nf[d] - 1
                        return gridf[tuple(slf)].ravel(order="F").tolist()
            out = []
            for d in range(3):          # edges along direction d
                o = [x for x in range(3) if x != d]
                for s0 in (0, 1):
                    for s1 in (0, 1):
                        sl = [slice(None)] * 3
                        sl[o[0]] = nce[o[0]][s0]
                        sl[o[1]] = nce[o[1]][s1]
                        if mask[tuple(sl)].all():
                            slf = [slice(None)] * 3
                            slf[o[0]] = 0 if s0 == 0 else nf[o[0]] - 1
                            slf[o[1]] = 0 if s1 == 0 else nf[o[1]] - 1
                            out.extend(gridf[tuple(slf)].ravel().tolist())
            if out:
                return out
            for c in np.argwhere(mask):  # vertices
                if all(c[d] in (0, nc[d] - 1) for d in range(3)):
                    out.append(int(gridf[tuple(0 if c[d] == 0 else nf[d] - 1 for d in range(3))]))
            return out

        def map_nodes(ids):
            ids = list(set(int(i) for i in ids))
            out = []
            for q in pts:
                mask = np.isin(q["grid"], ids)
                if mask.any():
                    out.extend(map_patch(q, mask))
            return np.asarray(list(dict.fromkeys(out)), np.int64)

        def map_vertices(ids):   # concentrated loads keep the patch vertices among their CPs
            out = []
            for cp in ids:
                for q in pts:
                    c = np.argwhere(q["grid"] == cp)
                    if c.size and all(c[0][d] in (0, q["nc"][d] - 1) for d in range(3)):
                        out.append(int(q["gridf"][tuple(0 if c[0][d] == 0 else q["nf"][d] - 1 for d in range(3))]))
                        break
            return np.asarray(out, np.int64)
        self._bcs = [(map_nodes(ids), dof, v) for ids, dof, v in self._bcs]
        self._cloads = [(map_vertices(ids), dof, v) for ids, dof, v in self._cloads]
        self._update_dofs()
        return [_patch_refinement_matrices(q["T"]) for q in pts]

    def refine_and_getTransformationMatrices(self, nb_refinementByDirection,
                                             nb_degreeElevationByDirection=np.array([0, 0, 0]),
                                             additional_knots=None):
        """refine, returning per patch [S_xi, S_eta, S_zeta] (scipy CSC): whole-patch matrices
        that refine one direction each, so that S_zeta S_eta S_xi maps the coarse homogeneous
        control points (x w, w) of the patch, xi fastest, to the refined ones (the product the
        reference forms in parameterizationRefinement, IGA_refinementFcts.py:340-355, and applies
        in _updateRefinementMatHistory, IGA_parametrization.py:2456-2461; returned by
        IGA_parametrization.py:2294-2423)."""
        return self.refine(nb_refinementByDirection, nb_degreeElevationByDirection, additional_knots)

    @staticmethod
    def _local_nodes(p):
        return [(i, j, k) for k in range(p[2] + 1) for j in range(p[1] + 1) for i in range(p[0] + 1)]

    def _cp_grid(self, ip):
        """CP ids of patch ip on its (n0, n1, n2) grid, from the elements' spans and IEN (local
        node (i, j, k) of the element with spans s is grid point s - 1 - (i, j, k))."""
        p = self._Jpqr[:, ip]
        n = [self._Ukv[ip][d].size - p[d] - 1 for d in range(3)]
        grid = np.zeros(n, np.int64)
        loc = self._local_nodes(p)
        e0 = int(np.sum(self._elementsByPatch[:ip]))
        for el in range(int(self._elementsByPatch[ip])):
            s = self._Nijk[:, e0 + el]
            for l, (i, j, k) in enumerate(loc):
                grid[s[0] - 1 - i, s[1] - 1 - j, s[2] - 1 - k] = self._IEN[ip][el][l]
        if (grid == 0).any():
            raise ValueError("patch %d: control points not covered by its elements" % (ip + 1))
        return grid

    # -- f2py argument lists -------------------------------------------------------------------
    def _flat(self):
        ne = self.nb_elem
        bc_target = np.concatenate([ids for ids, _, _ in self._bcs]) if self._bcs else np.zeros(1, np.int64)
        bc_nb = np.array([ids.size for ids, _, _ in self._bcs], np.int64)
        bc_values = np.array([[dof for _, dof, _ in self._bcs], [v for _, _, v in self._bcs]], np.float64)
        # loads in file order; a concentrated load is (CP ids, JDLType = dof, magnitude)
        loads = [self._dloads[i] if kind == "d" else (self._cloads[i][0], self._cloads[i][1], self._cloads[i][2],
                                                       np.zeros(0)) for kind, i in self._load_order]
        indDLoad = np.concatenate([ids for ids, _, _, _ in loads]) if loads else np.zeros(1, np.int64)
        extra = [x for _, _, _, x in loads]
        nmat = np.full(self._nb_patch, 2, np.int64)
        d = dict(
            activeElement=np.ones(ne, np.int64),
            nb_data=float(sum(int(self._elementsByPatch[i]) * (3 * int(self._nnode[i])) * (3 * int(self._nnode[i]) + 1)
                              // 2 for i in range(self._nb_patch))),
            COORDS3D=self._COORDS, IEN=np.concatenate([x.ravel() for x in self._IEN]),
            nb_elem_patch=self._elementsByPatch, Nkv=self._Nkv,
            Ukv=np.concatenate([np.concatenate(kv) for kv in self._Ukv]), Nijk=self._Nijk,
            weight=np.concatenate([np.asarray(w, np.float64).ravel() for w in self._weight]), Jpqr=self._Jpqr,
            ELT_TYPE="".join(self._elt_type), PROPS=np.concatenate(self._props) if self._props else np.zeros(1),
            JPROPS=np.array([p.size for p in self._props], np.int64),
            MATERIAL_PROPERTIES=self._MATERIAL_PROPERTIES[:2], n_mat_props=nmat,
            RHO=self._MATERIAL_PROPERTIES[2], TENSOR="".join("/" + t for t in self._TENSOR),
            indDLoad=indDLoad, JDLType=np.array([c for _, c, _, _ in loads], np.int64),
            ADLMAG=np.array([mgn for _, _, mgn, _ in loads], np.float64),
            load_target_nbelem=np.array([ids.size for ids, _, _, _ in loads], np.int64),
            load_additionalInfos=np.concatenate(extra) if extra else np.zeros(0),
            nb_load_additionalInfos=np.array([x.size for x in extra], np.int64),
            bc_values=bc_values, bc_target=bc_target, bc_target_nbelem=bc_nb,
            ind_dof_free=self._ind_dof_free, nb_dof_free=self._nb_dof_free, MCRD=self._mcrd, NBINT=self._NBPINT,
            NNODE=self._nnode, nodal_dist=self._nodal_dist(),
            nb_bc=len(self._bcs), nb_load=len(loads), nb_patch=self._nb_patch, nb_elem=ne, nb_cp=self._nb_cp,
            nb_dof_tot=self._nb_dof_tot,
        )
        d["nb_n_dist"] = int(d["nodal_dist"].shape[0])
        d["nb_cp_n_dist"] = int(d["nodal_dist"].shape[1])
        return d

    def _nodal_dist(self):
        """The nodal distributions as rows of (rows, nb_cp), row k = the distribution with label k
        (1-based, file order), up to the highest label a Dload names. The reference stacks only
        the distributions a Dload names (_get_load_info, IGA_parametrization.py:592-611) while
        the loads address rows by label; both agree whenever the named ones come first, which is
        when the reference's own lookup stays in range."""
        last = max([k for k, (_, _, u) in enumerate(self._dists, start=1) if u], default=0)
        if last == 0:
            return np.zeros((1, self._nb_cp), order="F")
        return np.asfortranarray(np.vstack([f.T for _, f, _ in self._dists[:last]]))

    def get_inputs4system_elemStorage(self):
        """Positional argument list of sys_linmat_lindef_static (IGA_parametrization.py:1705-1716)."""
        d = self._flat()
        return [d[k] for k in F2PY_ARGS] + [d[k] for k in F2PY_OPTIONAL if k in d]

    def get_inputs4system_elemStorage_OMP(self, num_threads=None, COORDS=None, activeElem=None):
        """Keyword arguments of sys_linmat_lindef_static_omp (IGA_parametrization.py:1561-1650):
        the same inputs under their lower-case f2py names, plus num_threads."""
        d = self._flat()
        if COORDS is not None and np.shape(COORDS) == np.shape(self._COORDS):
            d["COORDS3D"] = COORDS
        if activeElem is not None:
            d["activeElement"] = np.asarray(activeElem)
        kw = {k.lower(): d[k] for k in F2PY_ARGS + F2PY_OPTIONAL if k in d}
        kw["num_threads"] = -1 if num_threads is None else num_threads
        return kw

    def _get_inputs4gradTotalwork(self, sol, ADJ, COORDS=None, activeElem=None, activeDir=np.ones(3, np.intp),
                                  computeWint=True, computeWext=True):
        """Positional argument list of gradadjointwork.gradlinelastwork_an
        (IGA_parametrization.py:1507-1527)."""
        from .gradadjointwork import inputs4gradtotalwork
        return inputs4gradtotalwork(self._flat(), sol, ADJ, COORDS=COORDS, activeElem=activeElem, activeDir=activeDir,
                                    computeWint=computeWint, computeWext=computeWext)

    def get_model_dict(self):
        """The same inputs by name (what yeti_amd.assembly.Assembler also accepts)."""
        return self._flat()

    def get_inputs4solution(self, U_inv):
        """Keyword arguments of reconstructionSOL.reconstruction (IGA_parametrization.py:664-682)."""
        d = self._flat()
        return {"u_inv": U_inv, "coords": self._COORDS, "ind_dof_free": self._ind_dof_free,
                "bc_target": d["bc_target"], "bc_target_nbelem": d["bc_target_nbelem"], "bc_values": d["bc_values"],
                "couplg_flag": False, "nb_bc": d["nb_bc"], "mcrd": self._mcrd, "nb_cp": self._nb_cp,
                "nb_dof_free": self._nb_dof_free}

    def get_inputs4massmat(self):
        """Positional argument list of massmtrx.build_cmassmatrix (IGA_parametrization.py:1720-1765)."""
        d = self._flat()
        return [d["activeElement"], d["nb_data"], self._COORDS, d["IEN"], self._elementsByPatch, self._Nkv, d["Ukv"],
                self._Nijk, d["weight"], self._Jpqr, d["ELT_TYPE"], d["PROPS"], d["JPROPS"],
                self._MATERIAL_PROPERTIES[2], d["TENSOR"], self._ind_dof_free, self._nb_dof_free, self._mcrd,
                self._NBPINT, self._nnode, self._nb_patch, self.nb_elem, self._nb_cp, self._nb_dof_tot]

    def get_inputs4area(self, COORDS=None, activepatch=None):
        """Positional argument list of volume.computevolume / computegradvolume
        (IGA_parametrization.py:1286-1312): every patch active unless activepatch (one flag per
        patch) is given."""
        d = self._flat()
        X = self._COORDS if COORDS is None or np.shape(COORDS) != np.shape(self._COORDS) else COORDS
        if activepatch is None or np.size(activepatch) != self._nb_patch:
            activepatch = np.ones(self._nb_patch, np.int64)
        return [np.asarray(activepatch), X, d["IEN"], self._elementsByPatch, self._Nkv, d["Ukv"], self._Nijk,
                d["weight"], self._Jpqr, d["ELT_TYPE"], d["PROPS"], d["JPROPS"], d["TENSOR"], self._mcrd,
                self._NBPINT, self._nnode, self._nb_patch, self.nb_elem, self._nb_cp]

    def get_inputs4evaldisp(self, sol, xi, numpatch=1):
        """Positional argument list of postproc.evaldisp (IGA_parametrization.py:908-938)."""
        d = self._flat()
        return [numpatch, xi, sol, self._COORDS, d["IEN"], self._elementsByPatch, self._Nkv, d["Ukv"], self._Nijk,
                d["weight"], self._Jpqr, d["ELT_TYPE"], self._MATERIAL_PROPERTIES[:2], d["TENSOR"], d["PROPS"],
                d["JPROPS"], self._nnode, self._nb_patch, self.nb_elem, self._nb_cp, self._mcrd]

    def get_inputs4postprocVTU(self, filename, sol, nb_ref=np.ones(3, dtype=int), Flag=np.array([True, True, True]),
                               output_path="results"):
        """Positional argument list of postproc.generatevtu (IGA_parametrization.py:717-752)."""
        d = self._flat()
        nb_ref = np.maximum(nb_ref, np.ones(3, dtype=int))
        return [filename, output_path, Flag, nb_ref, sol, self._COORDS, d["IEN"], self._elementsByPatch, self._Nkv,
                d["Ukv"], self._Nijk, d["weight"], self._Jpqr, d["ELT_TYPE"], self._MATERIAL_PROPERTIES[:2],
                d["TENSOR"], d["PROPS"], d["JPROPS"], self._nnode, self._nb_patch, self.nb_elem, self._nb_cp,
                self._mcrd]
