"""Model builders shared by the tests: extra load cases, multi-patch models, inactive elements.

All of them edit the f2py argument dict produced by yeti_amd.geometry.make_box_patch, keeping
YETI's conventions (1-based ids, Fortran shapes)."""
from __future__ import annotations

import copy

import numpy as np

from yeti_amd.geometry import make_box_patch


def face_elements(bp, face):
    """1-based ids of the elements of a box patch touching parametric face 1..6
    (1/2: xi min/max, 3/4: eta min/max, 5/6: zeta min/max)."""
    n1, n2, n3 = bp.nel
    e = np.arange(n1 * n2 * n3)
    e1, e2, e3 = e % n1, (e // n1) % n2, e // (n1 * n2)
    sel = {1: e1 == 0, 2: e1 == n1 - 1, 3: e2 == 0, 4: e2 == n2 - 1, 5: e3 == 0, 6: e3 == n3 - 1}[face]
    return (e[sel] + 1).astype(np.int32)


def set_loads(args, loads, nodal_dist=None):
    """Replace the load tables. ``loads`` = list of (JDLType, ADLMAG, targets, additional_infos)."""
    a = dict(args)
    a["indDLoad"] = np.concatenate([np.asarray(t, np.int32) for _, _, t, _ in loads]).astype(np.int32)
    a["JDLType"] = np.array([j for j, _, _, _ in loads], np.int32)
    a["ADLMAG"] = np.array([m for _, m, _, _ in loads], np.float64)
    a["load_target_nbelem"] = np.array([len(t) for _, _, t, _ in loads], np.int32)
    infos = [np.asarray(x, np.float64) for _, _, _, x in loads]
    a["load_additionalInfos"] = np.concatenate(infos) if infos else np.array([], np.float64)
    a["nb_load_additionalInfos"] = np.array([x.size for x in infos], np.int32)
    a["nb_load"] = len(loads)
    if nodal_dist is not None:
        a["nodal_dist"] = np.asfortranarray(nodal_dist)
    return a


def concat_patches(p1, p2):
    """Two box patches as one two-patch model (CPs of p2 renumbered after p1's)."""
    a, b = p1.args, p2.args
    ncp_a = a["COORDS3D"].shape[1]
    ne_a = int(a["nb_elem"])
    m = {}
    m["activeElement"] = np.concatenate([a["activeElement"], b["activeElement"]])
    m["nb_data"] = float(a["nb_data"] + b["nb_data"])
    m["COORDS3D"] = np.asfortranarray(np.hstack([a["COORDS3D"], b["COORDS3D"]]))
    m["IEN"] = np.concatenate([a["IEN"], b["IEN"] + ncp_a]).astype(np.int32)
    m["nb_elem_patch"] = np.array([a["nb_elem"], b["nb_elem"]], np.int32)
    m["Nkv"] = np.asfortranarray(np.hstack([a["Nkv"], b["Nkv"]]))
    m["Ukv"] = np.concatenate([a["Ukv"], b["Ukv"]])
    m["Nijk"] = np.asfortranarray(np.hstack([a["Nijk"], b["Nijk"]]))
    m["weight"] = np.concatenate([a["weight"], b["weight"]])
    m["Jpqr"] = np.asfortranarray(np.hstack([a["Jpqr"], b["Jpqr"]]))
    m["ELT_TYPE"] = a["ELT_TYPE"] + b["ELT_TYPE"]
    m["PROPS"] = np.concatenate([a["PROPS"], b["PROPS"]])
    m["JPROPS"] = np.concatenate([a["JPROPS"], b["JPROPS"]])
    m["MATERIAL_PROPERTIES"] = np.asfortranarray(np.hstack([a["MATERIAL_PROPERTIES"], b["MATERIAL_PROPERTIES"]]))
    m["n_mat_props"] = np.concatenate([a["n_mat_props"], b["n_mat_props"]])
    m["RHO"] = np.concatenate([a["RHO"], b["RHO"]])
    m["TENSOR"] = a["TENSOR"] + b["TENSOR"]
    # loads: p1's then p2's (element ids shifted)
    m["indDLoad"] = np.concatenate([a["indDLoad"], b["indDLoad"] + ne_a]).astype(np.int32)
    m["JDLType"] = np.concatenate([a["JDLType"], b["JDLType"]])
    m["ADLMAG"] = np.concatenate([a["ADLMAG"], b["ADLMAG"]])
    m["load_target_nbelem"] = np.concatenate([a["load_target_nbelem"], b["load_target_nbelem"]])
    m["load_additionalInfos"] = np.concatenate([a["load_additionalInfos"], b["load_additionalInfos"]])
    m["nb_load_additionalInfos"] = np.concatenate([a["nb_load_additionalInfos"], b["nb_load_additionalInfos"]])
    m["nb_load"] = int(a["nb_load"] + b["nb_load"])
    # bcs: clamp both x=0 faces
    m["bc_values"] = np.asfortranarray(np.hstack([a["bc_values"], b["bc_values"]]))
    m["bc_target"] = np.concatenate([a["bc_target"], b["bc_target"] + ncp_a]).astype(np.int32)
    m["bc_target_nbelem"] = np.concatenate([a["bc_target_nbelem"], b["bc_target_nbelem"]])
    m["nb_bc"] = int(a["nb_bc"] + b["nb_bc"])
    ncp = ncp_a + b["COORDS3D"].shape[1]
    fixed = np.zeros(3 * ncp, bool)
    off = 0
    for k in range(m["nb_bc"]):
        n = m["bc_target_nbelem"][k]
        d = int(m["bc_values"][0, k])
        fixed[(m["bc_target"][off:off + n] - 1) * 3 + d - 1] = True
        off += n
    free = np.where(~fixed)[0] + 1
    idf = np.zeros(3 * ncp, np.int32)
    idf[: free.size] = free
    m["ind_dof_free"] = idf
    m["nb_dof_free"] = int(free.size)
    m["MCRD"] = 3
    m["NBINT"] = np.concatenate([a["NBINT"], b["NBINT"]])
    m["NNODE"] = np.concatenate([a["NNODE"], b["NNODE"]])
    m["nodal_dist"] = np.zeros((1, ncp), order="F")
    m["nb_patch"] = 2
    m["nb_elem"] = ne_a + int(b["nb_elem"])
    m["nb_cp"] = ncp
    m["nb_dof_tot"] = 3 * ncp
    return m


def all_face_loads(bp, seed=7):
    """Every surface-load code (faces 1..6 x types 0,1,2,3,4) plus centrifugal (101) and volume
    (102) body loads on a box patch; returns the edited argument dict."""
    rng = np.random.default_rng(seed)
    ncp = bp.nb_cp
    loads = []
    for face in range(1, 7):
        for typ in (0, 1, 2, 3):
            loads.append((10 * face + typ, float(rng.uniform(-2, 2)), face_elements(bp, face), []))
        loads.append((10 * face + 4, 0.0, face_elements(bp, face), [2.0]))
    ne = bp.nb_elem
    body = np.arange(1, ne + 1, dtype=np.int32)[rng.random(ne) < 0.5]
    loads.append((101, 3.0, body, [0.1, 0.2, 0.3, 1.1, 0.4, 0.9]))
    loads.append((102, -1.5, np.arange(1, ne + 1, dtype=np.int32), [0.2, -0.7, 0.4]))
    nd = np.asfortranarray(rng.uniform(-1, 1, size=(2, ncp)))
    a = set_loads(bp.args, loads, nodal_dist=nd)
    a["RHO"] = np.array([7.8], np.float64)
    return a


def with_inactive(args, frac=0.3, seed=3):
    a = dict(args)
    rng = np.random.default_rng(seed)
    act = (rng.random(int(a["nb_elem"])) >= frac).astype(np.int32)
    a["activeElement"] = act
    n = int(a["NNODE"][0])
    nde = 3 * n * (3 * n + 1) // 2
    a["nb_data"] = float(act.sum() * nde)
    return a
