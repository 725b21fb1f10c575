"""Drop-in for YETI's f2py module ``volume``: ``computevolume`` (optim/volume.f90:19-210) and
``computegradvolume`` (optim/gradvolume.f90:16-355), called as
``volume.computevolume(*iga.get_inputs4area(activepatch=...))`` (OPT_modelling.py:508-530,
IGA_parametrization.py:1286-1312). The volume of the listed solid (U1) patches and its shape
derivative are computed on the GPU (``yeti_asm_volume_grad``).
"""
from __future__ import annotations

import re

import numpy as np

from .assembly import Assembler


def _assembler(listpatch, coords3d, ien, nb_elem_patch, nkv, ukv, nijk, weight, jpqr, elt_type, props, jprops, tensor,
               mcrd, nbint, nnode, nb_patch=None, nb_elem=None, nb_cp=None):
    nep = np.asarray(nb_elem_patch, np.int64).reshape(-1)
    npatch = int(nb_patch) if nb_patch is not None else nep.size
    if int(mcrd) != 3:
        raise NotImplementedError("volume: only MCRD = 3 (solid) models are supported")
    lp = np.asarray(listpatch, np.int64).reshape(-1)
    if lp.size != npatch:
        raise ValueError("listpatch must hold one flag per patch (%d)" % npatch)
    types = re.findall(r"U\d+", str(elt_type))
    if len(types) != npatch:
        raise ValueError("ELT_TYPE names %d element types for %d patches" % (len(types), npatch))
    for i in range(npatch):
        if lp[i] and types[i] not in ("U1",):
            raise NotImplementedError("volume: patch %d is %s; only solid U1 patches are supported"
                                      % (i + 1, types[i]))
    ncp = int(nb_cp) if nb_cp is not None else np.asarray(coords3d).shape[1]
    active = np.repeat((lp != 0).astype(np.int32), nep[:npatch])
    d = {
        "activeElement": active, "nb_data": 1, "COORDS3D": coords3d, "IEN": ien,
        "nb_elem_patch": nep, "Nkv": nkv, "Ukv": ukv, "Nijk": nijk, "weight": weight, "Jpqr": jpqr,
        "ELT_TYPE": elt_type, "PROPS": props, "JPROPS": jprops,
        # the material plays no part in the volume
        "MATERIAL_PROPERTIES": np.tile(np.array([[1.0], [0.25]]), (1, npatch)),
        "n_mat_props": np.full(npatch, 2, np.int32), "RHO": np.ones(npatch), "TENSOR": tensor,
        "indDLoad": np.zeros(1, np.int32), "JDLType": np.zeros(0, np.int32), "ADLMAG": np.zeros(0),
        "load_target_nbelem": np.zeros(0, np.int32), "load_additionalInfos": np.zeros(1),
        "nb_load_additionalInfos": np.zeros(0, np.int32), "bc_values": np.zeros((2, 0)),
        "bc_target": np.zeros(1, np.int32), "bc_target_nbelem": np.zeros(0, np.int32),
        "ind_dof_free": np.arange(1, 3 * ncp + 1, dtype=np.int32), "nb_dof_free": 3 * ncp, "MCRD": 3,
        "NBINT": nbint, "NNODE": nnode, "nodal_dist": np.zeros((1, 1)),
        "nb_load": 0, "nb_bc": 0, "nb_n_dist": 1, "nb_cp_n_dist": 1, "nb_cp": ncp, "nb_dof_tot": 3 * ncp,
        "nb_patch": npatch,
    }
    if nb_elem is not None:
        d["nb_elem"] = int(nb_elem)
    return Assembler(d), ncp


def computevolume(*args, **kwargs):
    """Volume of the patches flagged in listpatch (signed Jacobian, as the reference)."""
    asm, _ = _assembler(*args, **kwargs)
    try:
        v, _ = asm.volume_grad()
        return float(v.item())
    finally:
        asm.close()


def computegradvolume(*args, **kwargs):
    """gradV (3, nb_cp), column-major like the f2py routine: dV/dP of the flagged patches."""
    asm, ncp = _assembler(*args, **kwargs)
    try:
        _, g = asm.volume_grad()
        return np.asfortranarray(g.cpu().numpy().reshape(ncp, 3).T)
    finally:
        asm.close()
