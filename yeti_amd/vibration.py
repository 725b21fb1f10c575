"""Drop-in for YETI's f2py module ``vibration`` (src/yeti_iga/optim/CMakeLists.txt:40-43):
``gradvibration_an`` (optim/gradvibration_fullAN.f:22-189), called as
``vibration.gradvibration_an(*iga.get_inputs4gradVibration(VECT, vals, activeElem=...,
activeDir=...))`` (OPT_modelling.py:823-854, IGA_parametrization.py:1354-1392). The shape
gradient of the natural frequencies of solid (U1) patches is computed on the GPU
(``yeti_asm_run_vib_grad``).
"""
from __future__ import annotations

import numpy as np

from .assembly import Assembler


def gradvibration_an(vect, vals, activeelement, activedir, coords3d, ien, nb_elem_patch, nkv, ukv, nijk, weight, jpqr,
                     elt_type, props, jprops, material_properties, rho, tensor, mcrd, nbint, nnode, nb_frq=None,
                     nb_patch=None, nb_elem=None, nb_cp=None):
    """Returns gradV (nb_frq, 3, nb_cp), column-major like the f2py routine: d lambda_f / d P_cp
    over the active elements, divided per mode by u_f^T M u_f of those elements. vect is
    (nb_frq, 3, nb_cp), vals (nb_frq,). activedir plays no part (the reference's element routine
    does not read it)."""
    npatch = int(nb_patch) if nb_patch is not None else np.asarray(nb_elem_patch).size
    if int(mcrd) != 3:
        raise NotImplementedError("gradvibration_an: only MCRD = 3 (solid) models are supported")
    vals = np.asarray(vals, np.float64).reshape(-1)
    nf = int(nb_frq) if nb_frq is not None else vals.size
    vect = np.asarray(vect, np.float64)
    ncp = int(nb_cp) if nb_cp is not None else np.asarray(coords3d).shape[1]
    if vect.shape != (nf, 3, ncp):
        raise ValueError("vect must be (nb_frq, 3, nb_cp) = (%d, 3, %d), got %s" % (nf, ncp, vect.shape))
    mat = np.asarray(material_properties, np.float64).reshape(2, -1, order="F")
    d = {
        "activeElement": activeelement, "nb_data": 1, "COORDS3D": coords3d, "IEN": ien,
        "nb_elem_patch": nb_elem_patch, "Nkv": nkv, "Ukv": ukv, "Nijk": nijk, "weight": weight, "Jpqr": jpqr,
        "ELT_TYPE": elt_type, "PROPS": props, "JPROPS": jprops, "MATERIAL_PROPERTIES": mat,
        "n_mat_props": np.full(npatch, 2, np.int32), "RHO": rho, "TENSOR": tensor,
        # loads and boundary conditions play no part in the gradient
        "indDLoad": np.zeros(1, np.int32), "JDLType": np.zeros(0, np.int32), "ADLMAG": np.zeros(0),
        "load_target_nbelem": np.zeros(0, np.int32), "load_additionalInfos": np.zeros(1),
        "nb_load_additionalInfos": np.zeros(0, np.int32), "bc_values": np.zeros((2, 0)),
        "bc_target": np.zeros(1, np.int32), "bc_target_nbelem": np.zeros(0, np.int32),
        "ind_dof_free": np.arange(1, 3 * ncp + 1, dtype=np.int32), "nb_dof_free": 3 * ncp, "MCRD": 3,
        "NBINT": nbint, "NNODE": nnode, "nodal_dist": np.zeros((1, 1)),
        "nb_load": 0, "nb_bc": 0, "nb_n_dist": 1, "nb_cp_n_dist": 1, "nb_cp": ncp, "nb_dof_tot": 3 * ncp,
    }
    if nb_patch is not None:
        d["nb_patch"] = int(nb_patch)
    if nb_elem is not None:
        d["nb_elem"] = int(nb_elem)
    asm = Assembler(d)
    try:
        g = asm.run_vib_grad(vect, vals)
        return np.asfortranarray(g.cpu().numpy())
    finally:
        asm.close()
