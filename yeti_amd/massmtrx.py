"""Drop-in for YETI's f2py module ``massmtrx`` (src/yeti_iga/CMakeLists.txt:74-77):
``build_cmassmatrix`` (compute_CMASSMatrix.f:22-215), called as
``data, row, col = build_cmassmatrix(*iga.get_inputs4massmat())`` (OPT_modelling.py:369-372).
The consistent mass matrix of solid (U1) patches is built on the GPU in the stiffness COO layout.
"""
from __future__ import annotations

import numpy as np

from .assembly import Assembler


def build_cmassmatrix(activeelement, nb_data, coords3d, ien, nb_elem_patch, nkv, ukv, nijk, weight, jpqr,
                      elt_type, props, jprops, rho, tensor, ind_dof_free, nb_dof_free, mcrd, nbint, nnode,
                      nb_patch=None, nb_elem=None, nb_cp=None, nb_dof_tot=None):
    """Returns (Mdata f8, Mrow i4, Mcol i4) of nb_data entries, like the f2py routine."""
    npatch = int(nb_patch) if nb_patch is not None else np.asarray(nb_elem_patch).size
    d = {
        "activeElement": activeelement, "nb_data": int(nb_data), "COORDS3D": coords3d, "IEN": ien,
        "nb_elem_patch": nb_elem_patch, "Nkv": nkv, "Ukv": ukv, "Nijk": nijk, "weight": weight, "Jpqr": jpqr,
        "ELT_TYPE": elt_type, "PROPS": props, "JPROPS": jprops,
        # the stiffness material and the loads and boundary conditions play no part in the mass
        "MATERIAL_PROPERTIES": np.vstack([np.ones(npatch), np.zeros(npatch)]),
        "n_mat_props": np.full(npatch, 2, np.int32), "RHO": rho, "TENSOR": tensor,
        "indDLoad": np.zeros(1, np.int32), "JDLType": np.zeros(0, np.int32), "ADLMAG": np.zeros(0),
        "load_target_nbelem": np.zeros(0, np.int32), "load_additionalInfos": np.zeros(1),
        "nb_load_additionalInfos": np.zeros(0, np.int32), "bc_values": np.zeros((2, 0)),
        "bc_target": np.zeros(1, np.int32), "bc_target_nbelem": np.zeros(0, np.int32),
        "ind_dof_free": ind_dof_free, "nb_dof_free": int(nb_dof_free), "MCRD": int(mcrd), "NBINT": nbint,
        "NNODE": nnode, "nodal_dist": np.zeros((1, 1)),
    }
    for k, v in (("nb_patch", nb_patch), ("nb_elem", nb_elem), ("nb_cp", nb_cp), ("nb_dof_tot", nb_dof_tot)):
        if v is not None:
            d[k] = int(v)
    d["nb_load"] = 0
    d["nb_bc"] = 0
    d["nb_n_dist"] = 1
    d["nb_cp_n_dist"] = 1
    asm = Assembler(d)
    try:
        M, r, c = asm.run_mass()
        n = int(nb_data)
        return M[:n].cpu().numpy(), r[:n].cpu().numpy(), c[:n].cpu().numpy()
    finally:
        asm.close()
