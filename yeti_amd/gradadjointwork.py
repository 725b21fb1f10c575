"""Drop-in for YETI's f2py module ``optim.gradadjointwork`` (src/yeti_iga/optim/CMakeLists.txt:68-69):
``gradlinelastwork_an`` (optim/gradtotalwork_fullAN.f90:17-300), called as

    dervWint, dervWext = gradlinelastwork_an(*iga._get_inputs4gradTotalwork(SOL.T, ADJ, activeElem=...,
                                             activeDir=..., computeWint=..., computeWext=...))

(OPT_modelling.py:447-481, IGA_parametrization.py:1507-1527). The derivatives of the internal
and external work of adjoint fields with respect to the control-point coordinates of solid (U1)
patches are computed on the GPU (``yeti_grad_lin_elast_work``):

    gradWint[k] = -ADJ_k^T (dK/dX) SOL,   gradWext[k] = ADJ_k^T (dF/dX)

over the active elements, with the reference's external terms: the centrifugal body loads
(type 101), the k-th of which reads its axis points at values 6k+1..6k+6 of
load_additionalInfos (gradUELMAT_solid.f90:378-398); its surface-load term is a stub
(gradUELMAT_solid.f90:463-477) and stays zero here too. activeDir plays no part: the
reference's solid routine declares it and never reads it.
"""
from __future__ import annotations

import numpy as np
import torch

from .assembly import Assembler


def gradlinelastwork_an(computewint, computewext, sol, adj, activeelement, activedir, coords3d, ien, nb_elem_patch,
                        nkv, ukv, nijk, weight, jpqr, elt_type, props, jprops, material_properties, rho, tensor,
                        inddload, jdltype, adlmag, load_target_nbelem, load_additionalinfos, nbint, nnode, nadj=None,
                        mcrd=None, nb_load=None, nb_patch=None, nb_elem=None, nb_cp=None):
    """Returns (gradWint, gradWext), each (nadj, 3, nb_cp) float64 in Fortran order like the f2py
    routine. sol is (mcrd, nb_cp), adj (nadj, mcrd, nb_cp); material_properties is (2, nb_patch)
    (E, nu) and rho (nb_patch,), as _get_inputs4gradTotalwork passes them."""
    sol = np.asarray(sol, np.float64)
    adj = np.asarray(adj, np.float64)
    ncp = int(nb_cp) if nb_cp is not None else np.asarray(coords3d).shape[1]
    mc = int(mcrd) if mcrd is not None else sol.shape[0]
    if mc != 3:
        raise NotImplementedError("gradlinelastwork_an: only MCRD = 3 (solid) models are supported")
    if adj.ndim == 2:
        adj = adj[None]
    na = int(nadj) if nadj is not None else adj.shape[0]
    if sol.shape != (3, ncp) or adj.shape != (na, 3, ncp):
        raise ValueError("sol must be (3, nb_cp) and adj (nadj, 3, nb_cp) with nb_cp = %d, got %s and %s"
                         % (ncp, sol.shape, adj.shape))
    npatch = int(nb_patch) if nb_patch is not None else np.asarray(nb_elem_patch).size
    jdl = np.asarray(jdltype, np.int32).reshape(-1)
    nl = int(nb_load) if nb_load is not None else jdl.size
    jdl = jdl[:nl]
    infos = np.asarray(load_additionalinfos, np.float64).reshape(-1)
    # only the centrifugal loads enter the reference's gradient; keep them with their targets
    cnt = np.asarray(load_target_nbelem, np.int64).reshape(-1)[:nl]
    off = np.concatenate([[0], np.cumsum(cnt)])
    targets = np.asarray(inddload, np.int64).reshape(-1)
    keep = np.flatnonzero(jdl == 101)
    ind = [targets[off[i]:off[i + 1]] for i in keep]
    mat = np.asarray(material_properties, np.float64).reshape(2, -1, order="F")
    d = {
        "activeElement": activeelement, "nb_data": 1, "COORDS3D": coords3d, "IEN": ien,
        "nb_elem_patch": nb_elem_patch, "Nkv": nkv, "Ukv": ukv, "Nijk": nijk, "weight": weight, "Jpqr": jpqr,
        "ELT_TYPE": elt_type, "PROPS": props, "JPROPS": jprops, "MATERIAL_PROPERTIES": mat,
        "n_mat_props": np.full(npatch, 2, np.int32), "RHO": rho, "TENSOR": tensor,
        "indDLoad": np.concatenate(ind).astype(np.int32) if ind else np.zeros(1, np.int32),
        "JDLType": jdl[keep], "ADLMAG": np.asarray(adlmag, np.float64).reshape(-1)[:nl][keep],
        "load_target_nbelem": cnt[keep].astype(np.int32),
        # the reference's gradient reads the axis points by its own rule (6 values per
        # centrifugal load from the start of the array); other loads' counts play no part in it
        "load_additionalInfos": infos if infos.size else np.zeros(1),
        "nb_load_additionalInfos": np.zeros(keep.size, np.int32),
        "bc_values": np.zeros((2, 0)), "bc_target": np.zeros(1, np.int32), "bc_target_nbelem": np.zeros(0, np.int32),
        "ind_dof_free": np.arange(1, 3 * ncp + 1, dtype=np.int32), "nb_dof_free": 3 * ncp, "MCRD": 3,
        "NBINT": nbint, "NNODE": nnode, "nodal_dist": np.zeros((1, 1)),
        "nb_load": int(keep.size), "nb_bc": 0, "nb_n_dist": 1, "nb_cp_n_dist": 1, "nb_cp": ncp, "nb_dof_tot": 3 * ncp,
    }
    if nb_patch is not None:
        d["nb_patch"] = int(nb_patch)
    if nb_elem is not None:
        d["nb_elem"] = int(nb_elem)
    asm = Assembler(d)
    try:
        u = torch.from_numpy(np.ascontiguousarray(sol.T).reshape(-1))
        a = torch.from_numpy(np.ascontiguousarray(np.transpose(adj, (0, 2, 1))).reshape(na, -1))
        gi, ge = asm.grad_lin_elast_work(u, a, bool(computewint), bool(computewext))

        def back(t):   # nadj x (3 nb_cp) -> (nadj, 3, nb_cp)
            return np.asfortranarray(t.cpu().numpy().reshape(na, ncp, 3).transpose(0, 2, 1))
        return back(gi), back(ge)
    finally:
        asm.close()


def inputs4gradtotalwork(d, sol, ADJ, COORDS=None, activeElem=None, activeDir=np.ones(3, np.intp), computeWint=True,
                         computeWext=True):
    """Positional argument list of gradlinelastwork_an for a model given by its system inputs
    ``d`` (the argument dict of sys_linmat_lindef_static), in the order of
    IGA_parametrization._get_inputs4gradTotalwork (IGA_parametrization.py:1507-1527): a COORDS of
    another shape than the model's, or an activeElem of another size, falls back to the
    model's own (all elements active)."""
    X = np.asarray(d["COORDS3D"])
    if COORDS is None or np.shape(COORDS) != X.shape:
        COORDS = X
    ne = int(d["nb_elem"]) if "nb_elem" in d else int(np.asarray(d["activeElement"]).size)
    if activeElem is None or np.size(activeElem) != ne:
        activeElem = np.ones(ne, np.intp)
    mat = np.asarray(d["MATERIAL_PROPERTIES"], np.float64)
    nadj = np.size(ADJ, 0)
    npatch = int(np.asarray(d["nb_elem_patch"]).size)
    return [computeWint, computeWext, sol, ADJ, activeElem, activeDir, COORDS, d["IEN"], d["nb_elem_patch"], d["Nkv"],
            d["Ukv"], d["Nijk"], d["weight"], d["Jpqr"], d["ELT_TYPE"], d["PROPS"], d["JPROPS"], mat[:2, :],
            np.asarray(d["RHO"]), d["TENSOR"], d["indDLoad"], d["JDLType"], d["ADLMAG"], d["load_target_nbelem"],
            d["load_additionalInfos"], d["NBINT"], d["NNODE"], nadj, 3, int(np.asarray(d["JDLType"]).size), npatch, ne,
            int(X.shape[1])]
