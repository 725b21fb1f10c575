"""The oracle's marshalling (oracle/inputs.py) against the product's (yeti_amd/_abi.py).

The two are written independently: the oracle reads the struct's layout from the header and
infers sizes from the first array Fortran declares with each dimension; the product declares
its ctypes struct by hand and infers from other arrays. On every model here they must give the
same struct layout, the same scalars and the same array contents, field by field, and the
oracle's patch slicing must match the product's.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import inputs as OI
from tests.helpers import all_face_loads, concat_patches, with_inactive
from yeti_amd import _abi
from yeti_amd.geometry import make_box_patch
from yeti_amd.model import IGAparametrization

HERE = os.path.dirname(os.path.abspath(__file__))
CENTRIF2 = os.path.join(HERE, "golden", "models", "centrif_two_patches", "centrif_U1_U1_2loads")


def _models():
    bp1 = make_box_patch(nel=(4, 3, 2), p=1)
    bp2 = make_box_patch(nel=(3, 2, 2), p=2)
    m = IGAparametrization(filename=CENTRIF2)
    m.refine(np.full((3, m.nb_patch), 1, np.intp), np.full((3, m.nb_patch), 1, np.intp))
    return {
        "box_p1": (bp1.args, None),
        "box_p2_all_loads": (all_face_loads(bp2), None),
        "two_patches": (concat_patches(bp1, bp2), None),
        "inactive": (with_inactive(bp2.args), None),
        "centrif_positional": (None, m.get_inputs4system_elemStorage()),
    }


MODELS = _models()


def _both(name):
    kw, pos = MODELS[name]
    if pos is not None:
        return OI.to_dict(pos), _abi.args_to_dict(pos, {})
    return OI.to_dict((), kw), _abi.args_to_dict((), kw)


def test_struct_layout_matches_header():
    """The product's hand-written struct has the header's fields, in order, at the same
    offsets as the struct the oracle builds from the header."""
    names = [n for n, _, _ in OI.header_fields()]
    assert [f[0] for f in _abi.YetiSystemInputs._fields_] == names
    for n in names:
        assert getattr(_abi.YetiSystemInputs, n).offset == getattr(OI.OracleSystemInputs, n).offset, n
    assert C.sizeof(_abi.YetiSystemInputs) == C.sizeof(OI.OracleSystemInputs)
    assert OI.YETI_LAYOUT_PACKED == _abi.YETI_LAYOUT_PACKED
    assert OI.YETI_LAYOUT_SLOTTED == _abi.YETI_LAYOUT_SLOTTED


def _array_len(s, name, fields):
    """Element count of pointer field `name`: its len_ field, else the dimension the header's
    comment gives it, read from the struct's own scalars."""
    if "len_" + name in fields:
        return int(getattr(s, "len_" + name))
    per = {"activeElement": s.nb_elem, "COORDS3D": 3 * s.nb_cp, "nb_elem_patch": s.nb_patch,
           "Nkv": 3 * s.nb_patch, "Nijk": 3 * s.nb_elem, "Jpqr": 3 * s.nb_patch, "JPROPS": s.nb_patch,
           "MATERIAL_PROPERTIES": s.n_mat_props_max * s.nb_patch, "n_mat_props": s.nb_patch,
           "RHO": s.nb_patch, "JDLType": s.nb_load, "ADLMAG": s.nb_load, "load_target_nbelem": s.nb_load,
           "nb_load_additionalInfos": s.nb_load, "bc_values": 2 * s.nb_bc, "bc_target_nbelem": s.nb_bc,
           "ind_dof_free": s.nb_dof_tot, "NBINT": s.nb_patch, "NNODE": s.nb_patch,
           "nodal_dist": s.nb_n_dist * s.nb_cp_n_dist}
    return int(per[name])


@pytest.mark.parametrize("name", sorted(MODELS))
def test_packers_agree(name):
    d_o, d_p = _both(name)
    assert sorted(d_o) == sorted(d_p)
    so, ko = OI.pack(d_o)
    sp_, kp = _abi.pack_system_inputs(d_p)
    fields = {n: (kind, base) for n, kind, base in OI.header_fields()}
    for n, (kind, base) in fields.items():
        if kind == "val":
            assert getattr(so, n) == getattr(sp_, n), n
    for n, (kind, base) in fields.items():
        if kind != "ptr":
            continue
        k = _array_len(so, n, fields)
        assert k == _array_len(sp_, n, fields), n
        if base == "char":
            a = C.string_at(C.cast(getattr(so, n), C.c_void_p), k)
            b = C.string_at(C.cast(getattr(sp_, n), C.c_void_p), k)
            assert a == b, n
            continue
        ct = {"int32_t": C.c_int32, "double": C.c_double}[base]
        pa = C.cast(getattr(so, n), C.POINTER(ct))
        pb = C.cast(getattr(sp_, n), C.POINTER(ct))
        if k == 0:
            continue
        a = np.ctypeslib.as_array(pa, shape=(k,))
        b = np.ctypeslib.as_array(pb, shape=(k,))
        assert np.array_equal(a, b), n


def test_sizes_inferred_when_omitted():
    """Without the optional sizes both packers infer the same values from the arrays."""
    kw = dict(concat_patches(make_box_patch(nel=(2, 2, 2), p=1), make_box_patch(nel=(2, 3, 2), p=2)))
    for n in OI.OPTIONAL_SIZES:
        kw.pop(n, None)
    so, _ = OI.pack(OI.to_dict((), kw))
    sp_, _ = _abi.pack_system_inputs(_abi.args_to_dict((), kw))
    for n in OI.OPTIONAL_SIZES:
        assert getattr(so, n) == getattr(sp_, n), n
    assert so.nb_patch == 2 and so.nb_elem == 8 + 12


def test_patch_slice_matches_product():
    from yeti_amd.postproc import _patch
    a = concat_patches(make_box_patch(nel=(4, 3, 2), p=1), make_box_patch(nel=(3, 2, 2), p=2))
    for ip in (1, 2):
        o = OI.patch_slice(ip, a["IEN"], a["nb_elem_patch"], a["Nkv"], a["Ukv"], a["Nijk"], a["weight"],
                           a["Jpqr"], a["NNODE"])
        s, k = _patch(ip, a["IEN"], a["nb_elem_patch"], a["Nkv"], a["Ukv"], a["Nijk"], a["weight"], a["Jpqr"],
                      a["NNODE"])
        assert list(o["p"]) == list(s.p) and list(o["nkv"]) == list(s.nkv)
        assert o["nelem"] == s.nelem and o["nnode"] == s.nnode
        for dd in range(3):
            assert np.array_equal(o["U"][dd], k["U"][dd])
        for key in ("nijk", "ien", "w"):
            assert np.array_equal(o[key], k[key]), key
