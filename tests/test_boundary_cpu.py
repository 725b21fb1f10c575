"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every function
include/yeti_amd.h declares, the f2py-signature marshalling, and the box generator's conventions.
No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from yeti_amd import _abi, _lib
from yeti_amd.geometry import make_box_patch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "yeti_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(yeti_[a-z0-9_]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_abi():
    names = header_functions()
    for n in ("yeti_sys_linmat_lindef_static", "yeti_asm_create", "yeti_asm_run_coo", "yeti_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, "libyeti_hip.so lacks %s" % missing
    # every declared symbol has a Python signature
    assert set(header_functions()) <= set(_lib.SIGNATURES)


def test_library_reports_no_device_without_gpu():
    L = _lib.lib()
    assert L.yeti_version() == 0x000100
    if not os.path.exists("/dev/kfd"):
        assert L.yeti_device_count() == 0
        with pytest.raises(RuntimeError):
            _lib.require_device()


def test_struct_layout_matches_header_order():
    txt = open(os.path.join(ROOT, "include", "yeti_amd.h")).read()
    body = txt[txt.index("typedef struct yeti_system_inputs {"):txt.index("} yeti_system_inputs;")]
    fields = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*;", body)
    assert fields == [f[0] for f in _abi.YetiSystemInputs._fields_]


def test_args_to_dict_errors_like_a_python_call():
    bp = make_box_patch(nel=(2, 1, 1), p=2)
    args = bp.f2py_args()
    d = _abi.args_to_dict(args)
    assert d["nb_dof_tot"] == bp.nb_dof_tot and d["MCRD"] == 3
    with pytest.raises(TypeError):
        _abi.args_to_dict(args[:5])
    with pytest.raises(TypeError):
        _abi.args_to_dict(args, {"bogus": 1})
    with pytest.raises(TypeError):
        _abi.args_to_dict(args, {"coords3d": 1})
    s, keep = _abi.pack_system_inputs(d)
    assert s.nb_cp == bp.nb_cp and s.nb_elem == bp.nb_elem and s.len_IEN == bp.nb_elem * 27
    assert s.nb_data == bp.nb_data and s.len_ELT_TYPE == 2 and s.len_TENSOR == 7


def test_generator_conventions():
    p = 2
    bp = make_box_patch(nel=(3, 2, 2), p=p, jitter=0.0, weight_range=(1.0, 1.0))
    a = bp.args
    n1, n2, n3 = bp.n_cp_dir
    n = (p + 1) ** 3
    ien = a["IEN"].reshape(-1, n)
    # local node 1 is the highest (i, j, k) corner of the element, xi fastest, reversed
    for e in range(bp.nb_elem):
        e1, e2, e3 = e % 3, (e // 3) % 2, e // 6
        cp0 = ien[e] - 1
        i1, i2, i3 = cp0 % n1, (cp0 // n1) % n2, cp0 // (n1 * n2)
        assert (i1[0], i2[0], i3[0]) == (e1 + p, e2 + p, e3 + p)
        assert (i1[-1], i2[-1], i3[-1]) == (e1, e2, e3)
        assert np.all(np.diff(i1[: p + 1]) == -1)
        assert tuple(a["Nijk"][:, e]) == (e1 + p + 1, e2 + p + 1, e3 + p + 1)
    # clamped face x = 0, free dofs ascending and 1-based
    free = a["ind_dof_free"][: a["nb_dof_free"]]
    assert np.all(np.diff(free) > 0) and free.min() >= 1
    X = a["COORDS3D"]
    fixed = np.setdiff1d(np.arange(1, bp.nb_dof_tot + 1), free)
    assert np.allclose(X[0, (fixed - 1) // 3], 0.0)
    # Greville control points reproduce the box
    assert np.isclose(X[0].max(), 10.0) and np.isclose(X[1].max(), 1.0)
    # traction on the 4 elements at x = 10, code 23
    assert a["JDLType"][0] == 23 and a["indDLoad"].size == n2 * n3 // (n2 * n3) * 4


def _counting_create(fn, d, extra):
    """Run a *_create_in entry point in counting mode (null arena: no device is touched)."""
    L = _lib.lib()
    s, keep = _abi.pack_system_inputs(d)
    out = ctypes.c_void_p()
    used = ctypes.c_size_t(0)
    rc = getattr(L, fn)(ctypes.byref(s), extra, None, 0, None, ctypes.byref(out), ctypes.byref(used))
    msg = L.yeti_last_error().decode()
    return rc, msg, used.value, out.value


@pytest.mark.parametrize("fn,extra", [("yeti_asm_create_in", 0), ("yeti_mf_create_in", 1)])
def test_set_up_rejects_out_of_range_ids_before_any_device_call(fn, extra):
    bp = make_box_patch(nel=(3, 2, 2), p=2)
    good = dict(bp.args)
    rc, msg, used, out = _counting_create(fn, good, extra)
    assert rc == 0 and used > 0 and out is None, msg

    def bad(key, idx, value):
        d = dict(good)
        arr = np.array(d[key], copy=True, order="F")
        arr.reshape(-1, order="F")[idx] = value
        d[key] = arr
        return d

    n1 = bp.n_cp_dir[0]
    cases = [
        (bad("IEN", 5, bp.nb_cp + 1), "IEN"),
        (bad("IEN", 0, 0), "IEN"),
        (bad("Nijk", 3, 2), "span"),                        # below p + 1 in xi
        (bad("Nijk", 4, bp.nel[1] + 3), "span"),            # beyond nk - p - 1 in eta
        (bad("indDLoad", 0, bp.nb_elem + 1), "load"),       # face load on a missing element
        (bad("bc_target", 0, bp.nb_cp + 7), "bc_target"),
        (bad("Ukv", n1, -1.0), "knot"),                     # decreasing knot vector
    ]
    for d, word in cases:
        rc, msg, _, out = _counting_create(fn, d, extra)
        assert rc == _abi.YETI_ERR_INVALID_ARGUMENT, (word, rc, msg)
        assert word in msg and out is None, msg
    d = dict(good)
    d["nb_elem_patch"] = np.array([bp.nb_elem + 1], np.int32)
    rc, msg, _, _ = _counting_create(fn, d, extra)
    assert rc == _abi.YETI_ERR_INVALID_ARGUMENT and "nb_elem" in msg


def test_host_drop_in_rejects_bad_connectivity_without_a_device():
    """The host-buffer drop-in validates before it uploads anything."""
    L = _lib.lib()
    bp = make_box_patch(nel=(2, 1, 1), p=2)
    d = dict(bp.args)
    ien = np.array(d["IEN"], copy=True)
    ien[7] = bp.nb_cp + 100
    d["IEN"] = ien
    s, keep = _abi.pack_system_inputs(d)
    nd = int(bp.nb_data)
    K = np.zeros(nd)
    r = np.zeros(nd, np.int32)
    c = np.zeros(nd, np.int32)
    F = np.zeros(bp.nb_dof_tot)
    p64 = ctypes.POINTER(ctypes.c_double)
    p32 = ctypes.POINTER(ctypes.c_int32)
    rc = L.yeti_sys_linmat_lindef_static(ctypes.byref(s), 0, K.ctypes.data_as(p64), r.ctypes.data_as(p32),
                                         c.ctypes.data_as(p32), F.ctypes.data_as(p64))
    assert rc == _abi.YETI_ERR_INVALID_ARGUMENT
    assert "IEN" in L.yeti_last_error().decode()


def test_gradient_argument_list_follows_the_reference_order():
    """inputs4gradtotalwork builds gradlinelastwork_an's positional list in the order of
    IGA_parametrization._get_inputs4gradTotalwork (IGA_parametrization.py:1516-1525), which is the
    drop-in's parameter order; the trailing sizes are the f2py optional arguments."""
    import inspect
    from yeti_amd.gradadjointwork import gradlinelastwork_an, inputs4gradtotalwork
    bp = make_box_patch(nel=(2, 1, 1), p=2)
    SOL = np.zeros((3, bp.nb_cp))
    ADJ = np.zeros((2, 3, bp.nb_cp))
    lst = inputs4gradtotalwork(dict(bp.args), SOL, ADJ, computeWint=False)
    names = list(inspect.signature(gradlinelastwork_an).parameters)
    assert names[:7] == ["computewint", "computewext", "sol", "adj", "activeelement", "activedir", "coords3d"]
    assert names[-6:] == ["nadj", "mcrd", "nb_load", "nb_patch", "nb_elem", "nb_cp"]
    assert len(lst) == len(names) == 33
    assert lst[0] is False and lst[2] is SOL and lst[3] is ADJ
    assert np.array_equal(lst[4], np.ones(bp.nb_elem)) and np.array_equal(lst[5], np.ones(3))
    assert lst[-6:] == [2, 3, 1, 1, bp.nb_elem, bp.nb_cp]
    # a wrong-sized activeElem or COORDS falls back to the model's, as the reference's marshaller
    lst = inputs4gradtotalwork(dict(bp.args), SOL, ADJ, COORDS=np.zeros((3, 2)), activeElem=np.ones(3))
    assert lst[6] is bp.args["COORDS3D"] and lst[4].size == bp.nb_elem
