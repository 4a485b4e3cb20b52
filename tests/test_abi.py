"""C ABI surface checks that need no GPU: the library loads, exports every
function include/*.h declares, and rejects bad arguments before touching a
device."""
import ctypes
import os
import re

import numpy as np
import pytest

import brmerge_amd as bm
from brmerge_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("brmerge.h", "brmerge_gen.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(brm_\w+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_loads_and_exports_every_declared_symbol():
    L = bm.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTED)


def test_version_string():
    assert b"gfx950" in bm.lib().brm_version()


def test_params_roundtrip_and_rejects_unknown():
    L = bm.lib()
    assert L.brm_set_param(b"no_such_knob", 1) == _lib.BRM_E_INVALID
    assert L.brm_set_param(b"med_cap", 10 ** 9) == _lib.BRM_E_INVALID
    old = bm.get_param("leaf_cap")
    bm.set_param("leaf_cap", 123)
    assert bm.get_param("leaf_cap") == 123
    bm.set_param("leaf_cap", old)


def test_null_and_negative_arguments_rejected_without_device():
    L = bm.lib()
    r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nnz = ctypes.c_int()
    st = L.brm_spgemm_precise_i32(None, None, None, None, None, None, ctypes.byref(r), ctypes.byref(c),
                                  ctypes.byref(v), -1, 0, 0, ctypes.byref(nnz), None)
    assert st == _lib.BRM_E_INVALID
    st = L.brm_spgemm_upper_i32(None, None, None, None, None, None, None, None, None, 1, 1, 1,
                                ctypes.byref(nnz), None)
    assert st == _lib.BRM_E_INVALID
    assert "null" in _lib.last_error()


def test_timing_struct_layout_matches_header():
    assert ctypes.sizeof(_lib.PreciseTimingC) == 12 * 8
    assert ctypes.sizeof(_lib.UpperTimingC) == 14 * 8
    t = bm.Precise_Timing()
    assert set(bm.Precise_Timing.FIELDS) <= set(t.as_dict())


def test_csr_equality_semantics():
    A = bm.gen.er(20, 30, 4, 0)
    B = bm.CSR(A.rpt.copy(), A.col.copy(), A.val + 1e-12, A.M, A.N)
    assert A == B
    B.val[3] += 1e-3
    assert not (A == B)
    C = bm.gen.er(20, 30, 5, 0)
    with pytest.raises(RuntimeError):
        A == C  # noqa: B015  (reference operator== throws on nnz mismatch)


def test_pair_struct_layout_and_batch_argument_checks():
    assert ctypes.sizeof(_lib.PairC) == 80
    L = bm.lib()
    r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nnz = ctypes.c_int64()
    row_off = np.zeros(2, np.int64)
    assert L.brm_spgemm_batched_dev(None, 1, ctypes.byref(r), ctypes.byref(c), ctypes.byref(v),
                                    ctypes.byref(nnz), row_off.ctypes.data, None, None) == _lib.BRM_E_INVALID
    bad = (_lib.PairC * 1)()
    bad[0] = _lib.PairC(None, None, None, 0, None, None, None, 0, -1, 0, 0)
    assert L.brm_spgemm_batched_dev(ctypes.cast(bad, ctypes.c_void_p), 1, ctypes.byref(r), ctypes.byref(c),
                                    ctypes.byref(v), ctypes.byref(nnz), row_off.ctypes.data, None,
                                    None) == _lib.BRM_E_INVALID
    bad[0] = _lib.PairC(None, None, None, 5, None, None, None, 0, 3, 2, 2)
    h = ctypes.c_void_p()
    assert L.brm_batch_plan_create(ctypes.cast(bad, ctypes.c_void_p), 1, 0, ctypes.byref(h),
                                   None) == _lib.BRM_E_INVALID
    assert "null" in _lib.last_error()
