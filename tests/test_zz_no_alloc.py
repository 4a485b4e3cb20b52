"""brm_spgemm_precise_dev_noalloc: BRMerge-Precise into caller-owned outputs
with every scratch buffer carved from the library's fixed device array (no
device allocation inside the call).  Its C must equal brm_spgemm_precise_dev's
bit for bit, on two calls in a row (the second on a smaller product, so
stale contents of the shared scratch from the first call would show), and
it must refuse products whose scratch or output exceeds its limits."""
import ctypes

import numpy as np
import pytest

from tests._util import bits

pytestmark = pytest.mark.gpu


def _fn(bm):
    L = bm.lib()
    f = L.brm_spgemm_precise_dev_noalloc
    P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    f.argtypes = [P, P, P, I64, P, P, P, I64, I, I, I, P, P, P, I64, ctypes.POINTER(I64), P]
    f.restype = I
    L.brm_noalloc_scratch_bytes.restype = ctypes.c_uint64
    return f


def _noalloc(bm, dA, dB, cap):
    f = _fn(bm)
    crpt = bm.DeviceArray.alloc(dA.M + 1, np.int64)
    ccol = bm.DeviceArray.alloc(max(cap, 1), np.int32)
    cval = bm.DeviceArray.alloc(max(cap, 1), np.float64)
    nnz = ctypes.c_int64(-7)
    st = f(dA.rpt.ptr, dA.col.ptr, dA.val.ptr, dA.nnz, dB.rpt.ptr, dB.col.ptr, dB.val.ptr, dB.nnz,
           dA.M, dA.N, dB.N, crpt.ptr, ccol.ptr, cval.ptr, cap, ctypes.byref(nnz), None)
    return st, nnz.value, crpt, ccol, cval


def _check_same(bm, dA, dB, slack):
    ref = bm.spgemm_dev(dA, dB)
    R = ref.download()
    ref.free()
    st, nnz, crpt, ccol, cval = _noalloc(bm, dA, dB, R.nnz + slack)
    try:
        assert st == 0, bm._lib.last_error()
        assert nnz == R.nnz
        np.testing.assert_array_equal(crpt.to_host(), np.asarray(R.rpt, np.int64))
        np.testing.assert_array_equal(ccol.to_host(0, nnz), R.col)
        assert np.array_equal(bits(cval.to_host(0, nnz)), bits(R.val))
    finally:
        for a in (crpt, ccol, cval):
            a.free()
    return R.nnz


def test_noalloc_matches_precise_dev_twice(gpu):
    bm = gpu
    big = bm.gen.rmat(14, seed=5)      # every tier, long rows with several passes
    small = bm.gen.er(3000, 3000, 7, 9)
    dbig, dsmall = bm.DeviceCSR.upload(big), bm.DeviceCSR.upload(small)
    try:
        n1 = _check_same(bm, dbig, dbig, 0)
        n2 = _check_same(bm, dsmall, dsmall, 1000)
        assert n2 < n1
        # the long-row engine under the fixed buffer, both engines
        old = bm.get_param("group_cfg")
        try:
            for eng in (0, 4):
                bm.set_param("group_cfg", eng)
                _check_same(bm, dbig, dbig, 5)
        finally:
            bm.set_param("group_cfg", old)
    finally:
        dbig.free()
        dsmall.free()


def test_noalloc_refuses_beyond_limits(gpu):
    bm = gpu
    A = bm.gen.er(2000, 2000, 6, 1)
    dA = bm.DeviceCSR.upload(A)
    try:
        ref = bm.spgemm_dev(dA, dA)
        need = ref.nnz
        ref.free()
        # output capacity too small: BRM_E_OVERFLOW and the needed size
        st, nnz, *arrs = _noalloc(bm, dA, dA, need - 1)
        for a in arrs:
            a.free()
        assert st == 3 and nnz == need
        # still correct afterwards
        _check_same(bm, dA, dA, 0)
    finally:
        dA.free()
    # more rows than the fixed scratch can bin: refused before any kernel runs
    limit = int(bm.lib().brm_noalloc_scratch_bytes())
    M = limit // 40
    rpt = bm.DeviceArray.from_host(np.zeros(M + 1, np.int32))
    B = bm.DeviceCSR.upload(bm.gen.er(8, 8, 2, 0))
    f = _fn(bm)
    out = bm.DeviceArray.alloc(M + 1, np.int64)
    nnz = ctypes.c_int64(0)
    try:
        st = f(rpt.ptr, None, None, 0, B.rpt.ptr, B.col.ptr, B.val.ptr, B.nnz, M, 8, 8, out.ptr, None, None, 0,
               ctypes.byref(nnz), None)
        assert st == 2 and "scratch" in bm._lib.last_error()
    finally:
        for x in (rpt, out):
            x.free()
        B.free()
