"""ctypes wrapper of libbrm_xcheck.so: C = A*B by rocSPARSE (vendor
cross-check; test and bench infrastructure, not part of the product)."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "brmerge_amd", "libbrm_xcheck.so")
_L = None


def lib():
    global _L
    if _L is None:
        _L = ctypes.CDLL(PATH)
        P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        PP, PD = ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_double)
        _L.brm_xcheck_spgemm.argtypes = [P, P, P, I64, P, P, P, I64, I, I, I, PP, PP, PP,
                                         ctypes.POINTER(I64), PD, PD]
        _L.brm_xcheck_spgemm.restype = I
        _L.brm_xcheck_free.argtypes = [P]
        _L.brm_xcheck_free.restype = None
    return _L


def spgemm(bm, dA, dB):
    """rocSPARSE product of two DeviceCSR operands -> (host CSR, total_s, compute_s)."""
    L = lib()
    r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nnz = ctypes.c_int64()
    tot, comp = ctypes.c_double(), ctypes.c_double()
    st = L.brm_xcheck_spgemm(dA.rpt.ptr, dA.col.ptr, dA.val.ptr, dA.nnz, dB.rpt.ptr, dB.col.ptr, dB.val.ptr, dB.nnz,
                             dA.M, dA.N, dB.N, ctypes.byref(r), ctypes.byref(c), ctypes.byref(v), ctypes.byref(nnz),
                             ctypes.byref(tot), ctypes.byref(comp))
    if st != 0:
        raise RuntimeError("rocSPARSE SpGEMM failed")
    n = nnz.value
    rpt = bm.DeviceArray(r.value, dA.M + 1, np.int32).to_host()
    col = bm.DeviceArray(c.value, n, np.int32).to_host()
    val = bm.DeviceArray(v.value, n, np.float64).to_host()
    for p in (r, c, v):
        L.brm_xcheck_free(p.value)
    return bm.CSR(rpt, col, val, dA.M, dB.N), tot.value, comp.value
