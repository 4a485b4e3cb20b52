"""ctypes binding of liboracle.so (brm_oracle.c), the CPU restatement of
BRMerge-SpGEMM/inc/spgemm.h.  Test infrastructure only."""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def load():
    """Load (not build) liboracle.so; ``make -C oracle`` builds it."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
    lib = ctypes.CDLL(path)
    lib.brmo_upper_log.argtypes = [ctypes.c_int]
    lib.brmo_upper_log.restype = ctypes.c_int
    lib.brmo_compute_flop.argtypes = [_i32p, _i32p, _i32p, ctypes.c_int, _i32p, _i64p, ctypes.c_int]
    lib.brmo_compute_flop.restype = ctypes.c_int64
    lib.brmo_prefix_sum.argtypes = [_i64p, _i64p, ctypes.c_int, ctypes.c_int]
    lib.brmo_prefix_sum.restype = None
    lib.brmo_binary_search_approx.argtypes = [_i64p, ctypes.c_int, ctypes.c_int64]
    lib.brmo_binary_search_approx.restype = ctypes.c_int
    lib.brmo_max_threads.restype = ctypes.c_int
    lib.brmo_free.argtypes = [ctypes.c_void_p]
    for name in ("brmo_brmerge_precise", "brmo_brmerge_upper"):
        f = getattr(lib, name)
        f.argtypes = [_i32p, _i32p, _f64p, _i32p, _i32p, _f64p,
                      ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                      ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int64),
                      ctypes.c_int, _f64p]
        f.restype = ctypes.c_int
    lib.brmo_hash_numeric.argtypes = [_i32p, _i32p, _f64p, _i32p, _i32p, _f64p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
    lib.brmo_hash_numeric.restype = ctypes.c_int
    _LIB = lib
    return lib


@dataclass
class OracleCSR:
    M: int
    N: int
    rpt: np.ndarray  # int64 [M+1]
    col: np.ndarray  # int32 [nnz]
    val: np.ndarray  # float64 [nnz]

    @property
    def nnz(self):
        return int(self.rpt[-1])


def max_threads():
    return load().brmo_max_threads()


def upper_log(n):
    return load().brmo_upper_log(int(n))


def compute_flop(arpt, acol, brpt, nthreads=0):
    lib = load()
    M = len(arpt) - 1
    nnzrA = np.zeros(max(M, 1), np.int32)
    flop = np.zeros(max(M, 1), np.int64)
    tot = lib.brmo_compute_flop(np.ascontiguousarray(arpt, np.int32), np.ascontiguousarray(acol, np.int32),
                                np.ascontiguousarray(brpt, np.int32), M, nnzrA, flop, nthreads)
    return int(tot), nnzrA[:M], flop[:M]


def prefix_sum(x, nthreads=1):
    x = np.ascontiguousarray(x, np.int64)
    out = np.zeros(len(x) + 1, np.int64)
    load().brmo_prefix_sum(x, out, len(x), nthreads)
    return out


def binary_search_approx(arr, elem):
    arr = np.ascontiguousarray(arr, np.int64)
    return load().brmo_binary_search_approx(arr, len(arr), int(elem))


def _run(name, A, B, nthreads):
    lib = load()
    arpt, acol, aval = (np.ascontiguousarray(A[0], np.int32), np.ascontiguousarray(A[1], np.int32),
                        np.ascontiguousarray(A[2], np.float64))
    brpt, bcol, bval = (np.ascontiguousarray(B[0], np.int32), np.ascontiguousarray(B[1], np.int32),
                        np.ascontiguousarray(B[2], np.float64))
    M = len(arpt) - 1
    K = len(brpt) - 1
    N = int(B[3]) if len(B) > 3 else (int(bcol.max()) + 1 if len(bcol) else 0)
    # numpy refuses zero-length ndpointer arguments on some versions; pad.
    if len(acol) == 0:
        acol, aval = np.zeros(1, np.int32), np.zeros(1)
    if len(bcol) == 0:
        bcol, bval = np.zeros(1, np.int32), np.zeros(1)
    pr, pc, pv = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nnz = ctypes.c_int64()
    tm = np.zeros(16)
    rc = getattr(lib, name)(arpt, acol, aval, brpt, bcol, bval, M, K, N,
                            ctypes.byref(pr), ctypes.byref(pc), ctypes.byref(pv), ctypes.byref(nnz),
                            nthreads, tm)
    if rc != 0:
        raise RuntimeError(f"{name} failed with {rc}")
    n = nnz.value
    rpt = np.ctypeslib.as_array(ctypes.cast(pr, ctypes.POINTER(ctypes.c_int64)), (M + 1,)).copy()
    if n > 0:
        col = np.ctypeslib.as_array(ctypes.cast(pc, ctypes.POINTER(ctypes.c_int32)), (n,)).copy()
        val = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(ctypes.c_double)), (n,)).copy()
    else:
        col, val = np.zeros(0, np.int32), np.zeros(0)
    for p in (pr, pc, pv):
        lib.brmo_free(p)
    return OracleCSR(M, N, rpt, col, val), tm


def brmerge_precise(A, B, nthreads=0, timing=False):
    """C = A*B by the BRMerge-Precise restatement.  A and B are tuples
    (rpt, col, val[, ncols]).  Returns OracleCSR (and phase times)."""
    C, tm = _run("brmo_brmerge_precise", A, B, nthreads)
    return (C, tm) if timing else C


def hash_spgemm(A, B, cap=-1, lmax=-1, nthreads=0):
    """C = A*B with Hash SpGEMM's accumulation order (Existing-SpGEMM/inc/
    hash_mult.h:558-608: each column's contributions added in list order,
    sorted output) for rows with at most `cap` products and `lmax` A entries
    (no limit when negative); the other rows by BRMerge-Precise.  Returns
    OracleCSR."""
    lib = load()
    arpt, acol, aval = (np.ascontiguousarray(A[0], np.int32), np.ascontiguousarray(A[1], np.int32),
                        np.ascontiguousarray(A[2], np.float64))
    brpt, bcol, bval = (np.ascontiguousarray(B[0], np.int32), np.ascontiguousarray(B[1], np.int32),
                        np.ascontiguousarray(B[2], np.float64))
    M, K = len(arpt) - 1, len(brpt) - 1
    N = int(B[3]) if len(B) > 3 else (int(bcol.max()) + 1 if len(bcol) else 0)
    if len(acol) == 0:
        acol, aval = np.zeros(1, np.int32), np.zeros(1)
    if len(bcol) == 0:
        bcol, bval = np.zeros(1, np.int32), np.zeros(1)
    pr, pc, pv = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nnz = ctypes.c_int64()
    rc = lib.brmo_hash_numeric(arpt, acol, aval, brpt, bcol, bval, M, K, N, int(cap), int(lmax), ctypes.byref(pr),
                               ctypes.byref(pc), ctypes.byref(pv), ctypes.byref(nnz), nthreads)
    if rc != 0:
        raise RuntimeError(f"brmo_hash_numeric failed with {rc}")
    n = nnz.value
    rpt = np.ctypeslib.as_array(ctypes.cast(pr, ctypes.POINTER(ctypes.c_int64)), (M + 1,)).copy()
    if n > 0:
        col = np.ctypeslib.as_array(ctypes.cast(pc, ctypes.POINTER(ctypes.c_int32)), (n,)).copy()
        val = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(ctypes.c_double)), (n,)).copy()
    else:
        col, val = np.zeros(0, np.int32), np.zeros(0)
    for p in (pr, pc, pv):
        lib.brmo_free(p)
    return OracleCSR(M, N, rpt, col, val)


def brmerge_upper(A, B, nthreads=0, timing=False):
    C, tm = _run("brmo_brmerge_upper", A, B, nthreads)
    return (C, tm) if timing else C
