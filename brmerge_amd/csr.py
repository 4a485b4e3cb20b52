"""Host-side CSR container and timing records, shaped like the reference's.

``CSR`` mirrors ``class CSR`` (BRMerge-SpGEMM/inc/CSR.h:7-28): M, N, nnz and
the rpt/col/val arrays (int32 / int32 / float64, numpy).  ``Precise_Timing``
and ``Upper_Timing`` mirror BRMerge-SpGEMM/inc/Timings.h:4-46 and print the
same tables as src/Timings.cpp.
"""
import ctypes

import numpy as np

from . import _lib


def _take(ptr, n, ctype, dtype, free):
    """Copy n elements from a library-allocated (new[]) array and free it."""
    if n > 0:
        arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (n,)).astype(dtype, copy=True)
    else:
        arr = np.zeros(0, dtype)
    if ptr:
        free(ptr)
    return arr


class CSR:
    """CSR matrix with 0-based int32 indices and float64 values."""

    def __init__(self, rpt=None, col=None, val=None, M=None, N=None):
        if rpt is None:
            self.M, self.N = 0, 0
            self.rpt = np.zeros(1, np.int32)
            self.col = np.zeros(0, np.int32)
            self.val = np.zeros(0, np.float64)
            return
        self.rpt = np.ascontiguousarray(rpt, dtype=np.int64 if np.asarray(rpt).dtype == np.int64 else np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.M = int(M if M is not None else len(self.rpt) - 1)
        self.N = int(N if N is not None else (int(self.col.max()) + 1 if len(self.col) else 0))

    @property
    def nnz(self):
        return int(self.rpt[self.M]) if len(self.rpt) else 0

    # CSR::construct (src/CSR.cpp:259-479)
    @classmethod
    def construct(cls, path):
        L = _lib.lib()
        M, N, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(L.brm_read_mtx(str(path).encode(), ctypes.byref(M), ctypes.byref(N), ctypes.byref(nnz),
                                  ctypes.byref(r), ctypes.byref(c), ctypes.byref(v)))
        return cls.from_owned(r, c, v, M.value, N.value)

    @classmethod
    def from_owned(cls, r, c, v, M, N, rpt_i64=False):
        """Wrap arrays allocated with new[] by the library (copies, then frees)."""
        L = _lib.lib()
        if rpt_i64:
            rpt = _take(r, M + 1, ctypes.c_int64, np.int64, L.brm_host_free_i64)
        else:
            rpt = _take(r, M + 1, ctypes.c_int32, np.int32, L.brm_host_free_i32)
        nnz = int(rpt[M])
        col = _take(c, nnz, ctypes.c_int32, np.int32, L.brm_host_free_i32)
        val = _take(v, nnz, ctypes.c_double, np.float64, L.brm_host_free_f64)
        return cls(rpt, col, val, M, N)

    def write_mtx(self, path):
        rpt = np.ascontiguousarray(self.rpt, np.int32)
        _lib.check(_lib.lib().brm_write_mtx(str(path).encode(), self.M, self.N, rpt.ctypes.data,
                                            self.col.ctypes.data, self.val.ctypes.data))

    def release(self):
        self.__init__()

    def transpose(self):
        L = _lib.lib()
        r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        rpt = np.ascontiguousarray(self.rpt, np.int32)
        _lib.check(L.brm_csr_transpose(self.M, self.N, rpt.ctypes.data, self.col.ctypes.data,
                                       self.val.ctypes.data, ctypes.byref(r), ctypes.byref(c), ctypes.byref(v)))
        return CSR.from_owned(r, c, v, self.N, self.M)

    def astuple(self):
        return (self.rpt, self.col, self.val, self.N)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rpt.astype(np.int64)), shape=(self.M, self.N))

    def row(self, i):
        b, e = int(self.rpt[i]), int(self.rpt[i + 1])
        return self.col[b:e], self.val[b:e]

    # CSR::operator== (src/CSR.cpp:214-253): structure exact, values within
    # 1e-9 absolute or relative; raises when nnz differs.
    def __eq__(self, other):
        if self.nnz != other.nnz:
            raise RuntimeError(f"nnz not equal {self.nnz} {other.nnz}")
        if self.M != other.M or self.N != other.N:
            return False
        if not np.array_equal(np.asarray(self.rpt, np.int64), np.asarray(other.rpt, np.int64)):
            return False
        if not np.array_equal(self.col, other.col):
            return False
        d = np.abs(self.val - other.val)
        ok = (d < 1e-9) | (d < 1e-9 * np.abs(self.val))
        return bool(ok.all())

    __hash__ = None

    def __repr__(self):
        return f"CSR(M={self.M}, N={self.N}, nnz={self.nnz})"


class _Timing:
    FIELDS = ()
    SUM_FIELDS = ()

    def __init__(self):
        self.measure_separate = True
        self.measure_whole = True
        for f in self.FIELDS:
            setattr(self, f, 0.0)
        self.h2d = 0.0
        self.d2h = 0.0

    def _load(self, c):
        for f in self.FIELDS + ("h2d", "d2h"):
            setattr(self, f, getattr(c, f))

    def __iadd__(self, o):
        for f in self.FIELDS + ("h2d", "d2h"):
            setattr(self, f, getattr(self, f) + getattr(o, f))
        return self

    def __itruediv__(self, n):
        for f in self.FIELDS + ("h2d", "d2h"):
            setattr(self, f, getattr(self, f) / n)
        return self

    def as_dict(self):
        return {f: getattr(self, f) for f in self.FIELDS + ("h2d", "d2h")}

    def print(self, total_flop):
        g = total_flop / 1e9
        print(f"total flop {total_flop:f}")
        s = sum(getattr(self, f) for f in self.SUM_FIELDS)
        tot = self.total if self.total else float("nan")
        if self.measure_separate:
            print("time(ms):")
            for f in self.FIELDS:
                if f == "total":
                    continue
                v = getattr(self, f)
                print(f"    {f:<20}{1000 * v:8.3f}ms {v / tot * 100:6.2f}%")
            print(f"    {'sum_total':<20}{1000 * s:8.3f}ms {s / tot * 100:6.2f}%")
            print("perf(GFLOPS):")
            for f in self.FIELDS:
                v = getattr(self, f)
                print(f"    {f:<20}{(g / v) if v else float('inf'):6.2f}")

    def reg_print(self, total_flop):
        print(f"{total_flop / 1e9 / self.total:f}")


class Precise_Timing(_Timing):  # noqa: N801  (reference name, Timings.h:27)
    FIELDS = ("pre_allocate", "compute_flop", "prefix_sum_flop", "load_balance", "symbolic",
              "prefix_sum_nnz", "allocate_C", "compute", "cleanup", "total")
    SUM_FIELDS = FIELDS[:8]  # Timings.cpp:135 excludes cleanup


class Upper_Timing(_Timing):  # noqa: N801  (reference name, Timings.h:4)
    FIELDS = ("thread", "pre_allocate", "compute_flop", "prefix_sum_flop", "load_balance", "allocate_hook",
              "compute", "prefix_sum_nnz", "allocate_C", "copy", "cleanup", "total")
    SUM_FIELDS = FIELDS[:11]
