"""SpGEMM entry points mirroring the reference's interface.

    brmerge_precise(A, B, C, timing)   ~ spgemm.h:737-741 (CSR overload)
    brmerge_upper(A, B, C, timing)     ~ spgemm.h:460-464
    compute_flop(A, B)                 ~ spgemm.h:59-66

All of them call libbrmerge.so (HIP kernels on the GPU); there is no CPU
path.  ``DeviceCSR`` and ``spgemm_dev`` keep operands and results in HBM.
"""
import ctypes

import numpy as np

from . import _lib
from .csr import CSR, Precise_Timing, Upper_Timing


def _ptr(a):
    return a.ctypes.data if a is not None and len(a) else None


def _host_call(A, B, upper, i64, timing):
    L = _lib.lib()
    arpt = np.ascontiguousarray(A.rpt, np.int32)
    brpt = np.ascontiguousarray(B.rpt, np.int32)
    acol, aval = np.ascontiguousarray(A.col, np.int32), np.ascontiguousarray(A.val, np.float64)
    bcol, bval = np.ascontiguousarray(B.col, np.int32), np.ascontiguousarray(B.val, np.float64)
    r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    tc = _lib.UpperTimingC() if upper else _lib.PreciseTimingC()
    nnz = ctypes.c_int64() if i64 else ctypes.c_int()
    fn = {(False, False): L.brm_spgemm_precise_i32, (True, False): L.brm_spgemm_upper_i32,
          (False, True): L.brm_spgemm_precise_i64, (True, True): L.brm_spgemm_upper_i64}[(upper, i64)]
    st = fn(_ptr(arpt), _ptr(acol), _ptr(aval), _ptr(brpt), _ptr(bcol), _ptr(bval),
            ctypes.byref(r), ctypes.byref(c), ctypes.byref(v), A.M, A.N, B.N, ctypes.byref(nnz),
            ctypes.cast(ctypes.pointer(tc), ctypes.c_void_p))
    _lib.check(st)
    C = CSR.from_owned(r, c, v, A.M, B.N, rpt_i64=i64)
    if timing is not None:
        timing._load(tc)
    return C


def _assign(C, R):
    C.M, C.N, C.rpt, C.col, C.val = R.M, R.N, R.rpt, R.col, R.val


def brmerge_precise(A, B, C=None, timing=None, i64=False):
    """C = A*B with BRMerge-Precise (exact symbolic count, then numeric merge).

    As the reference's CSR overload, C (if given) receives M = A.M, N = B.N
    and the arrays.  Returns C.  Raises BrmError (BRM_E_OVERFLOW when nnz(C)
    exceeds the int32 row pointer; pass i64=True for an int64 row pointer)."""
    R = _host_call(A, B, False, i64, timing)
    if C is None:
        return R
    _assign(C, R)
    return C


def brmerge_upper(A, B, C=None, timing=None, i64=False):
    """C = A*B with BRMerge-Upper (merge into a flop-sized buffer, then compact)."""
    R = _host_call(A, B, True, i64, timing)
    if C is None:
        return R
    _assign(C, R)
    return C


def compute_flop(A, B):
    """Number of intermediate products of A*B (the reference's 'flop')."""
    L = _lib.lib()
    arpt = np.ascontiguousarray(A.rpt, np.int32)
    brpt = np.ascontiguousarray(B.rpt, np.int32)
    acol = np.ascontiguousarray(A.col, np.int32)
    tot = ctypes.c_int64()
    _lib.check(L.brm_flop_count(_ptr(arpt), _ptr(acol), _ptr(brpt), A.M, B.M, ctypes.byref(tot)))
    return tot.value


# ------------------------------------------------------------------ device side
class DeviceArray:
    """A buffer from the library's caching device allocator."""

    def __init__(self, ptr, n, dtype):
        self.ptr = ptr
        self.n = int(n)
        self.dtype = np.dtype(dtype)

    @classmethod
    def alloc(cls, n, dtype):
        p = _lib.lib().brm_dev_malloc(max(int(n), 1) * np.dtype(dtype).itemsize)
        if not p:
            raise _lib.BrmError(_lib.BRM_E_NOMEM, "brm_dev_malloc failed")
        return cls(p, n, dtype)

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls.alloc(len(arr), arr.dtype)
        if len(arr):
            _lib.check(_lib.lib().brm_memcpy_h2d(d.ptr, arr.ctypes.data, arr.nbytes))
        return d

    def to_host(self, start=0, stop=None):
        stop = self.n if stop is None else stop
        out = np.empty(max(stop - start, 0), self.dtype)
        if len(out):
            _lib.check(_lib.lib().brm_memcpy_d2h(out.ctypes.data, self.ptr + start * self.dtype.itemsize,
                                                 out.nbytes))
        return out

    def free(self):
        if self.ptr:
            _lib.lib().brm_dev_free(self.ptr)
            self.ptr = None

    def checksum(self, stream=None):
        """Order-independent 64-bit checksum of the array's bytes (device)."""
        out = ctypes.c_uint64()
        _lib.check(_lib.lib().brm_checksum_dev(self.ptr, self.n * self.dtype.itemsize, ctypes.byref(out), stream))
        return out.value


class DeviceCSR:
    """CSR resident in HBM (rpt int32 for operands, int64 for results)."""

    def __init__(self, rpt, col, val, M, N, nnz):
        self.rpt, self.col, self.val = rpt, col, val
        self.M, self.N, self.nnz = int(M), int(N), int(nnz)

    @classmethod
    def upload(cls, A):
        return cls(DeviceArray.from_host(np.ascontiguousarray(A.rpt, np.int32)),
                   DeviceArray.from_host(A.col), DeviceArray.from_host(A.val), A.M, A.N, A.nnz)

    def download(self):
        return CSR(self.rpt.to_host(), self.col.to_host(), self.val.to_host(), self.M, self.N)

    def download_panel(self, rows):
        """Host CSR of the selected rows (ascending row ids), as an
        (rpt int64, col, val) triple."""
        rows = np.asarray(rows, np.int64)
        rpt = self.rpt.to_host()
        b, e = rpt[rows], rpt[rows + 1]
        cols = [self.col.to_host(int(x), int(y)) for x, y in zip(b, e)]
        vals = [self.val.to_host(int(x), int(y)) for x, y in zip(b, e)]
        prpt = np.concatenate([[0], np.cumsum(e - b)]).astype(np.int64)
        return (prpt, np.concatenate(cols) if cols else np.zeros(0, np.int32),
                np.concatenate(vals) if vals else np.zeros(0))

    def download_rows(self, rows):
        """Host copy of selected rows: list of (cols, vals)."""
        rpt = self.rpt.to_host()
        out = []
        for r in rows:
            b, e = int(rpt[r]), int(rpt[r + 1])
            out.append((self.col.to_host(b, e), self.val.to_host(b, e)))
        return out

    def free(self):
        for a in (self.rpt, self.col, self.val):
            a.free()

    def transpose(self, stream=None):
        """A^T as a new DeviceCSR (int32 row pointer), computed on the GPU."""
        src = self.as_operand(stream)
        t = DeviceCSR(DeviceArray.alloc(self.N + 1, np.int32), DeviceArray.alloc(self.nnz, np.int32),
                      DeviceArray.alloc(self.nnz, np.float64), self.N, self.M, self.nnz)
        _lib.check(_lib.lib().brm_csr_transpose_dev(src.rpt.ptr, src.col.ptr, src.val.ptr, self.nnz, self.M, self.N,
                                                    t.rpt.ptr, t.col.ptr, t.val.ptr, stream))
        if src is not self:
            src.rpt.free()
        return t

    def as_operand(self, stream=None):
        """This CSR with an int32 row pointer (a new device array; the
        column and value arrays are shared), usable as an operand."""
        if self.rpt.dtype == np.int32:
            return self
        r32 = DeviceArray.alloc(self.M + 1, np.int32)
        _lib.check(_lib.lib().brm_rpt_to_i32(self.rpt.ptr, self.M + 1, r32.ptr, stream))
        return DeviceCSR(r32, self.col, self.val, self.M, self.N, self.nnz)


def spgemm_dev(dA, dB, upper=False, stream=None, timing=None):
    """C = A*B for operands already in HBM; returns a DeviceCSR (int64 rpt)."""
    L = _lib.lib()
    r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nnz = ctypes.c_int64()
    # phase timing (hipEvents around each phase) only when asked for
    tc = (_lib.UpperTimingC() if upper else _lib.PreciseTimingC()) if timing is not None else None
    fn = L.brm_spgemm_upper_dev if upper else L.brm_spgemm_precise_dev
    st = fn(dA.rpt.ptr, dA.col.ptr, dA.val.ptr, dA.nnz, dB.rpt.ptr, dB.col.ptr, dB.val.ptr, dB.nnz,
            dA.M, dA.N, dB.N, ctypes.byref(r), ctypes.byref(c), ctypes.byref(v), ctypes.byref(nnz),
            ctypes.cast(ctypes.pointer(tc), ctypes.c_void_p) if tc is not None else None, stream)
    _lib.check(st)
    if timing is not None:
        timing._load(tc)
    n = nnz.value
    return DeviceCSR(DeviceArray(r.value, dA.M + 1, np.int64), DeviceArray(c.value, n, np.int32),
                     DeviceArray(v.value, n, np.float64), dA.M, dB.N, n)


def set_param(name, value):
    _lib.check(_lib.lib().brm_set_param(name.encode(), int(value)))


def get_param(name):
    return int(_lib.lib().brm_get_param(name.encode()))


def last_stats():
    """Work counters of this thread's last SpGEMM call (dict)."""
    st = _lib.StatsC()
    _lib.check(_lib.lib().brm_last_stats(ctypes.cast(ctypes.pointer(st), ctypes.c_void_p)))
    d = {f: getattr(st, f) for f, _ in _lib.StatsC._fields_ if f != "rows_by_tier"}
    d["rows_by_tier"] = list(st.rows_by_tier)
    return d


def prof_reset():
    _lib.lib().brm_prof_reset()


def prof_report():
    import json
    buf = ctypes.create_string_buffer(1 << 16)
    _lib.check(_lib.lib().brm_prof_report(buf, len(buf)))
    return {k: (int(v[0]), float(v[1])) for k, v in json.loads(buf.value.decode()).items()}


__all__ = ["brmerge_precise", "brmerge_upper", "compute_flop", "DeviceArray", "DeviceCSR", "spgemm_dev",
           "Precise_Timing", "Upper_Timing", "set_param", "get_param", "prof_reset", "prof_report",
           "last_stats"]
