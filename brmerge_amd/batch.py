"""Batched SpGEMM: many independent products C_i = A_i * B_i in one pass.

The reference multiplies one pair per call (BRMerge-SpGEMM/inc/spgemm.h:468-471).
Here a batch runs as the block-diagonal product diag(A_i) * diag(B_i) on the
GPU; every C_i is bit-identical to a single call on (A_i, B_i).

    C, row_off = spgemm_batched([(dA0, dB0), (dA1, dB1), ...])
    plan = BatchPlan(pairs, autotune=True); plan.execute(); plan.product(i)
"""
import ctypes
import json

import numpy as np

from . import _lib
from .csr import CSR
from .spgemm import DeviceArray, DeviceCSR, Precise_Timing


def _pairs_array(pairs):
    arr = (_lib.PairC * max(len(pairs), 1))()
    for i, (dA, dB) in enumerate(pairs):
        if dA.N != dB.M:
            raise ValueError(f"pair {i}: A is {dA.M}x{dA.N}, B is {dB.M}x{dB.N}")
        arr[i] = _lib.PairC(dA.rpt.ptr, dA.col.ptr, dA.val.ptr, dA.nnz, dB.rpt.ptr, dB.col.ptr, dB.val.ptr,
                            dB.nnz, dA.M, dA.N, dB.N)
    return arr


def set_tune_cache(path):
    """Name the JSON file of tuned tier shapes that BatchPlan(autotune=True)
    reads (a hit skips the tuner) and records its choices in; None turns
    it off.  See brm_set_tune_cache in include/brmerge.h."""
    _lib.check(_lib.lib().brm_set_tune_cache(path.encode() if path else None))


def spgemm_batched(pairs, stream=None, timing=None):
    """C_i = A_i * B_i for a list of (DeviceCSR A_i, DeviceCSR B_i).

    Returns (DeviceCSR C, row_off): C holds all products' rows in order
    (int64 row pointer, each product's own columns); product i is rows
    row_off[i] .. row_off[i+1]-1."""
    L = _lib.lib()
    arr = _pairs_array(pairs)
    n = len(pairs)
    row_off = np.zeros(n + 1, np.int64)
    r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    nnz = ctypes.c_int64()
    tc = _lib.PreciseTimingC()
    _lib.check(L.brm_spgemm_batched_dev(ctypes.cast(arr, ctypes.c_void_p), n, ctypes.byref(r), ctypes.byref(c),
                                        ctypes.byref(v), ctypes.byref(nnz), row_off.ctypes.data,
                                        ctypes.cast(ctypes.pointer(tc), ctypes.c_void_p), stream))
    if timing is not None:
        timing._load(tc)
    M = int(row_off[-1])
    N = max([dB.N for _, dB in pairs], default=0)
    C = DeviceCSR(DeviceArray(r.value, M + 1, np.int64), DeviceArray(c.value, nnz.value, np.int32),
                  DeviceArray(v.value, nnz.value, np.float64), M, N, nnz.value)
    return C, row_off


def split_products(C, row_off, ncols):
    """Host copies of the products of a batched result: list of CSR."""
    rpt = C.rpt.to_host()
    out = []
    for i in range(len(row_off) - 1):
        r0, r1 = int(row_off[i]), int(row_off[i + 1])
        b, e = int(rpt[r0]), int(rpt[r1])
        out.append(CSR(rpt[r0:r1 + 1] - b, C.col.to_host(b, e), C.val.to_host(b, e), r1 - r0, ncols[i]))
    return out


class BatchPlan:
    """A batch over fixed sparsity patterns, executed by one HIP graph replay.

    The caller's operand arrays must stay alive (and at the same addresses)
    for the plan's lifetime; their values may change between executions."""

    def __init__(self, pairs, autotune=True, stream=None):
        self._keep = list(pairs)
        self._ncols = [dB.N for _, dB in pairs]
        arr = _pairs_array(pairs)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().brm_batch_plan_create(ctypes.cast(arr, ctypes.c_void_p), len(pairs),
                                                    1 if autotune else 0, ctypes.byref(h), stream))
        self._h = h.value

    def execute(self, stream=None):
        _lib.check(_lib.lib().brm_batch_plan_execute(self._h, stream))

    def _result(self):
        r, c, v, ro = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        nnz = ctypes.c_int64()
        _lib.check(_lib.lib().brm_batch_plan_result(self._h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(v),
                                                    ctypes.byref(nnz), ctypes.byref(ro)))
        n = len(self._ncols)
        row_off = np.ctypeslib.as_array(ctypes.cast(ro.value, ctypes.POINTER(ctypes.c_int64)), (n + 1,)).copy()
        M = int(row_off[-1])
        C = DeviceCSR(DeviceArray(r.value, M + 1, np.int64), DeviceArray(c.value, nnz.value, np.int32),
                      DeviceArray(v.value, nnz.value, np.float64), M, max(self._ncols, default=0), nnz.value)
        return C, row_off

    def result(self):
        """(DeviceCSR C, row_off) -- views owned by the plan (do not free)."""
        return self._result()

    def products(self):
        """Host copies of every product (list of CSR)."""
        C, row_off = self._result()
        return split_products(C, row_off, self._ncols)

    def info(self):
        buf = ctypes.create_string_buffer(1 << 14)
        _lib.check(_lib.lib().brm_batch_plan_info(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def close(self):
        if self._h:
            _lib.lib().brm_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["spgemm_batched", "split_products", "BatchPlan", "Precise_Timing"]
