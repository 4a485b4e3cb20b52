"""Request server: resident B operands, A*B requests batched as they arrive.

The reference has no serving path; this wraps the brm_server_* C ABI
(include/brmerge.h).  Requests are admitted against an HBM budget by their
BRMerge-Upper bound, fused with whatever else is queued, and run on worker
streams that overlap on the GPU.  Every result equals a single call on the
same operands, bit for bit.

    srv = Server(budget_bytes=64 << 30)
    b = srv.add_b(DeviceCSR.upload(B))
    rid = srv.submit(b, DeviceCSR.upload(A))
    C = srv.result(rid)          # host CSR; srv.release(rid) frees HBM
"""
import ctypes
import json

import numpy as np

from . import _lib
from .csr import CSR
from .spgemm import DeviceArray


class Server:
    def __init__(self, budget_bytes=64 << 30, max_batch=64, streams=2):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().brm_server_create(int(budget_bytes), int(max_batch), int(streams), ctypes.byref(h)))
        self._h = h.value
        self._keep = {}  # request id -> A (kept alive until the request is done)

    def add_b(self, dB):
        b = ctypes.c_int()
        _lib.check(_lib.lib().brm_server_add_b(self._h, dB.rpt.ptr, dB.col.ptr, dB.val.ptr, dB.nnz, dB.M, dB.N,
                                               ctypes.byref(b)))
        return b.value

    def submit(self, b_id, dA):
        rid = ctypes.c_int64()
        _lib.check(_lib.lib().brm_server_submit(self._h, int(b_id), dA.rpt.ptr, dA.col.ptr, dA.val.ptr, dA.nnz,
                                                dA.M, ctypes.byref(rid)))
        self._keep[rid.value] = dA
        return rid.value

    def poll(self, rid):
        d = ctypes.c_int()
        _lib.check(_lib.lib().brm_server_poll(self._h, int(rid), ctypes.byref(d)))
        return bool(d.value)

    def wait(self, rid):
        """Blocks until request rid is done; returns its ServerResultC
        (device views and timings).  Raises BrmError if its batch failed."""
        r = _lib.ServerResultC()
        st = _lib.lib().brm_server_wait(self._h, int(rid), ctypes.cast(ctypes.pointer(r), ctypes.c_void_p))
        self._keep.pop(int(rid), None)
        _lib.check(st)
        return r

    def result(self, rid):
        """Host CSR copy of request rid's product (waits for it)."""
        r = self.wait(rid)
        rpt = DeviceArray(r.crpt, r.M + 1, np.int64).to_host()
        col = DeviceArray(r.ccol, r.nnz, np.int32).to_host()
        val = DeviceArray(r.cval, r.nnz, np.float64).to_host()
        return CSR(rpt, col, val, r.M, r.N)

    def release(self, rid):
        _lib.check(_lib.lib().brm_server_release(self._h, int(rid)))

    def stats(self):
        buf = ctypes.create_string_buffer(4096)
        _lib.check(_lib.lib().brm_server_stats(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def close(self):
        if self._h:
            _lib.lib().brm_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Server"]
