"""Deterministic synthetic inputs (libbrmerge.so host generators).

Every generator is counter-seeded (SplitMix64 -> xoshiro256**), so results do
not depend on thread count or machine.  The workload builders name the
sizes they produce.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .csr import CSR


def _gen(fn, M, N, *args):
    r, c, v = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(fn(*args, ctypes.byref(r), ctypes.byref(c), ctypes.byref(v)))
    return CSR.from_owned(r, c, v, M, N)


def er(M, N, d, seed=0):
    """M x N, d distinct uniform columns per row, values U[-1,1]."""
    return _gen(_lib.lib().brm_gen_er, M, N, M, N, d, seed)


def er_var(M, N, dmin, dmax, seed=0):
    return _gen(_lib.lib().brm_gen_er_var, M, N, M, N, dmin, dmax, seed)


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0):
    """R-MAT 2^scale square, duplicates removed, self-loops kept, values U[-1,1]."""
    n = 1 << scale
    return _gen(_lib.lib().brm_gen_rmat, n, n, scale, edge_factor, a, b, c, seed)


def stencil27(n):
    """27-point stencil on an n^3 grid: 26 on the diagonal, -1 off it."""
    return _gen(_lib.lib().brm_gen_stencil27, n ** 3, n ** 3, n)


def aggregation(n, g=3):
    """n^3 x ceil(n/g)^3 aggregation prolongator, one 1.0 per row."""
    m = (n + g - 1) // g
    return _gen(_lib.lib().brm_gen_aggregation, n ** 3, m ** 3, n, g)


def banded(M, bw, dmin, dmax, seed=0):
    return _gen(_lib.lib().brm_gen_banded, M, M, M, bw, dmin, dmax, seed)


def batched_pair(i, base_seed=0):
    """Pair i of the batched workload: square, rows log-uniform in
    [256, 65536], 4-32 nnz per row; even i Erdos-Renyi, odd i banded
    (half-width 64).  A uses seed 2i, B seed 2i+1 (offset by base_seed)."""
    x = (base_seed * 0x9E3779B97F4A7C15 + i * 0xD1B54A32D192ED03 + 0x2545F4914F6CDD1D) % (1 << 64)
    x ^= x >> 31
    x = (x * 0xBF58476D1CE4E5B9) % (1 << 64)
    x ^= x >> 29
    u = (x >> 11) * 2.0 ** -53
    M = int(round(math.exp(math.log(256) + u * (math.log(65536) - math.log(256)))))
    sa, sb = base_seed + 2 * i, base_seed + 2 * i + 1
    if i % 2 == 0:
        return er_var(M, M, 4, 32, sa), er_var(M, M, 4, 32, sb)
    return banded(M, 64, 4, 32, sa), banded(M, 64, 4, 32, sb)


def workload(name):
    """(A, B) operands of a named workload."""
    if name == "er16k":          # C = A*A, ER 16,384^2, 8 per row, seed 0
        A = er(16384, 16384, 8, 0)
        return A, A
    if name.startswith("rmat"):  # rmatS: R-MAT scale S, A*A
        A = rmat(int(name[4:]))
        return A, A
    if name.startswith("amg"):   # amgN: stencil27(N) * aggregation(N)
        n = int(name[3:])
        return stencil27(n), aggregation(n)
    raise ValueError(name)
