"""Vendor cross-check and the device transpose.

rocSPARSE computes the same products on the GPU independently (the
analogue of the reference's MKL check, mkl_spgemm.h:8-67): the structure
must match exactly, the values within 1e-12 relative (rocSPARSE sums in its
own order).  The device CSR transpose must equal the host transpose."""
import numpy as np
import pytest

from tests import _rocsparse
from tests._util import assert_csr_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["er", "rmat12", "rmat14", "banded", "amg"])
def test_rocsparse_agrees(gpu, name):
    bm = gpu
    if name == "er":
        A = B = bm.gen.er(16384, 16384, 8, 0)
    elif name.startswith("rmat"):
        A = B = bm.gen.rmat(int(name[4:]), seed=1)
    elif name == "banded":
        A, B = bm.gen.banded(6000, 64, 4, 32, 1), bm.gen.banded(6000, 64, 4, 32, 2)
    else:
        A, B = bm.gen.stencil27(24), bm.gen.aggregation(24)
    dA = bm.DeviceCSR.upload(A)
    dB = dA if B is A else bm.DeviceCSR.upload(B)
    ours = bm.spgemm_dev(dA, dB)
    C = ours.download()
    ours.free()
    V, _, _ = _rocsparse.spgemm(bm, dA, dB)
    assert_csr_equal(C, V, bitwise=False, rtol=1e-12)
    dA.free()
    if dB is not dA:
        dB.free()


@pytest.mark.parametrize("shape", [(1, 1), (5000, 300), (300, 5000), (20000, 20000)])
def test_device_transpose(gpu, shape):
    bm = gpu
    M, N = shape
    A = bm.gen.er_var(M, N, 0, min(N, 9), 3)
    T = A.transpose()
    dA = bm.DeviceCSR.upload(A)
    dT = dA.transpose()
    H = dT.download()
    assert H.M == N and H.N == M
    assert_csr_equal(H, T)
    dA.free()
    dT.free()


def test_device_transpose_aggregation(gpu):
    bm = gpu
    P = bm.gen.aggregation(30)
    dP = bm.DeviceCSR.upload(P)
    dPT = dP.transpose()
    assert_csr_equal(dPT.download(), P.transpose())
    dP.free()
    dPT.free()
