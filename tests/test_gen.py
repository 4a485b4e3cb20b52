"""Generators, MatrixMarket I/O and transpose (host code, CPU only)."""
import hashlib
import json
import os

import numpy as np
import pytest

import brmerge_amd as bm

GOLD = os.path.join(os.path.dirname(__file__), "golden", "generator_hashes.json")


def digest(A):
    h = hashlib.sha256()
    for a in (np.asarray(A.rpt, np.int64), A.col, A.val):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def check_csr(A):
    rpt = np.asarray(A.rpt, np.int64)
    assert rpt[0] == 0 and np.all(np.diff(rpt) >= 0) and rpt[-1] == len(A.col)
    for i in range(A.M):
        c = A.col[rpt[i]:rpt[i + 1]]
        assert np.all(np.diff(c) > 0)
    if len(A.col):
        assert A.col.min() >= 0 and A.col.max() < A.N


CASES = {
    "er_64_seed0": lambda: bm.gen.er(64, 64, 8, 0),
    "er_1024_seed1": lambda: bm.gen.er(1024, 1024, 8, 1),
    "rmat_10": lambda: bm.gen.rmat(10),
    "banded_500": lambda: bm.gen.banded(500, 64, 4, 32, 3),
    "stencil_5": lambda: bm.gen.stencil27(5),
    "agg_7": lambda: bm.gen.aggregation(7),
    "er_var_300": lambda: bm.gen.er_var(300, 400, 4, 32, 5),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_generators_valid_and_deterministic(name):
    A = CASES[name]()
    check_csr(A)
    assert digest(A) == digest(CASES[name]())
    with open(GOLD) as f:
        gold = json.load(f)
    assert gold[name] == digest(A), "generator output changed"


def test_er_shape_and_values():
    A = bm.gen.er(16384, 16384, 8, 0)
    assert A.nnz == 131072 and np.all(np.diff(A.rpt) == 8)
    assert A.val.min() >= -1 and A.val.max() < 1


def test_rmat_statistics_scale12():
    A = bm.gen.rmat(12)
    check_csr(A)
    deg = np.diff(A.rpt)
    # a = 0.57 concentrates edges at low indices: row 0 is among the densest
    assert deg[0] >= np.percentile(deg, 99)
    assert 0.25 < (deg == 0).mean() < 0.6
    assert A.nnz <= 16 * 4096


def test_stencil_counts():
    S = bm.gen.stencil27(96)
    assert S.M == 884736 and S.nnz == 23393656
    P = bm.gen.aggregation(96)
    assert P.N == 32768 and P.nnz == 884736
    assert np.all(P.val == 1.0)


def test_transpose_roundtrip():
    A = bm.gen.er(300, 200, 7, 4)
    T = A.transpose()
    check_csr(T)
    assert (T.M, T.N) == (200, 300)
    TT = T.transpose()
    assert np.array_equal(TT.rpt, A.rpt) and np.array_equal(TT.col, A.col) and np.array_equal(TT.val, A.val)
    np.testing.assert_array_equal(T.to_scipy().toarray(), A.to_scipy().toarray().T)


def test_mtx_roundtrip_and_reader_semantics(tmp_path):
    A = bm.gen.er(40, 50, 5, 2)
    p = tmp_path / "a.mtx"
    A.write_mtx(p)
    B = bm.CSR.construct(p)
    assert np.array_equal(B.rpt, A.rpt) and np.array_equal(B.col, A.col) and np.array_equal(B.val, A.val)
    # symmetric + pattern, entries out of order (CSR.cpp:358-438)
    q = tmp_path / "s.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n3 3 3\n3 1\n1 1\n2 1\n")
    S = bm.CSR.construct(q)
    assert S.rpt.tolist() == [0, 3, 4, 5] and S.col.tolist() == [0, 1, 2, 0, 0]
    assert np.all(S.val == 1.0)
    # skew-symmetric negates the mirrored entry; complex keeps the real part
    q.write_text("%%MatrixMarket matrix coordinate complex skew-symmetric\n2 2 1\n2 1 3.5 9.0\n")
    K = bm.CSR.construct(q)
    assert K.col.tolist() == [1, 0] and K.val.tolist() == [-3.5, 3.5]
    # duplicates are rejected like the reference (CSR.cpp:458-462)
    q.write_text("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n1 1 2.0\n")
    with pytest.raises(bm.BrmError):
        bm.CSR.construct(q)
    q.write_text("%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n1 1 1.0\n")
    with pytest.raises(bm.BrmError):
        bm.CSR.construct(q)
