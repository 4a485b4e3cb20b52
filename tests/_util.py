"""Shared helpers for the parity tests."""
import numpy as np


def bits(x):
    return np.asarray(x, np.float64).view(np.uint64)


def as_tuple(A):
    """CSR -> (rpt int32, col int32, val f64, ncols) for the oracle."""
    return (np.ascontiguousarray(A.rpt, np.int32), A.col, A.val, A.N)


def assert_csr_equal(C, R, bitwise=True, rtol=1e-12):
    """C: product under test; R: oracle result (OracleCSR or CSR).
    Structure must match exactly.  Values: bit-identical when bitwise,
    otherwise max-abs-normalised relative error <= rtol."""
    np.testing.assert_array_equal(np.asarray(C.rpt, np.int64), np.asarray(R.rpt, np.int64))
    np.testing.assert_array_equal(np.asarray(C.col, np.int32), np.asarray(R.col, np.int32))
    cv, rv = np.asarray(C.val, np.float64), np.asarray(R.val, np.float64)
    if len(rv):
        scale = max(np.abs(rv).max(), 1e-300)
        assert np.abs(cv - rv).max() / scale <= rtol
    if bitwise:
        mism = np.nonzero(bits(cv) != bits(rv))[0]
        assert len(mism) == 0, f"{len(mism)} values differ in bits, first at {mism[:5]}"


def rand_csr_tuple(rng, M, K, per_row_max, empty_frac=0.2, adversarial=True):
    rpt, col, val = [0], [], []
    for _ in range(M):
        d = 0 if rng.random() < empty_frac else int(rng.integers(0, per_row_max + 1))
        cols = np.sort(rng.choice(K, size=min(d, K), replace=False))
        col.extend(cols.tolist())
        if adversarial:
            mags = rng.choice([1.0, 2.0 ** 30, 2.0 ** -30, 3.0, 2.0 ** 52], size=len(cols))
            vals = rng.uniform(-1, 1, len(cols)) * mags
        else:
            vals = rng.uniform(-1, 1, len(cols))
        val.extend(vals.tolist())
        rpt.append(len(col))
    return (np.array(rpt, np.int32), np.array(col, np.int32), np.array(val, np.float64), K)


def csr_from_tuple(T, ncols=None):
    from brmerge_amd import CSR
    n = T[3] if len(T) > 3 else ncols
    return CSR(T[0], T[1], T[2], len(T[0]) - 1, n)
