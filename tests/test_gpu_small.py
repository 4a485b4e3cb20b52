"""GPU parity of the fused short-row path (small.hip) against the oracle.

The fused path takes products of at most 2^20 rows whose average row has at
most 64 products.  It runs the rows in groups of 32 lanes when the average is
at most 32, retries in groups of 64 when a row is longer than 32, and hands
the product to the general pipeline when a row is longer than 64.  Each case
below is built so that it ends in one of these three, and each is compared
with the oracle and with the general pipeline (small_fused = 0) bit for bit.
"""
import numpy as np
import pytest

import oracle
from tests._util import assert_csr_equal, csr_from_tuple, rand_csr_tuple

pytestmark = pytest.mark.gpu


def _rows_with(rng, K, N, lens, kuse=None):
    """A with the given row lengths over K columns (the first `kuse` of
    them), and B whose rows all have 1 entry: each A row's product count
    equals its length."""
    kuse = K if kuse is None else kuse
    rpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(kuse, size=n, replace=False)) for n in lens]).astype(np.int32)
    A = (rpt, col, rng.uniform(-1, 1, len(col)), K)
    bcol = rng.integers(0, N, size=K).astype(np.int32)
    B = (np.arange(K + 1, dtype=np.int32), bcol, rng.uniform(-1, 1, K), N)
    return A, B


def _check(bm, A, B):
    Ah, Bh = csr_from_tuple(A), csr_from_tuple(B)
    R = oracle.brmerge_precise(A, B, nthreads=4)
    C = bm.brmerge_precise(Ah, Bh)
    st = bm.last_stats()
    assert_csr_equal(C, R)
    ip = int(np.diff(B[0])[A[1]].sum())
    assert st["total_ip"] == ip
    assert st["nnz_c"] == R.nnz
    old = bm.get_param("small_fused")
    bm.set_param("small_fused", 0)
    try:
        assert_csr_equal(bm.brmerge_precise(Ah, Bh), R)
    finally:
        bm.set_param("small_fused", old)
    return st


@pytest.mark.parametrize("longest", [32, 33, 64, 65, 300])
def test_group_choice(gpu, longest):
    """Rows of 4-20 products (groups of 32 lanes) and one row of `longest`:
    32 stays in groups of 32, 33 and 64 retry in groups of 64, 65 and 300
    take the general pipeline."""
    rng = np.random.default_rng(longest)
    M, K, N = 3000, 4000, 500
    lens = rng.integers(4, 21, size=M)
    lens[1234] = longest
    A, B = _rows_with(rng, K, N, lens)
    _check(gpu, A, B)


@pytest.mark.parametrize("seed", range(6))
def test_random_short_rows(gpu, seed):
    """Random rows (adversarial magnitudes, empty rows) whose products all
    fit one wavefront, from 1 row to 70,000 rows."""
    rng = np.random.default_rng(500 + seed)
    M = [1, 63, 64, 65, 5000, 70000][seed]
    K, N = int(rng.integers(1, 3000)), int(rng.integers(1, 3000))
    A = rand_csr_tuple(rng, M, K, 8, empty_frac=0.3)
    B = rand_csr_tuple(rng, K, N, 8, empty_frac=0.3)
    _check(gpu, A, B)


def test_million_rows(gpu):
    """2^20 rows (the largest fused product) of 27-point stencil times
    aggregation shape: groups of 32 lanes, workgroup look-back over 2^17
    workgroups."""
    rng = np.random.default_rng(7)
    M, K, N = 1 << 20, 1 << 20, 40000
    lens = np.full(M, 27)
    rpt = np.arange(M + 1, dtype=np.int64) * 27
    base = np.clip(np.arange(M) - 13, 0, K - 27)
    col = (base[:, None] + np.arange(27)[None, :]).reshape(-1).astype(np.int32)
    A = (rpt.astype(np.int32), col, rng.uniform(-1, 1, len(col)), K)
    B = (np.arange(K + 1, dtype=np.int32), (np.arange(K) // 27 % N).astype(np.int32), rng.uniform(-1, 1, K), N)
    st = _check(gpu, A, B)
    assert st["total_ip"] == int(lens.sum())


def test_fused_invalid_b_row(gpu):
    """An unsorted B row that no A row uses is still refused (the fused
    kernel checks all of B), and the library stays usable."""
    bm = gpu
    rng = np.random.default_rng(3)
    A, B = _rows_with(rng, 100, 50, rng.integers(1, 10, size=200), kuse=99)
    spare = 99
    # B row `spare` gets two entries in the wrong order
    brpt = np.concatenate([[0], np.cumsum([2 if k == spare else 1 for k in range(100)])]).astype(np.int32)
    bcol = np.concatenate([[7, 3] if k == spare else [B[1][k]] for k in range(100)]).astype(np.int32)
    Bbad = (brpt, bcol, np.ones(len(bcol)), 50)
    with pytest.raises(bm.BrmError) as e:
        bm.brmerge_precise(csr_from_tuple(A), csr_from_tuple(Bbad))
    assert e.value.status == 1
    _check(bm, A, B)


@pytest.mark.parametrize("used", [False, True], ids=["unused", "used"])
def test_long_b_row(gpu, used):
    """A B row of 100 entries (longer than the fused kernel checks) sends
    the product to the general pipeline, whether an A row uses it or not."""
    rng = np.random.default_rng(11)
    K, N = 500, 800
    lens = np.ones(K, np.int64)
    lens[K - 1] = 100
    brpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    bcol = np.concatenate([np.sort(rng.choice(N, size=n, replace=False)) for n in lens]).astype(np.int32)
    B = (brpt, bcol, rng.uniform(-1, 1, len(bcol)), N)
    A, _ = _rows_with(rng, K, N, rng.integers(1, 8, size=2000), kuse=K - 1)
    if used:  # the last entry of A's row 5 becomes column K - 1 (no other row uses it)
        rpt, col = A[0], A[1]
        col[rpt[6] - 1] = K - 1
    _check(gpu, A, B)


@pytest.mark.parametrize("N", [(1 << 26) - 1, 1 << 26, (1 << 27) - 1, 1 << 27, (1 << 31) - 1])
@pytest.mark.parametrize("shape", ["g32", "g64"])
def test_wide_columns(gpu, N, shape):
    """Columns up to N - 1 around the widths where the register sort
    switches from 32-bit (column, lane) keys to 64-bit (column, position)
    keys: 2^27 for 32-lane groups, 2^26 for 64-lane groups.  Both the fused
    path and the general pipeline's short tier are checked."""
    rng = np.random.default_rng(N % 1000 + (shape == "g64"))
    K = 3000
    if shape == "g32":
        lens_a, per_b = rng.integers(4, 21, size=2000), 1
    else:
        lens_a, per_b = np.full(2000, 8), 8
    rpt = np.concatenate([[0], np.cumsum(lens_a)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(K, size=n, replace=False)) for n in lens_a]).astype(np.int32)
    A = (rpt, col, rng.uniform(-1, 1, len(col)), K)
    # B columns from a small pool near N - 1 (so that columns repeat across
    # lists and merge), N - 1 itself included
    pool = np.unique(np.concatenate([[N - 1], N - 1 - rng.integers(0, 200, size=60)]))
    brows = [np.sort(rng.choice(pool, size=per_b, replace=False)) for _ in range(K)]
    brpt = np.concatenate([[0], np.cumsum([len(r) for r in brows])]).astype(np.int32)
    B = (brpt, np.concatenate(brows).astype(np.int32), rng.uniform(-1, 1, K * per_b), N)
    _check(gpu, A, B)
