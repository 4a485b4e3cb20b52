"""Hash-SpGEMM tier (the paper's comparison baseline,
Existing-SpGEMM/inc/hash_mult.h:558-608), selected by the hash_cap knob:
rows with at most hash_cap products are merged by an LDS hash table in the
reference Hash SpGEMM's accumulation order.  Its C must equal the oracle's
hash restatement bit for bit (oracle.hash_spgemm with the same cap), keep
Precise's structure exactly, and stay within 1e-12 of Precise's values."""
import numpy as np
import pytest

import oracle
from tests._util import as_tuple, assert_csr_equal, csr_from_tuple, rand_csr_tuple

pytestmark = pytest.mark.gpu

CAP = 1024    # the knob's value here
LMAX = 1024   # the tier's list limit (rows with more A entries keep BRMerge's tree)


@pytest.fixture
def hash_on(gpu):
    old = gpu.get_param("hash_cap")
    gpu.set_param("hash_cap", CAP)
    yield gpu
    gpu.set_param("hash_cap", old)


def check(gpu, A, B, nthreads=0):
    """A, B: oracle tuples.  Precise and Upper in hash mode against the
    hash restatement; values also against BRMerge's."""
    H = oracle.hash_spgemm(A, B, cap=CAP, lmax=LMAX, nthreads=nthreads)
    P = oracle.brmerge_precise(A, B, nthreads=nthreads)
    Ah, Bh = csr_from_tuple(A, len(B[0]) - 1), csr_from_tuple(B)
    for f in (gpu.brmerge_precise, gpu.brmerge_upper):
        C = f(Ah, Bh)
        assert_csr_equal(C, H)
        assert_csr_equal(C, P, bitwise=False, rtol=1e-12)


@pytest.mark.parametrize("seed", range(8))
def test_hash_random(hash_on, seed):
    rng = np.random.default_rng(900 + seed)
    M, K, N = [int(x) for x in rng.integers(1, 400, 3)]
    A = rand_csr_tuple(rng, M, K, int(rng.integers(1, 90)), empty_frac=0.2, adversarial=False)
    B = rand_csr_tuple(rng, K, N, int(rng.integers(1, 40)), empty_frac=0.2, adversarial=False)
    check(hash_on, A, B, nthreads=4)


def test_hash_cancellation_orders(hash_on):
    """One column summed over four lists where list order, tree order and a
    right fold give 0, 1 and 2 (2^53 + 1 rounds to 2^53)."""
    b = 2.0 ** 53
    L = 4
    A = (np.array([0, L], np.int32), np.arange(L, dtype=np.int32), np.ones(L), L)
    B = (np.arange(L + 1, dtype=np.int32), np.zeros(L, np.int32), np.array([b, 1.0, 1.0, -b]), 1)
    H = oracle.hash_spgemm(A, B)
    assert H.val.tolist() == [0.0]
    C = hash_on.brmerge_precise(csr_from_tuple(A), csr_from_tuple(B))
    assert C.val.tolist() == [0.0]


@pytest.mark.parametrize("name", ["er", "banded", "rmat12", "amg15", "hubs"])
def test_hash_workloads(hash_on, name):
    """Configs' shapes: ER (config 1), banded and AMG (config 3; every row
    in the tier), R-MAT scale 12 (rows on both sides of the cap: the long
    ones keep BRMerge's tree), and rows of up to 1,024 single-entry lists
    into a few hub columns (deep sequential folds)."""
    gpu = hash_on
    if name == "er":
        A = B = as_tuple(gpu.gen.er(8192, 8192, 8, 3))
    elif name == "banded":
        A, B = as_tuple(gpu.gen.banded(6000, 64, 4, 32, 1)), as_tuple(gpu.gen.banded(6000, 64, 4, 32, 2))
    elif name == "rmat12":
        A = B = as_tuple(gpu.gen.rmat(12, seed=1))
    elif name == "amg15":
        A, B = as_tuple(gpu.gen.stencil27(15)), as_tuple(gpu.gen.aggregation(15))
    else:
        rng = np.random.default_rng(77)
        K, N = 3000, 50
        bc = rng.integers(0, 4, K).astype(np.int32)  # one entry per B row, 4 hub columns
        B = (np.arange(K + 1, dtype=np.int32), bc, rng.uniform(-1, 1, K) * rng.choice([1.0, 2.0 ** 40], K), N)
        # a fifth of the B rows empty: rows of up to 1,300 lists with fewer than
        # 1,024 products (more lists than the tier holds: BRMerge's tree)
        bc_rows = [[c] if r % 5 else [] for r, c in enumerate(bc)]
        brpt = np.concatenate([[0], np.cumsum([len(x) for x in bc_rows])]).astype(np.int32)
        bcol = np.array([c for x in bc_rows for c in x], np.int32)
        B = (brpt, bcol, rng.uniform(-1, 1, len(bcol)) * rng.choice([1.0, 2.0 ** 40], len(bcol)), N)
        acols = [np.sort(rng.choice(K, size=s, replace=False)) for s in (1024, 1025, 1300, 700, 64, 1)]
        arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
        A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
    check(gpu, A, B, nthreads=8)
