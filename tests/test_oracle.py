"""Pins the CPU oracle (oracle/brm_oracle.c) -- CPU only.

The reference ships no tests or golden vectors and its sources cannot be
compiled here (TBB header absent), so the oracle is pinned by:
  * hand-derived known answers for the merge bracketing (tests/golden/),
    which also reject plausible wrong algorithms (left fold, right fold,
    dropped empty lists, fused multiply-add) -- see test_kat_rejects_mutants;
  * an independent brute force by the tree definition (tests/_brute.py);
  * scipy's all-ones pattern product (structure) and FP64 product (values);
  * closed forms: A = I, one entry per row, flop totals, scan identities.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from tests._brute import brute_spgemm

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "kat_bracketing.json")


def _unhex(vals):
    return np.array([float.fromhex(v) for v in vals], np.float64)


def load_kats():
    with open(GOLDEN) as f:
        data = json.load(f)
    out = []
    for c in data["cases"]:
        A = (np.array(c["A"]["rpt"], np.int32), np.array(c["A"]["col"], np.int32), _unhex(c["A"]["val"]))
        B = (np.array(c["B"]["rpt"], np.int32), np.array(c["B"]["col"], np.int32), _unhex(c["B"]["val"]),
             c["B"]["ncols"])
        C = (np.array(c["C"]["rpt"], np.int64), np.array(c["C"]["col"], np.int32), _unhex(c["C"]["val"]))
        out.append((c["name"], A, B, C))
    return out


KATS = load_kats()


def bits(x):
    return np.asarray(x, np.float64).view(np.uint64)


def assert_same(C, rpt, col, val):
    np.testing.assert_array_equal(np.asarray(C.rpt, np.int64), np.asarray(rpt, np.int64))
    np.testing.assert_array_equal(C.col, np.asarray(col, np.int32))
    np.testing.assert_array_equal(bits(C.val), bits(np.asarray(val, np.float64)))


@pytest.mark.parametrize("case", KATS, ids=[k[0] for k in KATS])
@pytest.mark.parametrize("variant", ["precise", "upper"])
def test_kat(case, variant):
    _, A, B, C = case
    f = oracle.brmerge_precise if variant == "precise" else oracle.brmerge_upper
    R = f(A, B, nthreads=2)
    assert_same(R, *C)


@pytest.mark.parametrize("case", KATS, ids=[k[0] for k in KATS])
def test_kat_brute(case):
    _, A, B, C = case
    rpt, col, val = brute_spgemm(A, B)
    np.testing.assert_array_equal(rpt, C[0])
    np.testing.assert_array_equal(col, C[1])
    np.testing.assert_array_equal(bits(val), bits(C[2]))


# ---- plausible wrong algorithms must disagree with at least one KAT -------
def _contribs(A, B, i, fma=False):
    arpt, acol, aval = A[0], A[1], A[2]
    brpt, bcol, bval = B[0], B[1], B[2]
    L = int(arpt[i + 1] - arpt[i])
    per = {}
    for p in range(L):
        a = int(arpt[i]) + p
        k = int(acol[a])
        for q in range(int(brpt[k]), int(brpt[k + 1])):
            per.setdefault(int(bcol[q]), []).append((p, float(aval[a]), float(bval[q])))
    return L, per


def _mutant(A, B, kind):
    rows = []
    for i in range(len(A[0]) - 1):
        L, per = _contribs(A, B, i)
        row = []
        for c in sorted(per):
            terms = per[c]
            if kind == "left_fold":
                v = terms[0][1] * terms[0][2]
                for t in terms[1:]:
                    v = v + t[1] * t[2]
            elif kind == "right_fold":
                v = terms[-1][1] * terms[-1][2]
                for t in reversed(terms[:-1]):
                    v = t[1] * t[2] + v
            elif kind == "drop_empty":
                # compact positions of non-empty lists, then pairwise tree
                nonempty = sorted({p for tt in per.values() for (p, _, _) in tt})
                remap = {p: k for k, p in enumerate(nonempty)}
                from tests._brute import _tree
                size = 1
                while size < len(nonempty):
                    size *= 2
                v = _tree([(remap[p], a * b) for (p, a, b) in terms], 0, size)
            elif kind in ("fma_left", "fma_right"):
                # pairwise tree, but one leaf product of a two-term sum fused
                # into the add (what -ffp-contract=fast can produce)
                from tests._brute import _tree
                size = 1
                while size < L:
                    size *= 2
                if len(terms) == 2:
                    (p0, a0, b0), (p1, a1, b1) = terms
                    if kind == "fma_left":
                        v = float(Fraction(a0) * Fraction(b0) + Fraction(a1 * b1))
                    else:
                        v = float(Fraction(a1) * Fraction(b1) + Fraction(a0 * b0))
                else:
                    v = _tree([(p, a * b) for (p, a, b) in terms], 0, size)
            row.append((c, v))
        rows.append(row)
    return rows


@pytest.mark.parametrize("kind", ["left_fold", "right_fold", "drop_empty", "fma_left", "fma_right"])
def test_kat_rejects_mutants(kind):
    disagree = 0
    for _, A, B, C in KATS:
        rows = _mutant(A, B, kind)
        vals = np.array([v for r in rows for (_, v) in r], np.float64)
        if len(vals) != len(C[2]) or not np.array_equal(bits(vals), bits(C[2])):
            disagree += 1
    assert disagree >= 1, f"the known answers do not distinguish the {kind} mistake"


# ---- random inputs: oracle == brute force, bit for bit ---------------------
def rand_csr(rng, M, K, per_row_max, empty_frac=0.2, adversarial=True):
    rpt, col, val = [0], [], []
    for _ in range(M):
        if rng.random() < empty_frac:
            d = 0
        else:
            d = int(rng.integers(0, per_row_max + 1))
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


@pytest.mark.parametrize("seed", range(12))
def test_random_vs_brute(seed):
    rng = np.random.default_rng(seed)
    M, K, N = [int(x) for x in rng.integers(1, 40, 3)]
    A = rand_csr(rng, M, K, 12)
    B = rand_csr(rng, K, N, 10)
    rpt, col, val = brute_spgemm(A, B)
    for f in (oracle.brmerge_precise, oracle.brmerge_upper):
        C = f(A, B, nthreads=3)
        assert_same(C, rpt, col, val)


def _scipy(T):
    return sp.csr_matrix((T[2], T[1], T[0].astype(np.int64)), shape=(len(T[0]) - 1, T[3]))


@pytest.mark.parametrize("seed", range(4))
def test_structure_and_values_vs_scipy(seed):
    rng = np.random.default_rng(100 + seed)
    A = rand_csr(rng, 300, 200, 30, adversarial=False)
    B = rand_csr(rng, 200, 250, 40, adversarial=False)
    C = oracle.brmerge_precise(A, B, nthreads=4)
    ones = lambda T: (T[0], T[1], np.ones_like(T[2]), T[3])  # noqa: E731
    P = (_scipy(ones(A)) @ _scipy(ones(B))).tocsr()
    P.sort_indices()
    np.testing.assert_array_equal(C.rpt, P.indptr)
    np.testing.assert_array_equal(C.col, P.indices)
    V = (_scipy(A) @ _scipy(B)).toarray()
    dense = np.zeros_like(V)
    for i in range(C.M):
        dense[i, C.col[C.rpt[i]:C.rpt[i + 1]]] = C.val[C.rpt[i]:C.rpt[i + 1]]
    scale = max(np.abs(V).max(), 1e-300)
    assert np.abs(dense - V).max() / scale <= 1e-12


def test_identity_gives_B_bitwise():
    rng = np.random.default_rng(7)
    B = rand_csr(rng, 50, 60, 9)
    I = (np.arange(51, dtype=np.int32), np.arange(50, dtype=np.int32), np.ones(50), 50)
    C = oracle.brmerge_precise(I, B)
    assert_same(C, B[0], B[1], B[2])


def test_one_entry_per_row_is_scaled_rows():
    rng = np.random.default_rng(8)
    B = rand_csr(rng, 20, 30, 7)
    cols = rng.integers(0, 20, 40).astype(np.int32)
    a = rng.uniform(-3, 3, 40)
    A = (np.arange(41, dtype=np.int32), cols, a, 20)
    C = oracle.brmerge_upper(A, B)
    for i in range(40):
        k = cols[i]
        bc, bv = B[1][B[0][k]:B[0][k + 1]], B[2][B[0][k]:B[0][k + 1]]
        np.testing.assert_array_equal(C.col[C.rpt[i]:C.rpt[i + 1]], bc)
        np.testing.assert_array_equal(bits(C.val[C.rpt[i]:C.rpt[i + 1]]), bits(a[i] * bv))


def test_flop_closed_forms():
    rng = np.random.default_rng(9)
    A = rand_csr(rng, 80, 70, 11)
    B = rand_csr(rng, 70, 90, 13)
    tot, nnzrA, flop = oracle.compute_flop(A[0], A[1], B[0])
    blen = np.diff(B[0])
    assert tot == int(blen[A[1]].sum())
    np.testing.assert_array_equal(nnzrA, np.diff(A[0]))
    np.testing.assert_array_equal(flop, [blen[A[1][A[0][i]:A[0][i + 1]]].sum() for i in range(80)])


def test_upper_log_prefix_and_search():
    assert [oracle.upper_log(n) for n in (0, 1, 2, 3, 4, 5, 1023, 1024, 1025)] == [2, 1, 2, 4, 4, 8, 1024, 1024, 2048]
    x = np.random.default_rng(1).integers(0, 100, 20000)
    for t in (1, 4):
        out = oracle.prefix_sum(x, t)
        assert out[0] == 0
        np.testing.assert_array_equal(out[1:], np.cumsum(x))
    arr = oracle.prefix_sum(x, 1)
    for elem in (0, 5, 1234, int(arr[-1]), int(arr[-1]) + 10):
        m = oracle.binary_search_approx(arr, elem)
        assert arr[m] <= elem and (m == len(arr) - 1 or elem < arr[m + 1] or arr[m] == elem)


def test_cancellation_kept_and_empty_rows():
    A = (np.array([0, 2, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, -1.0]), 2)
    B = (np.array([0, 1, 2], np.int32), np.array([3, 3], np.int32), np.array([0.5, 0.5]), 4)
    C = oracle.brmerge_precise(A, B)
    np.testing.assert_array_equal(C.rpt, [0, 1, 1])
    assert C.col.tolist() == [3] and C.val[0] == 0.0 and not np.signbit(C.val[0])


def test_precise_equals_upper_bitwise_on_skewed_input():
    rng = np.random.default_rng(11)
    A = rand_csr(rng, 400, 400, 60, empty_frac=0.4)
    C1 = oracle.brmerge_precise(A, A, nthreads=5)
    C2 = oracle.brmerge_upper(A, A, nthreads=3)
    assert_same(C1, C2.rpt, C2.col, C2.val)


# ---- Hash SpGEMM's accumulation order (Existing-SpGEMM/inc/hash_mult.h:558-608)
def _one_column(vals):
    """A 1 x L times L x 1 product whose single output sums `vals` over the
    lists in order (A all ones, B row p holds vals[p] in column 0)."""
    L = len(vals)
    A = (np.array([0, L], np.int32), np.arange(L, dtype=np.int32), np.ones(L), L)
    B = (np.arange(L + 1, dtype=np.int32), np.zeros(L, np.int32), np.array(vals, np.float64), 1)
    return A, B


def test_hash_order_known_answers():
    """The hash order adds each contribution onto the running value in list
    order: ((v0 + v1) + v2) + v3.  With b = 2^53 (where the spacing of
    doubles becomes 2, so b + 1 rounds to b, ties to even, while 1 - b is
    exact):
      v = (b, 1, 1, -b): left fold ((b + 1) + 1) - b = (b + 1) - b = 0;
                         tree (b + 1) + (1 - b) = b - (b - 1) = 1;
                         right fold b + (1 + (1 - b)) = b + (2 - b) = 2.
      v = (1, 1e16, -1e16, 1) (spacing 2 at 1e16, so 1e16 +- 1 rounds back):
                         left fold ((1 + 1e16) - 1e16) + 1 = 1; tree
                         (1 + 1e16) + (1 - 1e16) = 1e16 - 1e16 = 0."""
    b = 2.0 ** 53
    A, B = _one_column([b, 1.0, 1.0, -b])
    assert oracle.hash_spgemm(A, B).val.tolist() == [0.0]
    assert oracle.brmerge_precise(A, B).val.tolist() == [1.0]
    A, B = _one_column([1.0, 1e16, -1e16, 1.0])
    assert oracle.hash_spgemm(A, B).val.tolist() == [1.0]
    assert oracle.brmerge_precise(A, B).val.tolist() == [0.0]
    # cap: rows above it keep BRMerge's tree
    assert oracle.hash_spgemm(A, B, cap=3).val.tolist() == [0.0]
    assert oracle.hash_spgemm(A, B, cap=4).val.tolist() == [1.0]
    assert oracle.hash_spgemm(A, B, cap=4, lmax=3).val.tolist() == [0.0]  # more lists than lmax: the tree


@pytest.mark.parametrize("seed", range(6))
def test_hash_order_structure_and_two_contributions(seed):
    """Same structure as Precise; values within 1e-12 of scipy; and bit-equal
    to Precise wherever a column has at most two contributions (one add: no
    order to choose)."""
    rng = np.random.default_rng(500 + seed)
    from tests._util import rand_csr_tuple
    M, K, N = [int(x) for x in rng.integers(1, 200, 3)]
    A = rand_csr_tuple(rng, M, K, int(rng.integers(1, 40)), empty_frac=0.2, adversarial=False)
    B = rand_csr_tuple(rng, K, N, int(rng.integers(1, 30)), empty_frac=0.2, adversarial=False)
    H = oracle.hash_spgemm(A, B, nthreads=2)
    P = oracle.brmerge_precise(A, B, nthreads=2)
    np.testing.assert_array_equal(H.rpt, P.rpt)
    np.testing.assert_array_equal(H.col, P.col)
    Asp = sp.csr_matrix((A[2], A[1], A[0]), shape=(M, K))
    Bsp = sp.csr_matrix((B[2], B[1], B[0]), shape=(K, N))
    ones = sp.csr_matrix((np.ones(len(A[1])), A[1], A[0]), shape=(M, K)) @ \
        sp.csr_matrix((np.ones(len(B[1])), B[1], B[0]), shape=(K, N))
    ones.sort_indices()
    np.testing.assert_array_equal(ones.indices, H.col)
    ref = (Asp @ Bsp).toarray()
    dense = np.zeros((M, N))
    for i in range(M):
        dense[i, H.col[H.rpt[i]:H.rpt[i + 1]]] = H.val[H.rpt[i]:H.rpt[i + 1]]
    scale = max(np.abs(ref).max(), 1e-300)
    assert np.abs(dense - ref).max() / scale <= 1e-12
    few = np.asarray(ones.data) <= 2  # contributions per output entry
    assert np.array_equal(bits(H.val[few]), bits(P.val[few]))
