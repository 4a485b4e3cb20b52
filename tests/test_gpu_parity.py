"""GPU parity: libbrmerge.so (HIP, gfx950) against the CPU oracle.

Structure (row pointer, columns) must match exactly; values are required to
be within 1e-12 max-abs-normalised relative error and are checked
bit-for-bit as well, since the GPU keeps the reference's tree bracketing and
rounds every product before adding it.
"""
import numpy as np
import pytest

import oracle
from tests._util import as_tuple, assert_csr_equal, csr_from_tuple, rand_csr_tuple
from tests.test_oracle import KATS

pytestmark = pytest.mark.gpu

TIER_SETTINGS = {
    "default": {},
    "all_long": {"force_long": 1, "group_cfg": 0},
    "tiny_leaves": {"force_long": 1, "leaf_cap": 8, "group_cfg": 0},
    "mergepath": {"group_cfg": 0},
    "small_medium": {"med_cap": 64, "leaf_cap": 100, "wmed_cap": 0},
    "wave_medium": {"wmed_cap": 1024},
    "no_wave_medium": {"wmed_cap": 0},
    "bitmap_all_long": {"force_long": 1, "group_cfg": 4},
    "bitmap": {"group_cfg": 4},
    "tiny_launches": {"launch_chunk": 3, "group_cfg": 4},
    "tiny_launches_mergepath": {"launch_chunk": 3, "group_cfg": 0},
    "register_medium_128": {"wmed_cap": 128, "wmed_engine": 1},
    "register_medium_512": {"wmed_cap": 512, "wmed_engine": 1, "med_cap": 512},
    "register_medium_1024": {"wmed_cap": 1024, "wmed_engine": 1},
    "lds_wave_medium": {"wmed_cap": 512, "wmed_engine": 0},
    "general_pipeline": {"small_fused": 0},
    "wave_medium_hash_count": {"wmed_cap": 1024, "wmed_engine": 1, "wmed_count": 0},
    "register_medium_sort": {"wmed_cap": 1024, "wmed_engine": 1, "wmed_bitmap": 0},
    "register_medium_2048": {"wmed_cap": 2048, "wmed_engine": 1},
    "register_medium_2048_hash_count": {"wmed_cap": 2048, "wmed_engine": 1, "wmed_count": 0},
}
KNOBS = ("force_long", "leaf_cap", "med_cap", "wmed_cap", "group_cfg", "launch_chunk", "wmed_engine", "small_fused",
         "wmed_count", "wmed_bitmap")


@pytest.fixture
def tier(request, gpu):
    name = request.param
    saved = {k: gpu.get_param(k) for k in KNOBS}
    for k, v in TIER_SETTINGS[name].items():
        gpu.set_param(k, v)
    yield name
    for k, v in saved.items():
        gpu.set_param(k, v)


@pytest.fixture(params=[0, 4], ids=["mergepath", "bitmap"])
def engine(request, gpu):
    """Long-row group-merge engine: merge-path levels (0) or bitmap ranks (4)."""
    old = gpu.get_param("group_cfg")
    gpu.set_param("group_cfg", request.param)
    yield request.param
    gpu.set_param("group_cfg", old)


def run_both(bm, A, B, i64=False):
    Ah, Bh = csr_from_tuple(A, len(B[0]) - 1), csr_from_tuple(B)
    Cp = bm.brmerge_precise(Ah, Bh, i64=i64)
    Cu = bm.brmerge_upper(Ah, Bh, i64=i64)
    return Cp, Cu


@pytest.mark.parametrize("case", KATS, ids=[k[0] for k in KATS])
@pytest.mark.parametrize("tier", list(TIER_SETTINGS), indirect=True)
def test_kat_gpu(gpu, case, tier):
    _, A, B, C = case
    Cp, Cu = run_both(gpu, A, B)
    from oracle import OracleCSR
    R = OracleCSR(len(C[0]) - 1, B[3], C[0], C[1], C[2])
    assert_csr_equal(Cp, R)
    assert_csr_equal(Cu, R)


@pytest.mark.parametrize("seed", range(16))
@pytest.mark.parametrize("tier", list(TIER_SETTINGS), indirect=True)
def test_random_small(gpu, seed, tier):
    rng = np.random.default_rng(1000 + seed)
    M, K, N = [int(x) for x in rng.integers(1, 300, 3)]
    A = rand_csr_tuple(rng, M, K, int(rng.integers(1, 80)), empty_frac=0.3)
    B = rand_csr_tuple(rng, K, N, int(rng.integers(1, 60)), empty_frac=0.3)
    R = oracle.brmerge_precise(A, B, nthreads=4)
    Cp, Cu = run_both(gpu, A, B)
    assert_csr_equal(Cp, R)
    assert_csr_equal(Cu, R)
    Cp64, Cu64 = run_both(gpu, A, B, i64=True)
    assert_csr_equal(Cp64, R)
    assert_csr_equal(Cu64, R)


def test_er16384_config(gpu):
    A = gpu.gen.er(16384, 16384, 8, 0)
    R = oracle.brmerge_precise(as_tuple(A), as_tuple(A))
    assert gpu.compute_flop(A, A) == 8 * 8 * 16384
    assert_csr_equal(gpu.brmerge_precise(A, A), R)
    assert_csr_equal(gpu.brmerge_upper(A, A), R)


@pytest.mark.parametrize("scale", [10, 13, 15])
def test_rmat(gpu, engine, scale):
    A = gpu.gen.rmat(scale)
    R = oracle.brmerge_precise(as_tuple(A), as_tuple(A))
    assert_csr_equal(gpu.brmerge_precise(A, A), R)
    assert_csr_equal(gpu.brmerge_upper(A, A), R)


def test_flop_count_many_tiles(gpu):
    """Product count of an A with 40M entries: the prefix sum takes its
    three-kernel path (more tiles than the down pass sums by itself), and
    the total matches numpy."""
    from types import SimpleNamespace
    rng = np.random.default_rng(5)
    M, per, K = 10_000_000, 4, 1000
    blen = rng.integers(0, 6, size=K)
    brpt = np.concatenate([[0], np.cumsum(blen)]).astype(np.int32)
    acol = rng.integers(0, K, size=M * per).astype(np.int32)
    arpt = (np.arange(M + 1, dtype=np.int64) * per).astype(np.int32)
    A = SimpleNamespace(rpt=arpt, col=acol, M=M)
    B = SimpleNamespace(rpt=brpt, M=K)
    assert gpu.compute_flop(A, B) == int(blen[acol].sum())


def test_amg_chain(gpu):
    n = 21
    S, P = gpu.gen.stencil27(n), gpu.gen.aggregation(n)
    AP = gpu.brmerge_precise(S, P)
    R1 = oracle.brmerge_precise(as_tuple(S), as_tuple(P))
    assert_csr_equal(AP, R1)
    assert gpu.compute_flop(S, P) == S.nnz
    PT = P.transpose()
    C = gpu.brmerge_precise(PT, AP)
    R2 = oracle.brmerge_precise(as_tuple(PT), (np.asarray(R1.rpt, np.int32), R1.col, R1.val, P.N))
    assert_csr_equal(C, R2)
    assert gpu.compute_flop(PT, AP) == AP.nnz


def test_banded(gpu):
    A = gpu.gen.banded(5000, 64, 4, 32, 3)
    B = gpu.gen.banded(5000, 64, 4, 32, 4)
    R = oracle.brmerge_precise(as_tuple(A), as_tuple(B))
    Cp = gpu.brmerge_precise(A, B)
    assert_csr_equal(Cp, R)


def test_edge_shapes(gpu):
    bm = gpu
    # empty A (M = 0)
    A = bm.CSR(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), 0, 5)
    B = bm.gen.er(5, 7, 2, 1)
    C = bm.brmerge_precise(A, B)
    assert C.M == 0 and C.nnz == 0
    # A without entries, B without entries
    A = bm.CSR(np.zeros(4, np.int32), np.zeros(0, np.int32), np.zeros(0), 3, 5)
    C = bm.brmerge_upper(A, B)
    assert C.nnz == 0 and list(C.rpt) == [0, 0, 0, 0]
    Bz = bm.CSR(np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0), 5, 7)
    A = bm.gen.er(3, 5, 4, 2)
    C = bm.brmerge_precise(A, Bz)
    assert C.nnz == 0


def test_row_with_many_empty_lists(gpu, engine):
    """Rows with thousands of A entries, almost all pointing at empty B rows."""
    rng = np.random.default_rng(5)
    K, N = 20000, 3000
    brows = [np.sort(rng.choice(N, size=(int(rng.integers(1, 40)) if r % 97 == 0 else 0), replace=False))
             for r in range(K)]
    brpt = np.concatenate([[0], np.cumsum([len(b) for b in brows])]).astype(np.int32)
    B = (brpt, np.concatenate(brows).astype(np.int32), rng.uniform(-1, 1, brpt[-1]), N)
    acols = [np.sort(rng.choice(K, size=s, replace=False)) for s in (5000, 700, 64, 1, 12000)]
    arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
    A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
    R = oracle.brmerge_precise(A, B)
    Cp, Cu = run_both(gpu, A, B)
    assert_csr_equal(Cp, R)
    assert_csr_equal(Cu, R)


def test_single_long_list(gpu, engine):
    """A row selecting one B row longer than any LDS tile (a copied list)."""
    rng = np.random.default_rng(6)
    N = 100000
    bc = np.sort(rng.choice(N, size=30000, replace=False)).astype(np.int32)
    B = (np.array([0, 30000, 30000], np.int32), bc, rng.uniform(-1, 1, 30000), N)
    A = (np.array([0, 1, 3, 4], np.int32), np.array([0, 0, 1, 1], np.int32), np.array([0.5, 3.0, 2.0, 1.0]), 2)
    R = oracle.brmerge_precise(A, B)
    Cp, Cu = run_both(gpu, A, B)
    assert_csr_equal(Cp, R)
    assert_csr_equal(Cu, R)


def test_long_rows_many_levels(gpu, engine):
    """Rows far beyond one tile with deep trees, dense column overlap."""
    rng = np.random.default_rng(7)
    K, N = 3000, 6000
    B = rand_csr_tuple(rng, K, N, 300, empty_frac=0.1)
    acols = [np.sort(rng.choice(K, size=s, replace=False)) for s in (2900, 1500, 257, 513, 100)]
    arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
    A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-2, 2, arpt[-1]), K)
    R = oracle.brmerge_precise(A, B)
    Cp, Cu = run_both(gpu, A, B)
    assert_csr_equal(Cp, R)
    assert_csr_equal(Cu, R)


def test_device_entry_and_determinism(gpu):
    bm = gpu
    A = bm.gen.rmat(12, seed=3)
    dA = bm.DeviceCSR.upload(A)
    try:
        t = bm.Precise_Timing()
        C1 = bm.spgemm_dev(dA, dA, timing=t)
        C2 = bm.spgemm_dev(dA, dA)
        C3 = bm.spgemm_dev(dA, dA, upper=True)
        h1, h2, h3 = C1.download(), C2.download(), C3.download()
        for C in (C1, C2, C3):
            C.free()
    finally:
        dA.free()
    R = oracle.brmerge_precise(as_tuple(A), as_tuple(A))
    for h in (h1, h2, h3):
        assert_csr_equal(h, R)
    assert t.total > 0 and t.compute > 0


def test_invalid_inputs(gpu):
    bm = gpu
    good = bm.gen.er(50, 50, 5, 0)
    # unsorted B row
    B = bm.gen.er(50, 50, 5, 1)
    B.col = B.col.copy()
    B.col[0], B.col[1] = B.col[1], B.col[0]
    with pytest.raises(bm.BrmError) as e:
        bm.brmerge_precise(good, B)
    assert e.value.status == 1
    # column out of range
    B = bm.gen.er(50, 50, 5, 1)
    B.col = B.col.copy()
    B.col[7] = 50
    with pytest.raises(bm.BrmError):
        bm.brmerge_upper(good, B)
    # A column beyond K
    A = bm.gen.er(50, 50, 5, 2)
    A.col = A.col.copy()
    A.col[-1] = 999
    with pytest.raises(bm.BrmError):
        bm.brmerge_precise(A, good)
    # non-monotone row pointer
    A = bm.gen.er(50, 50, 5, 2)
    A.rpt = A.rpt.copy()
    A.rpt[10] = A.rpt[11] + 2
    with pytest.raises(bm.BrmError):
        bm.brmerge_precise(A, good)
    # the library is still usable afterwards
    R = oracle.brmerge_precise(as_tuple(good), as_tuple(good))
    assert_csr_equal(bm.brmerge_precise(good, good), R)


def _skewed_b(rng):
    """B with a 3,000-entry first row, runs of empty rows and short rows
    (valid: columns strictly increase within every row, and often fall
    from one row to the next)."""
    lens = [3000]
    while len(lens) < 2000:
        lens += [0] * int(rng.integers(0, 100)) if rng.random() < 0.02 else [int(rng.integers(0, 40))]
    lens = np.array(lens[:2000], np.int64)
    rpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(5000, n, replace=False)) for n in lens]).astype(np.int32)
    return rpt, col, rng.uniform(-1, 1, len(col)), 5000


@pytest.mark.parametrize("launch_chunk", [0, 3], ids=["one_launch", "tiny_launches"])
def test_unsorted_b_rows_skewed(gpu, launch_chunk):
    """The sortedness check of B (fixed chunks of 512 entries per wavefront)
    refuses a descent or a repeated column anywhere inside a row: inside the
    hub row, across a chunk edge, after a run of empty rows, at the last pair,
    and in a resident B of the request server; it accepts the valid B."""
    bm = gpu
    rng = np.random.default_rng(5)
    B = _skewed_b(rng)
    rpt, col = B[0], B[1]
    nnz = int(rpt[-1])
    A = rand_csr_tuple(rng, 300, len(rpt) - 1, 6)
    row_of = np.repeat(np.arange(len(rpt) - 1), np.diff(rpt))
    same = np.nonzero(row_of[:-1] == row_of[1:])[0]  # pairs (k, k + 1) inside one row
    after_empty = [k for k in same if rpt[row_of[k]] == k and row_of[k] > 0 and rpt[row_of[k] - 1] == k]
    at_edge = [k for k in same if k % 512 in (0, 511)]
    picks = [0, 511, 512, 2998, int(same[-1])] + at_edge[-3:] + after_empty[:2]
    assert nnz > 4 * 512 and len(after_empty) > 0
    old = bm.get_param("launch_chunk")
    bm.set_param("launch_chunk", launch_chunk)
    try:
        Bh = csr_from_tuple(B)
        Ah = csr_from_tuple(A, len(rpt) - 1)
        assert_csr_equal(bm.brmerge_precise(Ah, Bh), oracle.brmerge_precise(A, B))
        for i, k in enumerate(picks):
            c = col.copy()
            if i % 2:
                c[k + 1] = c[k]  # repeated column
            else:
                c[k], c[k + 1] = c[k + 1], c[k]
            with pytest.raises(bm.BrmError) as e:
                bm.brmerge_precise(Ah, csr_from_tuple((rpt, c, B[2], B[3])))
            assert e.value.status == 1, k
        c = col.copy()
        c[picks[1]], c[picks[1] + 1] = c[picks[1] + 1], c[picks[1]]
        dB = bm.DeviceCSR.upload(csr_from_tuple((rpt, c, B[2], B[3])))
        try:
            with bm.Server(budget_bytes=1 << 30) as srv:
                with pytest.raises(bm.BrmError) as e:
                    srv.add_b(dB)
                assert e.value.status == 1
        finally:
            dB.free()
        assert_csr_equal(bm.brmerge_precise(Ah, Bh), oracle.brmerge_precise(A, B))
    finally:
        bm.set_param("launch_chunk", old)


def test_int32_overflow(gpu):
    """nnz(C) = 65,536 * 64 * 1,024 > INT32_MAX must be refused by the _i32 entry."""
    bm = gpu
    rows, lists, width = 65536, 64, 1024
    B = bm.CSR(np.arange(lists + 1, dtype=np.int32) * width, np.arange(lists * width, dtype=np.int32),
               np.ones(lists * width), lists, lists * width)
    A = bm.CSR(np.arange(rows + 1, dtype=np.int32) * lists, np.tile(np.arange(lists, dtype=np.int32), rows),
               np.ones(rows * lists), rows, lists)
    with pytest.raises(bm.BrmError) as e:
        bm.brmerge_precise(A, B)
    assert e.value.status == 3


def test_allocation_failure_reports_nomem(gpu):
    bm = gpu
    A = bm.gen.rmat(11, seed=9)
    L = bm.lib()
    for n in (0, 3, 7):
        L.brm_debug_fail_alloc_after(n)
        with pytest.raises(bm.BrmError) as e:
            bm.brmerge_precise(A, A)
        assert e.value.status == 2
    L.brm_debug_fail_alloc_after(-1)
    R = oracle.brmerge_precise(as_tuple(A), as_tuple(A))
    assert_csr_equal(bm.brmerge_precise(A, A), R)


@pytest.mark.parametrize("tier", list(TIER_SETTINGS), indirect=True)
def test_signed_zero_products(gpu, tier):
    """Explicit +0.0 / -0.0 A values: every product must still be rounded
    and summed (sign of zero included), in every tier."""
    rng = np.random.default_rng(21)
    K, N = 800, 2000
    B = rand_csr_tuple(rng, K, N, 120, empty_frac=0.1)
    acols = [np.sort(rng.choice(K, size=s, replace=False)) for s in (700, 90, 40, 3, 500)]
    arpt = np.concatenate([[0], np.cumsum([len(x) for x in acols])]).astype(np.int32)
    aval = rng.uniform(-1, 1, arpt[-1])
    aval[::3] = 0.0
    aval[1::7] = -0.0
    A = (arpt, np.concatenate(acols).astype(np.int32), aval, K)
    R = oracle.brmerge_precise(A, B)
    Cp, Cu = run_both(gpu, A, B)
    assert_csr_equal(Cp, R)
    assert_csr_equal(Cu, R)


@pytest.mark.parametrize("force_long", [0, 1])
def test_wide_columns_and_hub_columns(gpu, engine, force_long):
    """Columns spread over a million (several bitmap spans per node), a few
    columns shared by most B rows (many contributions per output, deep
    trees of pass-throughs), and rows on both sides of 2^5, 2^6, 2^8, 2^10
    and 2^12 lists (pass boundaries of the merge-path engine, fan-in 2^5,
    and of the bitmap engine, fan-in 2^6)."""
    old = gpu.get_param("force_long")
    gpu.set_param("force_long", force_long)
    try:
        rng = np.random.default_rng(11)
        K, N = 4500, 1_000_000
        hubs = np.array([0, 17, 500_000, 999_999])
        rows = []
        for r in range(K):
            d = int(rng.integers(0, 40))
            c = rng.choice(N, size=d, replace=False)
            c = np.union1d(c, hubs[rng.random(len(hubs)) < 0.8])
            rows.append(np.sort(c))
        brpt = np.concatenate([[0], np.cumsum([len(b) for b in rows])]).astype(np.int32)
        B = (brpt, np.concatenate(rows).astype(np.int32), rng.uniform(-1, 1, brpt[-1]), N)
        acols = [np.sort(rng.choice(K, size=sz, replace=False))
                 for sz in (4097, 4096, 3000, 1025, 1024, 600, 257, 256, 255, 65, 64, 40, 33, 32, 3)]
        arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
        A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
        R = oracle.brmerge_precise(A, B)
        Cp, Cu = run_both(gpu, A, B)
        assert_csr_equal(Cp, R)
        assert_csr_equal(Cu, R)
    finally:
        gpu.set_param("force_long", old)


def test_four_pass_rows(gpu, engine):
    """Rows on both sides of 2^15 lists: three passes at fan-in 2^5 (the
    merge-path engine) for 32,768 lists, four for 32,769 (three either way
    at the bitmap engine's 2^6), with B rows of 0 to 12 entries (empty
    lists among them) and columns over 300,000 (three bitmap spans)."""
    rng = np.random.default_rng(23)
    K, N = 40_000, 300_000
    lens = rng.integers(0, 13, K)
    brpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    bcol = np.concatenate([np.sort(rng.choice(N, size=int(d), replace=False)) for d in lens]).astype(np.int32)
    B = (brpt, bcol, rng.uniform(-1, 1, brpt[-1]), N)
    acols = [np.sort(rng.choice(K, size=sz, replace=False)) for sz in (32_769, 32_768, 1_025)]
    arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
    A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
    R = oracle.brmerge_precise(A, B)
    Cp, Cu = run_both(gpu, A, B)
    assert_csr_equal(Cp, R)
    assert_csr_equal(Cu, R)


def test_bitmap_four_pass_rows(gpu):
    """Rows on both sides of 2^18 lists: three passes of the bitmap engine
    (fan-in 2^6) for 262,144 lists, four for 262,145; B rows of 0 to 3
    entries over 200,000 columns."""
    old = gpu.get_param("group_cfg")
    gpu.set_param("group_cfg", 4)
    try:
        rng = np.random.default_rng(29)
        K, N = 270_000, 200_000
        lens = rng.integers(0, 4, K)
        brpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        bcol = np.concatenate([np.sort(rng.choice(N, size=int(d), replace=False)) for d in lens]).astype(np.int32)
        B = (brpt, bcol, rng.uniform(-1, 1, brpt[-1]), N)
        acols = [np.sort(rng.choice(K, size=sz, replace=False)) for sz in (262_145, 262_144)]
        arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
        A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
        R = oracle.brmerge_precise(A, B)
        Cp, Cu = run_both(gpu, A, B)
        assert gpu.last_stats()["passes"] == 4
        assert_csr_equal(Cp, R)
        assert_csr_equal(Cu, R)
    finally:
        gpu.set_param("group_cfg", old)


@pytest.mark.parametrize("ncols", [1 << 16, (1 << 16) + 1, 1 << 17, (1 << 17) + 1, 1 << 20, (1 << 20) + 1],
                         ids=lambda n: f"N{n}")
@pytest.mark.parametrize("wmed_cap", [128, 512, 1024, 2048])
@pytest.mark.parametrize("ranks", ["sort", "bitmap"])
def test_register_medium_key_widths(gpu, ncols, wmed_cap, ranks):
    """Medium rows of the register tier with columns up to N - 1 on both
    sides of the 32-bit key limit (2^16 columns at 1024 products, 2^17 at
    512) and of the bitmap variants' limits (2^16 columns flat, 2^20
    two-level, sort beyond): keys of either width and ranks of either kind
    give the same bits.  A cap of 2048 bins rows of 1025-2048 products into
    the register tier only where the bitmap ranks run (else 1024 holds)."""
    saved = {k: gpu.get_param(k) for k in ("wmed_cap", "wmed_engine", "wmed_bitmap")}
    gpu.set_param("wmed_cap", wmed_cap)
    gpu.set_param("wmed_engine", 1)
    gpu.set_param("wmed_bitmap", 1 if ranks == "bitmap" else 0)
    try:
        rng = np.random.default_rng(ncols + wmed_cap)
        K = 400
        rows = []
        for r in range(K):
            d = int(rng.integers(0, 24))
            top = rng.choice([ncols - 1, ncols - 2]) if r % 3 == 0 else None
            c = rng.choice(ncols, size=d, replace=False)
            if top is not None:
                c = np.unique(np.concatenate([c, [top]]))
            rows.append(np.sort(c))
        brpt = np.concatenate([[0], np.cumsum([len(b) for b in rows])]).astype(np.int32)
        B = (brpt, np.concatenate(rows).astype(np.int32), rng.uniform(-1, 1, brpt[-1]), ncols)
        acols = [np.sort(rng.choice(K, size=int(rng.integers(2, 64)), replace=False)) for _ in range(300)]
        arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
        A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
        R = oracle.brmerge_precise(A, B, nthreads=4)
        Cp, Cu = run_both(gpu, A, B)
        assert_csr_equal(Cp, R)
        assert_csr_equal(Cu, R)
    finally:
        for k, v in saved.items():
            gpu.set_param(k, v)


@pytest.mark.parametrize("stride", [5, 130001], ids=["narrow", "wide"])
@pytest.mark.parametrize("engine_wm", [0, 1], ids=["lds", "registers"])
def test_medium_rows_with_deep_column_trees(gpu, engine_wm, stride):
    """Rows of 64 lists whose B rows share a handful of columns: every output
    column sums up to 64 contributions through all six tree levels, with
    absent lists at irregular positions.  The columns lie `stride` apart: 40
    columns (flat bitmap ranks) or about 2^20 (two-level ranks)."""
    saved = {k: gpu.get_param(k) for k in ("wmed_cap", "wmed_engine")}
    gpu.set_param("wmed_cap", 1024)
    gpu.set_param("wmed_engine", engine_wm)
    try:
        rng = np.random.default_rng(13)
        K, N = 300, 8 * stride
        rows = []
        for r in range(K):
            d = 0 if r % 7 == 3 else int(rng.integers(1, 9))
            rows.append(np.sort(rng.choice(8, size=d, replace=False) * stride))
        brpt = np.concatenate([[0], np.cumsum([len(b) for b in rows])]).astype(np.int32)
        mags = rng.choice([1.0, 2.0 ** 40, 2.0 ** -40, 3.0], size=brpt[-1])
        B = (brpt, np.concatenate(rows).astype(np.int32), rng.uniform(-1, 1, brpt[-1]) * mags, N)
        acols = [np.sort(rng.choice(K, size=sz, replace=False)) for sz in [64] * 40 + [33, 17, 9, 63]]
        arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
        A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
        R = oracle.brmerge_precise(A, B, nthreads=4)
        Cp, Cu = run_both(gpu, A, B)
        assert_csr_equal(Cp, R)
        assert_csr_equal(Cu, R)
    finally:
        for k, v in saved.items():
            gpu.set_param(k, v)
