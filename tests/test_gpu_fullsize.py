"""Parity at the benchmarked sizes, and on the paths that only large inputs
reach: every comparison is bit for bit against the oracle (the OpenMP
restatement of BRMerge-SpGEMM/inc/spgemm.h:468-735).

- R-MAT scale 20, C = A*A (nnz(C) = 9.71e9, C offsets far past 2^31):
  Precise and Upper on the device, checked on a seeded random panel of
  rows plus the heaviest rows against the oracle on the same rows of A
  times A; the whole row pointer and device checksums of the whole C must
  agree between Precise and Upper (Upper runs in several C-bar row waves).
- AMG Galerkin product at 96^3 (A*P and P^T*(A*P)), whole, against the
  oracle.
- Long rows split over several workspace waves, and Upper over several
  C-bar waves, forced by small budgets.
- Columns beyond 2^20, so the long-row symbolic count runs several column
  spans of its LDS bitmap.
"""
import numpy as np
import pytest

import oracle
from tests._util import csr_from_tuple, as_tuple, assert_csr_equal, bits

pytestmark = pytest.mark.gpu


def row_ip(A, B):
    blen = np.diff(np.asarray(B.rpt, np.int64))
    cs = np.concatenate([[0], np.cumsum(blen[A.col])])
    rpt = np.asarray(A.rpt, np.int64)
    return cs[rpt[1:]] - cs[rpt[:-1]]


def panel_rows(M, ip, n, heavy, seed):
    rng = np.random.default_rng(seed)
    rows = set(rng.choice(M, size=n, replace=False).tolist()) | set(np.argsort(ip)[-heavy:].tolist())
    return np.array(sorted(rows), np.int64)


def oracle_rows(A, B, rows):
    """The oracle's product of rows `rows` of A with B."""
    rpt = np.asarray(A.rpt, np.int64)
    lens = rpt[rows + 1] - rpt[rows]
    prpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.arange(rpt[r], rpt[r + 1]) for r in rows])
    return oracle.brmerge_precise((prpt, A.col[idx], A.val[idx], A.N), as_tuple(B))


def assert_panel_equal(got, R):
    rpt, col, val = got
    np.testing.assert_array_equal(rpt, R.rpt)
    np.testing.assert_array_equal(col, R.col)
    mism = np.nonzero(bits(val) != bits(R.val))[0]
    assert len(mism) == 0, f"{len(mism)} values differ in bits, first at {mism[:5]}"


@pytest.fixture
def knobs(gpu):
    names = ("ws_budget_bytes", "cbar_budget_bytes", "group_cfg", "force_long", "sym_chunk_products", "upper_vm")
    saved = {k: gpu.get_param(k) for k in names}
    yield gpu
    for k, v in saved.items():
        gpu.set_param(k, v)


def test_rmat20_precise_and_upper_panels(gpu):
    A = gpu.gen.rmat(20)
    ip = row_ip(A, A)
    rows = panel_rows(A.M, ip, 2000, 8, seed=20)
    dA = gpu.DeviceCSR.upload(A)
    try:
        C = gpu.spgemm_dev(dA, dA)
        st = gpu.last_stats()
        nnz = C.nnz
        rpt = C.rpt.to_host()
        sums = (C.col.checksum(), C.val.checksum())
        got_p = C.download_panel(rows)
        C.free()
        assert nnz == 9_710_809_979 and st["total_ip"] == int(ip.sum())
        assert st["long_waves"] >= 1
        Cu = gpu.spgemm_dev(dA, dA, upper=True)
        stu = gpu.last_stats()
        rpt_u = Cu.rpt.to_host()
        sums_u = (Cu.col.checksum(), Cu.val.checksum())
        got_u = Cu.download_panel(rows)
        Cu.free()
    finally:
        dA.free()
    # C-bar (12 bytes per product, 251 GB) cannot be held at once: row waves.
    assert stu["upper_waves"] > 1
    np.testing.assert_array_equal(rpt_u, rpt)
    assert sums_u == sums
    # the panel reaches C offsets past 2^31 and the product's heaviest rows
    assert int((rpt[rows] > 2 ** 31).sum()) > 1000
    R = oracle_rows(A, A, rows)
    assert_panel_equal(got_p, R)
    assert_panel_equal(got_u, R)


def test_amg96_galerkin_whole(gpu):
    n = 96
    S, P = gpu.gen.stencil27(n), gpu.gen.aggregation(n)
    R1 = oracle.brmerge_precise(as_tuple(S), as_tuple(P))
    AP = gpu.brmerge_precise(S, P)
    assert_csr_equal(AP, R1)
    assert_csr_equal(gpu.brmerge_upper(S, P), R1)
    PT = P.transpose()
    R2 = oracle.brmerge_precise(as_tuple(PT), (np.asarray(R1.rpt, np.int32), R1.col, R1.val, P.N))
    C = gpu.brmerge_precise(PT, AP)
    assert_csr_equal(C, R2)
    assert_csr_equal(gpu.brmerge_upper(PT, AP), R2)


@pytest.mark.parametrize("engine,vm", [(0, 1), (4, 1), (4, 0)], ids=["mergepath", "bitmap", "bitmap-chunks"])
def test_rmat15_many_waves(knobs, engine, vm):
    """Long rows over at least three workspace waves and Upper over at least
    three C-bar waves, on both long-row engines; Upper's C grown in a
    reserved address range, or assembled from copied chunks."""
    gpu = knobs
    gpu.set_param("group_cfg", engine)
    gpu.set_param("upper_vm", vm)
    A = gpu.gen.rmat(15)
    ip = row_ip(A, A)
    R = oracle.brmerge_precise(as_tuple(A), as_tuple(A))
    long_bytes = 24 * int(ip[ip > 4096].sum())
    gpu.set_param("ws_budget_bytes", max(long_bytes // 6, 1))
    gpu.set_param("cbar_budget_bytes", 12 * int(ip.sum()) // 5)
    Cp = gpu.brmerge_precise(A, A)
    assert gpu.last_stats()["long_waves"] >= 3
    assert_csr_equal(Cp, R)
    Cu = gpu.brmerge_upper(A, A)
    st = gpu.last_stats()
    assert st["upper_waves"] >= 3 and st["long_waves"] >= 3
    assert_csr_equal(Cu, R)
    # a C-bar budget below one row still makes progress: one non-empty row per wave
    gpu.set_param("cbar_budget_bytes", 12)
    gpu.set_param("ws_budget_bytes", 0)
    A2 = gpu.gen.rmat(9, seed=4)
    R2 = oracle.brmerge_precise(as_tuple(A2), as_tuple(A2))
    assert_csr_equal(gpu.brmerge_upper(A2, A2), R2)
    assert gpu.last_stats()["upper_waves"] == int((row_ip(A2, A2) > 0).sum())


@pytest.mark.parametrize("budget", [1, 3 << 20], ids=["one-row-waves", "few-row-waves"])
def test_overlapped_waves_mixed_passes(knobs, budget):
    """Consecutive long-row waves overlap on two streams (bitmap engine),
    handing the workspace buffers over by events.  Rows of one, two and three
    passes (64, 65 to 4,096 and more than 4,096 lists) in mixed order over many
    waves: every handover order between a wave and the next one occurs."""
    gpu = knobs
    gpu.set_param("group_cfg", 4)
    gpu.set_param("force_long", 1)
    rng = np.random.default_rng(41)
    K, N = 6000, 400_000
    lens = rng.integers(0, 9, K)
    brpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    bcol = np.concatenate([np.sort(rng.choice(N, size=int(d), replace=False)) for d in lens]).astype(np.int32)
    B = (brpt, bcol, rng.uniform(-1, 1, brpt[-1]), N)
    sizes = (4500, 64, 3000, 65, 5000, 40, 4097, 1, 200, 4096, 5900, 2, 64)
    acols = [np.sort(rng.choice(K, size=sz, replace=False)) for sz in sizes]
    arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
    A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
    R = oracle.brmerge_precise(A, B)
    gpu.set_param("ws_budget_bytes", budget)
    Ah, Bh = csr_from_tuple(A), csr_from_tuple(B)
    for _ in range(2):
        assert_csr_equal(gpu.brmerge_precise(Ah, Bh), R)
        st = gpu.last_stats()
        assert st["long_waves"] >= (4 if budget == 1 else 2) and st["passes"] == 3
    assert_csr_equal(gpu.brmerge_upper(Ah, Bh), R)


@pytest.mark.parametrize("ncols", [(1 << 20) + (1 << 19), 3 * (1 << 20) + 5], ids=["N1.5M", "N3M"])
def test_long_rows_columns_beyond_2p20(knobs, ncols):
    """B rows with columns on both sides of 2^20 and 2^21 (several column
    spans of the long-row symbolic's 2^20-bit LDS bitmap), shared hub
    columns at the span edges, and rows of 300 to 3,000 lists."""
    gpu = knobs
    rng = np.random.default_rng(ncols)
    K = 5000
    edges = np.array([e for e in ((1 << 20) - 1, 1 << 20, (1 << 21) - 1, 1 << 21, 3 << 20) if e < ncols] +
                     [0, ncols - 1])
    brows = []
    for _ in range(K):
        d = int(rng.integers(0, 48))
        c = rng.choice(ncols, size=d, replace=False)
        near = (edges[rng.integers(0, len(edges), 4)] + rng.integers(-3, 4, 4)).clip(0, ncols - 1)
        c = np.union1d(c, near[rng.random(4) < 0.5])
        c = np.union1d(c, edges[rng.random(len(edges)) < 0.3])
        brows.append(np.sort(c))
    brpt = np.concatenate([[0], np.cumsum([len(b) for b in brows])]).astype(np.int32)
    B = (brpt, np.concatenate(brows).astype(np.int32), rng.uniform(-1, 1, brpt[-1]), ncols)
    acols = [np.sort(rng.choice(K, size=s, replace=False)) for s in (3000, 1200, 700, 300, 40)]
    arpt = np.concatenate([[0], np.cumsum([len(a) for a in acols])]).astype(np.int32)
    A = (arpt, np.concatenate(acols).astype(np.int32), rng.uniform(-1, 1, arpt[-1]), K)
    R = oracle.brmerge_precise(A, B)
    from tests._util import csr_from_tuple
    Ah, Bh = csr_from_tuple(A, K), csr_from_tuple(B)
    for engine in (4, 0):
        gpu.set_param("group_cfg", engine)
        assert_csr_equal(gpu.brmerge_precise(Ah, Bh), R)
        assert_csr_equal(gpu.brmerge_upper(Ah, Bh), R)
    # list chunks of the symbolic count cut by product count (the cut that
    # keeps a chunk's offsets inside int), down to one list per chunk
    for pcap in (1000, 1):
        gpu.set_param("sym_chunk_products", pcap)
        assert_csr_equal(gpu.brmerge_precise(Ah, Bh), R)
