"""GPU tests of the request server: each request's result must be the same
alone (one library call) and in the mix (fused with other requests in a
batch), bit for bit, and admission must respect the HBM budget."""
import ctypes

import numpy as np
import pytest

import oracle
from tests._util import as_tuple, assert_csr_equal

pytestmark = pytest.mark.gpu


def _requests(bm, Bs, n, seed, rows=(50, 1500)):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        b = int(rng.integers(0, len(Bs)))
        M = int(rng.integers(*rows))
        A = bm.gen.er_var(M, Bs[b].M, 1, 12, seed * 1000 + i)
        reqs.append((b, A))
    return reqs


def test_server_results_match_single_calls(gpu):
    Bs = [gpu.gen.rmat(10, seed=3), gpu.gen.er(700, 900, 6, 1), gpu.gen.banded(1024, 16, 2, 9, 4)]
    dBs = [gpu.DeviceCSR.upload(B) for B in Bs]
    reqs = _requests(gpu, Bs, 24, 5)
    dAs = [gpu.DeviceCSR.upload(A) for _, A in reqs]
    with gpu.Server(budget_bytes=8 << 30, max_batch=16, streams=2) as srv:
        ids = [srv.add_b(d) for d in dBs]
        assert ids == [0, 1, 2]
        rids = [srv.submit(b, dA) for (b, _), dA in zip(reqs, dAs)]
        sizes = []
        for k, rid in enumerate(rids):
            r = srv.wait(rid)
            sizes.append(r.batch_size)
            assert r.t_submit <= r.t_start <= r.t_done
            C = srv.result(rid)
            b, A = reqs[k]
            alone = gpu.spgemm_dev(dAs[k], dBs[b])
            assert_csr_equal(C, alone.download())
            alone.free()
            if k % 6 == 0:
                assert_csr_equal(C, oracle.brmerge_precise(as_tuple(A), as_tuple(Bs[b]), nthreads=8))
            srv.release(rid)
        st = srv.stats()
        assert st["completed"] == len(reqs)
        assert st["in_use_bytes"] == 0
        assert max(sizes) > 1, "no request was batched with another"
        assert st["p99_s"] >= st["p50_s"] > 0
    for d in dBs + dAs:
        d.free()


def test_server_wide_b_columns(gpu):
    """Resident B operands of up to 200,000 columns (the stacked B is wider
    than 65,536 columns: the register tier ranks with its two-level bitmap)
    and requests whose rows have 129 to about 2,400 products: every result
    equals its single call, and some equal the oracle."""
    Bs = [gpu.gen.er_var(2000, 200000, 5, 40, 61), gpu.gen.er_var(1500, 90000, 5, 40, 62)]
    dBs = [gpu.DeviceCSR.upload(B) for B in Bs]
    rng = np.random.default_rng(17)
    reqs = []
    for i in range(12):
        b = int(rng.integers(0, 2))
        reqs.append((b, gpu.gen.er_var(int(rng.integers(50, 400)), Bs[b].M, 20, 60, 700 + i)))
    dAs = [gpu.DeviceCSR.upload(A) for _, A in reqs]
    with gpu.Server(budget_bytes=8 << 30, max_batch=6, streams=2) as srv:
        ids = [srv.add_b(d) for d in dBs]
        rids = [srv.submit(ids[b], dA) for (b, _), dA in zip(reqs, dAs)]
        for k, rid in enumerate(rids):
            srv.wait(rid)
            C = srv.result(rid)
            b, A = reqs[k]
            alone = gpu.spgemm_dev(dAs[k], dBs[b])
            assert_csr_equal(C, alone.download())
            alone.free()
            if k % 4 == 0:
                assert_csr_equal(C, oracle.brmerge_precise(as_tuple(A), as_tuple(Bs[b]), nthreads=8))
            srv.release(rid)
    for d in dBs + dAs:
        d.free()


def test_server_admission_budget(gpu):
    B = gpu.gen.er(2000, 2000, 10, 7)
    dB = gpu.DeviceCSR.upload(B)
    reqs = _requests(gpu, [B], 10, 9, rows=(200, 400))
    dAs = [gpu.DeviceCSR.upload(A) for _, A in reqs]
    ips = [gpu.compute_flop(A, B) for _, A in reqs]
    need = [12 * ip + 8 * (A.M + 1) for ip, (_, A) in zip(ips, reqs)]
    budget = max(need) + 16
    with gpu.Server(budget_bytes=budget, max_batch=8, streams=2) as srv:
        b = srv.add_b(dB)
        # a request above the budget is refused at submission
        big = gpu.DeviceCSR.upload(gpu.gen.er(2000, 2000, 40, 1))
        with pytest.raises(gpu.BrmError) as e:
            srv.submit(b, big)
        assert e.value.status == 2
        rids = [srv.submit(b, dA) for dA in dAs]
        for rid, (_, A), ip in zip(rids, reqs, ips):
            r = srv.wait(rid)
            assert r.ip == ip
            C = srv.result(rid)
            assert_csr_equal(C, oracle.brmerge_precise(as_tuple(A), as_tuple(B), nthreads=8))
            srv.release(rid)
        assert srv.stats()["in_use_bytes"] == 0
        big.free()
    for d in [dB] + dAs:
        d.free()


def test_server_rejects_bad_requests(gpu):
    B = gpu.gen.er(100, 80, 5, 2)
    dB = gpu.DeviceCSR.upload(B)
    A = gpu.gen.er(30, 120, 4, 3)  # columns reach past B's 100 rows
    dA = gpu.DeviceCSR.upload(A)
    with gpu.Server(budget_bytes=1 << 30) as srv:
        b = srv.add_b(dB)
        with pytest.raises(gpu.BrmError) as e:
            srv.submit(b, dA)
        assert e.value.status == 1
        with pytest.raises(gpu.BrmError):
            srv.submit(b + 5, dA)
        ok = gpu.DeviceCSR.upload(gpu.gen.er(30, 100, 4, 3))
        # row pointers that do not end at the entry count given (one entry
        # fewer: every read stays inside the arrays)
        rid = ctypes.c_int64()
        st = gpu.lib().brm_server_submit(srv._h, b, ok.rpt.ptr, ok.col.ptr, ok.val.ptr, ok.nnz - 1, ok.M,
                                         ctypes.byref(rid))
        assert st == 1
        rid = srv.submit(b, ok)
        assert srv.result(rid).M == 30
        srv.release(rid)
        ok.free()
    dA.free()
    dB.free()


def test_server_small_batches_with_long_rows(gpu):
    """Many small batches (below 2^24 products, so the pipeline skips its
    read of nnz(C) after the symbolic phase) whose requests have long rows
    (their rows are sorted by product count before the long-row count), on
    two workers at once: every result equals its single call.  The sort's
    scratch must stay out of the allocator's cache until each call is done,
    since the other worker allocates meanwhile."""
    B = gpu.gen.rmat(12, seed=8)
    dB = gpu.DeviceCSR.upload(B)
    rng = np.random.default_rng(17)
    reqs = []
    for i in range(48):
        M = int(rng.integers(20, 120))
        # rows of 300 to 2,000 lists over B's rows: long rows of few products each
        lens = rng.integers(300, 2000, size=M)
        rpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        col = np.concatenate([np.sort(rng.choice(B.M, size=int(n), replace=False)) for n in lens]).astype(np.int32)
        reqs.append(gpu.CSR(rpt, col, rng.uniform(-1, 1, len(col)), M, B.M))
    dAs = [gpu.DeviceCSR.upload(A) for A in reqs]
    ips = [gpu.compute_flop(A, B) for A in reqs]
    assert max(ips) < (1 << 24) // 2
    with gpu.Server(budget_bytes=8 << 30, max_batch=2, streams=2) as srv:
        b = srv.add_b(dB)
        rids = [srv.submit(b, dA) for dA in dAs]
        for k, rid in enumerate(rids):
            srv.wait(rid)
            C = srv.result(rid)
            alone = gpu.spgemm_dev(dAs[k], dB)
            assert_csr_equal(C, alone.download())
            alone.free()
            srv.release(rid)
    for d in dAs + [dB]:
        d.free()
