"""GPU parity of the batched entry points: every product of a batch must be
bit-identical to the oracle's result for that pair alone (and hence to a
single library call), for the one-shot batch and for every replay of a
graph-captured plan."""
import numpy as np
import pytest

import oracle
from tests._util import as_tuple, assert_csr_equal, rand_csr_tuple

pytestmark = pytest.mark.gpu


def _rand_pairs(seed, n):
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n):
        M, K, N = [int(x) for x in rng.integers(1, 400, 3)]
        if i == 3:
            M = 0  # a product with no rows
        if i == 5:
            K = 1  # inner dimension one
        A = rand_csr_tuple(rng, M, K, int(rng.integers(0, 70)), empty_frac=0.3)
        B = rand_csr_tuple(rng, K, N, int(rng.integers(0, 50)), empty_frac=0.3)
        pairs.append((A, B))
    return pairs


def _upload(bm, T, ncols):
    from tests._util import csr_from_tuple
    return bm.DeviceCSR.upload(csr_from_tuple(T, ncols))


def _check_products(prods, pairs):
    assert len(prods) == len(pairs)
    for P, (A, B) in zip(prods, pairs):
        assert_csr_equal(P, oracle.brmerge_precise(A, B, nthreads=4))


def test_batched_random_pairs(gpu):
    pairs = _rand_pairs(7, 12)
    dev = [(_upload(gpu, A, len(B[0]) - 1), _upload(gpu, B, B[3])) for A, B in pairs]
    C, row_off = gpu.spgemm_batched(dev)
    assert row_off[-1] == sum(len(A[0]) - 1 for A, _ in pairs)
    prods = gpu.split_products(C, row_off, [B[3] for _, B in pairs])
    _check_products(prods, pairs)
    C.free()
    for a, b in dev:
        a.free()
        b.free()


@pytest.mark.parametrize("force_long", [0, 1])
def test_batched_matches_single_calls(gpu, force_long):
    """Batched products equal one call per pair, bit for bit, including when
    every row takes the long-row tier."""
    old = gpu.get_param("force_long")
    gpu.set_param("force_long", force_long)
    try:
        pairs = [gpu.gen.batched_pair(i) for i in range(4)]
        small = []
        for A, B in pairs:  # keep the CPU oracle fast: first 600 rows of A
            m = min(A.M, 600)
            A = gpu.CSR(A.rpt[:m + 1].copy(), A.col[:A.rpt[m]].copy(), A.val[:A.rpt[m]].copy(), m, A.N)
            small.append((A, B))
        dev = [(gpu.DeviceCSR.upload(A), gpu.DeviceCSR.upload(B)) for A, B in small]
        C, row_off = gpu.spgemm_batched(dev)
        prods = gpu.split_products(C, row_off, [B.N for _, B in small])
        for P, (dA, dB), (A, B) in zip(prods, dev, small):
            S = gpu.spgemm_dev(dA, dB).download()
            assert_csr_equal(P, S)
            assert_csr_equal(P, oracle.brmerge_precise(as_tuple(A), as_tuple(B), nthreads=8))
        C.free()
        for a, b in dev:
            a.free()
            b.free()
    finally:
        gpu.set_param("force_long", old)


def test_batched_wide_columns(gpu):
    """Pairs whose B has up to 200,000 columns (more than 65,536, at most
    2^20) and A rows of 20-60 entries over B rows of 5-40: the fused product's
    register tier ranks with its two-level bitmap at every row size from 129
    to 2,048 products, and rows above take the LDS tier.  Each C_i equals
    its single call and the oracle, bit for bit."""
    pairs = []
    for i in range(3):
        N = [200000, 70000, 1 << 20][i]
        B = gpu.gen.er_var(3000, N, 5, 40, 40 + i)
        A = gpu.gen.er_var(400, 3000, 20, 60, 50 + i)
        pairs.append((A, B))
    dev = [(gpu.DeviceCSR.upload(A), gpu.DeviceCSR.upload(B)) for A, B in pairs]
    C, row_off = gpu.spgemm_batched(dev)
    prods = gpu.split_products(C, row_off, [B.N for _, B in pairs])
    for P, (dA, dB), (A, B) in zip(prods, dev, pairs):
        S = gpu.spgemm_dev(dA, dB).download()
        assert_csr_equal(P, S)
        assert_csr_equal(P, oracle.brmerge_precise(as_tuple(A), as_tuple(B), nthreads=8))
    C.free()
    for a, b in dev:
        a.free()
        b.free()


def test_batched_rejects_bad_pair(gpu):
    rng = np.random.default_rng(3)
    A = rand_csr_tuple(rng, 20, 10, 5)
    B = rand_csr_tuple(rng, 10, 8, 4)
    dA, dB = _upload(gpu, A, 10), _upload(gpu, B, 8)
    cols = (A[1] + 3).astype(np.int32)  # some columns reach past K = 10
    assert cols.max() >= 10
    bad_col = gpu.DeviceCSR(gpu.DeviceArray.from_host(A[0]), gpu.DeviceArray.from_host(cols),
                            gpu.DeviceArray.from_host(A[2]), 20, 10, len(cols))
    with pytest.raises(gpu.BrmError) as e:
        gpu.spgemm_batched([(dA, dB), (bad_col, dB)])
    assert e.value.status == 1
    bad_nnz = gpu.DeviceCSR(dA.rpt, dA.col, dA.val, 20, 10, dA.nnz - 1)
    with pytest.raises(gpu.BrmError):
        gpu.spgemm_batched([(bad_nnz, dB)])
    for x in (dA, dB, bad_col):
        x.free()


def test_batched_skewed_b_and_unsorted_row(gpu):
    """A pair whose B has a 3,000-entry first row and runs of empty rows,
    between random pairs: the fused batch matches the oracle per product; a
    descent inside that hub row (across a 512-entry chunk edge of the fused
    B's sortedness check) refuses the batch, and the library stays usable."""
    from tests.test_gpu_parity import _skewed_b
    rng = np.random.default_rng(8)
    Bs = _skewed_b(rng)
    As = rand_csr_tuple(rng, 200, len(Bs[0]) - 1, 6)
    pairs = _rand_pairs(21, 4)
    pairs.insert(2, (As, Bs))
    dev = [(_upload(gpu, A, len(B[0]) - 1), _upload(gpu, B, B[3])) for A, B in pairs]
    C, ro = gpu.spgemm_batched(dev)
    try:
        _check_products(gpu.split_products(C, ro, [B[3] for _, B in pairs]), pairs)
    finally:
        C.free()
    # the fused B holds the pairs' B entries in pair order: put the descent's
    # second entry first in a 512-entry chunk of the fused B
    k = (511 - sum(len(B[1]) for _, B in pairs[:2])) % 512
    col = Bs[1].copy()
    col[k], col[k + 1] = col[k + 1], col[k]
    bad = _upload(gpu, (Bs[0], col, Bs[2], Bs[3]), Bs[3])
    try:
        with pytest.raises(gpu.BrmError) as e:
            gpu.spgemm_batched(dev[:2] + [(dev[2][0], bad)] + dev[3:])
        assert e.value.status == 1
    finally:
        bad.free()
    C, ro = gpu.spgemm_batched(dev)
    try:
        _check_products(gpu.split_products(C, ro, [B[3] for _, B in pairs]), pairs)
    finally:
        C.free()
        for a, b in dev:
            a.free()
            b.free()


def test_batch_plan_graph_replays(gpu):
    """A plan's graph is replayed several times with other library work in
    between; each replay must reproduce the oracle, and value changes of
    the operands must show up in the next replay."""
    pairs = _rand_pairs(11, 9)
    dev = [(_upload(gpu, A, len(B[0]) - 1), _upload(gpu, B, B[3])) for A, B in pairs]
    plan = gpu.BatchPlan(dev, autotune=True)
    info = plan.info()
    assert info["pairs"] == 9 and len(info["tuning_ms"]) >= 2
    other_A = gpu.DeviceCSR.upload(gpu.gen.er(2000, 2000, 8, 5))
    for rep in range(3):
        plan.execute()
        _check_products(plan.products(), pairs)
        gpu.spgemm_dev(other_A, other_A).free()  # eager work between replays
    # new values, same pattern and addresses
    scaled = []
    for (A, B), (dA, dB) in zip(pairs, dev):
        A2 = (A[0], A[1], A[2] * -3.0, A[3])
        B2 = (B[0], B[1], B[2] + 0.5, B[3])
        if len(A2[2]):
            gpu._lib.check(gpu.lib().brm_memcpy_h2d(dA.val.ptr, A2[2].ctypes.data, A2[2].nbytes))
        if len(B2[2]):
            gpu._lib.check(gpu.lib().brm_memcpy_h2d(dB.val.ptr, B2[2].ctypes.data, B2[2].nbytes))
        scaled.append((A2, B2))
    plan.execute()
    _check_products(plan.products(), scaled)
    plan.execute()
    _check_products(plan.products(), scaled)
    plan.close()
    other_A.free()
    for a, b in dev:
        a.free()
        b.free()


def test_batch_plan_with_long_rows(gpu):
    """Long rows run eagerly after the graph; results still exact."""
    old = gpu.get_param("force_long")
    gpu.set_param("force_long", 1)
    try:
        pairs = _rand_pairs(21, 5)
        dev = [(_upload(gpu, A, len(B[0]) - 1), _upload(gpu, B, B[3])) for A, B in pairs]
        plan = gpu.BatchPlan(dev, autotune=False)
        assert plan.info()["rows_by_tier"][3] > 0
        for _ in range(2):
            plan.execute()
            _check_products(plan.products(), pairs)
        plan.close()
        for a, b in dev:
            a.free()
            b.free()
    finally:
        gpu.set_param("force_long", old)


def test_batch_plan_tune_cache(gpu, tmp_path):
    """The tuner's choice is written to the tune cache under the batch's
    shape key; a second plan over a batch of the same shape takes the shape
    from the file without timing candidates, and both are exact."""
    import json as _json
    path = tmp_path / "tune.json"
    pairs = _rand_pairs(31, 6)
    dev = [(_upload(gpu, A, len(B[0]) - 1), _upload(gpu, B, B[3])) for A, B in pairs]
    gpu.set_tune_cache(str(path))
    try:
        p1 = gpu.BatchPlan(dev, autotune=True)
        i1 = p1.info()
        assert i1["tune_source"] == "timed" and len(i1["tuning_ms"]) >= 2
        doc = _json.loads(path.read_text())
        ent = [e for e in doc["entries"] if e["key"] == i1["tune_key"]]
        assert len(ent) == 1
        assert (ent[0]["lds_cfg"], ent[0]["med_cap"], ent[0]["wmed_cap"]) == (i1["lds_cfg"], i1["med_cap"],
                                                                              i1["wmed_cap"])
        p1.execute()
        _check_products(p1.products(), pairs)
        p1.close()
        p2 = gpu.BatchPlan(dev, autotune=True)
        i2 = p2.info()
        assert i2["tune_source"] == "cache" and i2["tuning_ms"] == [] and i2["tune_key"] == i1["tune_key"]
        assert (i2["lds_cfg"], i2["med_cap"], i2["wmed_cap"]) == (i1["lds_cfg"], i1["med_cap"], i1["wmed_cap"])
        p2.execute()
        _check_products(p2.products(), pairs)
        p2.close()
        # a different entry for another key is kept when the file is rewritten
        doc["entries"].append({"key": "other", "lds_cfg": 0, "med_cap": 4096, "wmed_cap": 0, "ms": 1.0})
        path.write_text(_json.dumps(doc))
        gpu.set_tune_cache(str(path))
        p3 = gpu.BatchPlan(dev, autotune=True)
        assert p3.info()["tune_source"] == "cache"
        p3.close()
        assert {e["key"] for e in _json.loads(path.read_text())["entries"]} == {i1["tune_key"], "other"}
    finally:
        gpu.set_tune_cache(None)
        for a, b in dev:
            a.free()
            b.free()
