"""Host-side logic of bench.py (no GPU): the one-GPU refusal and the
per-kernel algorithmic byte model the roofline is computed from."""
import os
import subprocess
import sys

import numpy as np

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr and r.stdout == ""


def test_numeric_kernel_bytes_partition_the_product():
    """Every row is finished by exactly one numeric kernel, so their bytes
    add up to the whole-product model 20 nnzA + 12 ip + 12 nnz(C) + 12 per
    row, over the rows with products (empty rows are written by no kernel);
    the symbolic kernels read 4 B per product instead of 12 and write no C."""
    rng = np.random.default_rng(1)
    M = 5000
    L = rng.integers(0, 400, M)
    ip = np.where(L > 0, L * rng.integers(1, 40, M), 0)
    nnz = np.minimum(ip, rng.integers(1, 3000, M))
    A = None  # kernel_bytes reads only the per-row arrays
    for med_cap, wcap in ((4096, 2048), (2048, 512), (4096, 0)):
        t = bench.tiers(L, ip, med_cap, wcap)
        numeric = ("k_short_numeric", "k_medium", "k_rmed", "k_bgroup")
        total = sum(bench.kernel_bytes(k, A, L, ip, nnz, t) for k in numeric)
        live = ip > 0
        assert total == 20 * int(L[live].sum()) + 12 * int(ip.sum()) + 12 * int(nnz.sum()) + 12 * int(live.sum())
        sym = sum(bench.kernel_bytes(k, A, L, ip, nnz, t) for k in ("k_short_count", "k_sym_hash", "k_sym_bitmap",
                                                                     "k_sym_hash_w"))
        assert sym == 12 * int(L[live].sum()) + 4 * int(ip.sum()) + 12 * int(live.sum())
        # the fused short-row kernel is credited with the whole product
        assert bench.kernel_bytes("k_small", A, L, ip, nnz, t) == (20 * int(L.sum()) + 12 * int(ip.sum())
                                                                  + 12 * int(nnz.sum()) + 12 * M)
        assert bench.kernel_bytes("k_plan", A, L, ip, nnz, t) is None


def test_tiers_follow_the_bins():
    """Short rows (1-64 products) stay short whatever the caps; rows of at
    most 64 lists go to the wavefront tier up to its cap, the rest of the
    rows up to med_cap products and 512 lists to the LDS tier, everything
    else long; force_long sends every row with products long."""
    L = np.array([0, 3, 64, 64, 65, 512, 513, 10, 64])
    ip = np.array([0, 64, 65, 2048, 65, 4096, 4096, 4097, 2049])
    assert bench.tiers(L, ip, 4096, 2048).tolist() == [0, 1, 4, 4, 2, 2, 3, 3, 2]
    assert bench.tiers(L, ip, 2048, 0).tolist() == [0, 1, 2, 2, 2, 3, 3, 3, 3]
    assert bench.tiers(L, ip, 4096, 2048, force_long=True).tolist() == [0, 3, 3, 3, 3, 3, 3, 3, 3]
