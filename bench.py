"""Headline benchmark: C = A*A for R-MAT scale 20 (1,048,576 rows,
a/b/c/d = 0.57/0.19/0.19/0.05, 16 edges per row drawn, duplicates removed,
FP64 values U[-1,1], seed 0) with BRMerge-Precise on one MI355X.

A "step" is one full SpGEMM (validation, product counts, binning, symbolic,
row pointer, allocation of C, numeric merge) on operands already resident in
HBM; C stays in HBM and is released after the step (the caching allocator
keeps the memory).  Prints one JSON line:

  value      = 2 * intermediate_products * steps / elapsed   (GFLOP/s)
  roofline   = the dominant kernel's algorithmic bytes / its time, from
               hipEvents recorded around every launch during the timed steps
  cpu_baseline = the CPU oracle (oracle/, an OpenMP restatement of the
               reference's BRMerge-Precise) on a bounded row panel of A
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def row_stats(A, B):
    """Per-row list count L and product count ip (numpy)."""
    blen = np.diff(np.asarray(B.rpt, np.int64))
    alen = blen[A.col]
    cs = np.concatenate([[0], np.cumsum(alen)])
    rpt = np.asarray(A.rpt, np.int64)
    ip = cs[rpt[1:]] - cs[rpt[:-1]]
    L = np.diff(rpt)
    return L, ip


def tiers(L, ip, med_cap, wmed_cap, force_long=False):
    """Row tiers as the library bins them: 0 empty, 1 short, 2 medium
    (workgroup LDS), 3 long, 4 medium (one wavefront per row: the register
    tier, or the wavefront LDS engine)."""
    t = np.full(len(ip), 3, np.int8)
    t[ip == 0] = 0
    if not force_long:
        t[(ip > 64) & (ip <= med_cap) & (L <= 512)] = 2
        t[(ip > 64) & (ip <= wmed_cap) & (L <= 64)] = 4
        t[(ip > 0) & (ip <= 64)] = 1
    return t


def bin_caps(bm, N):
    """(med_cap, wmed_cap) as the library's binning resolves them for B of N
    columns (src api.hip bin_launch): the LDS tier's cap is clipped by its
    shape (4,096 products for lds_cfg 0, 2,048 otherwise); the register
    tier takes rows of up to 2,048 products where its bitmap ranks run (B
    of at most 2^20 columns), else 1,024 while columns fit 16 bits, else
    512; the LDS wavefront engine stops at 512 by default, 1,024 at most."""
    med_cap = min(bm.get_param("med_cap"), 4096 if bm.get_param("lds_cfg") == 0 else 2048)
    reg = bm.get_param("wmed_engine") == 1
    bitmap = reg and bm.get_param("wmed_bitmap") == 1 and N <= (1 << 20)
    wcap = bm.get_param("wmed_cap")
    if wcap < 0:
        wcap = (2048 if bitmap else 1024 if N <= 65536 else 512) if reg else 512
    return med_cap, min(wcap, 2048 if bitmap else 1024)


def workload_label(w):
    """The product a workload name stands for, as the metric names it."""
    if w.startswith("rmat"):
        return "R-MAT A*A"
    if w == "er16k":
        return "Erdos-Renyi 16384^2 A*A"
    if w.startswith("amg"):
        return "AMG stencil*aggregation A*P"
    return w


def kernel_bytes(name, A, L, ip, nnz_row, tier):
    """Algorithmic bytes of all launches of kernel `name` in one SpGEMM, by
    the per-product model bytes_alg = 4(M+1) + 20 nnzA + 12 ip + 8(M+1) +
    12 nnz(C) restricted to the rows the kernel finishes: per row, its A
    entries (col 4 + val 8 + product offset 8 = 20 B each), one B gather per
    product (col 4 + val 8 = 12 B; the symbolic kernels read columns only,
    4 B), C written once (12 B per nnz) and its row pointers (4 B in, 8 B
    out).  Traffic between tree levels (LDS ping-pong, the long-row engine's
    workspace passes) is never counted: it is overhead, reported apart."""
    sel = {"k_short_numeric": 1, "k_short_count": 1, "k_medium": 2, "k_sym_hash": 2, "k_sym_bitmap": 3,
           "k_medium_w": 4, "k_rmed": 4, "k_sym_hash_w": 4,
           "k_bgroup": 3, "k_group": 3, "k_leaf": 3}
    if name == "k_small":  # the fused short-row path finishes every row
        return 20 * int(L.sum()) + 12 * int(ip.sum()) + 12 * int(nnz_row.sum()) + 12 * len(L)
    if name not in sel:
        return None
    m = tier == sel[name]
    nA, nip, nc, nr = int(L[m].sum()), int(ip[m].sum()), int(nnz_row[m].sum()), int(m.sum())
    if name in ("k_short_count", "k_sym_hash", "k_sym_bitmap", "k_sym_hash_w"):
        return 12 * nA + 4 * nip + 12 * nr
    return 20 * nA + 12 * nip + 12 * nc + 12 * nr


def engine_bytes(name, st):
    """Bytes the long-row merge kernels move by their own accounting: every
    element a pass reads (12 B) and writes (12 B), plus 48 B per node
    descriptor.  Minus the algorithmic bytes of the long rows, this is the
    workspace traffic of the intermediate tree levels."""
    if name == "k_leaf":
        return 12 * st["leaf_ip"] + 12 * st["leaf_out"] + 20 * st["leaf_tasks"]
    if name in ("k_group", "k_bgroup"):
        return 12 * st["group_in"] + 12 * st["group_out"] + 48 * st["group_tasks"]
    return None


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f).get(f"{kernel}|{workload}")
    except (OSError, ValueError):
        return None
    if not t:
        return None
    return int(t["read_bytes_per_launch"] + t["write_bytes_per_launch"])


def host_cpu():
    """The host's CPU model and socket count (lscpu, else /proc/cpuinfo)."""
    model, sockets = None, None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name" and model is None:
                model = v.strip()
            elif k.strip() == "Socket(s)":
                try:
                    sockets = int(v.strip())
                except ValueError:
                    pass
    except (OSError, ValueError):
        pass
    if model is None or sockets is None:
        try:
            ids = set()
            with open("/proc/cpuinfo") as f:
                for line in f:
                    k, _, v = line.partition(":")
                    if k.strip() == "model name" and model is None:
                        model = v.strip()
                    elif k.strip() == "physical id":
                        ids.add(v.strip())
            if sockets is None and ids:
                sockets = len(ids)
        except OSError:
            pass
    return model, sockets


PANEL_ROWS = 32768
PANEL_SEED = 12345


def panel(A, ip):
    """Rows the CPU baseline runs and the GPU result is checked on: all rows
    when the product has at most 2e8 products, else a seeded uniform random
    panel of PANEL_ROWS rows."""
    if int(ip.sum()) <= 200_000_000:
        return np.arange(A.M), "the whole product"
    rng = np.random.default_rng(PANEL_SEED)
    rows = np.sort(rng.choice(A.M, size=min(PANEL_ROWS, A.M), replace=False))
    return rows, f"a uniform random panel of {len(rows)} rows (numpy default_rng({PANEL_SEED}))"


def cpu_baseline(A, rows, what, calls=10):
    """Time the oracle's BRMerge-Precise (the OpenMP restatement of
    spgemm.h:468-735) on the host's cores beside the GPU run, with the
    reference driver's protocol (brmerge_precise.cpp:67-83): 1 warm-up call,
    then the mean of `calls` calls, each timed around the whole call
    (output allocation included).  The whole product runs when it has at
    most 2e8 products; otherwise a seeded uniform random panel of
    `panel_rows` rows of A times A (the reference's int32 row pointer and
    host memory cannot hold R-MAT-20's full C).  Threads: every CPU this
    process may run on, capped by OMP_NUM_THREADS when that is set."""
    visible = len(os.sched_getaffinity(0))
    env_thr = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(visible, int(env_thr)) if env_thr.isdigit() and int(env_thr) > 0 else visible
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle

    L, ip = row_stats(A, A)
    rpt = np.asarray(A.rpt, np.int64)
    lens = rpt[rows + 1] - rpt[rows]
    prpt = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = (np.concatenate([np.arange(rpt[r], rpt[r + 1]) for r in rows]) if len(rows)
           else np.zeros(0, np.int64))
    P = (prpt, A.col[idx], A.val[idx], A.N)
    B = (np.asarray(A.rpt, np.int32), A.col, A.val, A.N)
    ip_panel = int(ip[rows].sum())
    oracle.brmerge_precise(P, B, nthreads=threads)  # warm-up
    ts = []
    for _ in range(calls):
        t0 = time.perf_counter()
        C = oracle.brmerge_precise(P, B, nthreads=threads)
        ts.append(time.perf_counter() - t0)
    t = float(np.mean(ts))
    model, sockets = host_cpu()
    return {
        "value": round(2.0 * ip_panel / t / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
        "sample": (f"oracle BRMerge-Precise (OpenMP C restatement of spgemm.h:468-735) on {what} of the same A "
                   f"times A: {len(rows)} rows, {ip_panel} products, nnz(C) {C.nnz}; "
                   f"1 warm-up + mean of {calls} calls"),
        "seconds_per_call": round(t, 4), "calls": calls, "panel_rows": int(len(rows)), "panel_products": ip_panel,
        "visible_cpus": visible, "os_cpu_count": os.cpu_count(), "cpu_model": model, "sockets": sockets,
        "omp": {"OMP_NUM_THREADS": env_thr or None, "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
                "OMP_PLACES": os.environ.get("OMP_PLACES")},
    }, C


def download_panel(C, rows):
    """(row pointer diffs, columns, values) of `rows` of a DeviceCSR result."""
    rpt = C.rpt.to_host()
    lens = rpt[rows + 1] - rpt[rows]
    cols = [C.col.to_host(int(rpt[r]), int(rpt[r + 1])) for r in rows]
    vals = [C.val.to_host(int(rpt[r]), int(rpt[r + 1])) for r in rows]
    return (lens, np.concatenate(cols) if cols else np.zeros(0, np.int32),
            np.concatenate(vals) if vals else np.zeros(0))


def panel_matches(got, R):
    """Bit-for-bit equality of a downloaded panel with the oracle's C."""
    lens, col, val = got
    return (np.array_equal(lens, np.diff(R.rpt)) and np.array_equal(col, R.col)
            and np.array_equal(val.view(np.uint64), R.val.view(np.uint64)))


def _stats_ip(A, B):
    return int(row_stats(A, B)[1].sum())


def run_batched(args, bm):
    """Config: 4,096 independent products C_i = A_i * B_i (square, rows
    log-uniform in [256, 65536], 4-32 nnz per row, even i Erdos-Renyi, odd i
    banded with half-width 64; values U[-1,1]; A seed 2i, B seed 2i+1),
    one fused launch sequence per step, against one call per product."""
    n = args.pairs
    t0 = time.perf_counter()
    host = [bm.gen.batched_pair(i) for i in range(n)]
    tgen = time.perf_counter() - t0
    ip = sum(_stats_ip(A, B) for A, B in host)
    rows = sum(A.M for A, _ in host)
    dev = [(bm.DeviceCSR.upload(A), bm.DeviceCSR.upload(B)) for A, B in host]
    nnz_in = sum(A.nnz + B.nnz for A, B in host)
    del host
    print(f"[bench] batched: {n} pairs, {rows} rows, {nnz_in} input nnz, ip={ip} (generated in {tgen:.1f}s)",
          file=sys.stderr, flush=True)
    # one call per product on a sample (the same distribution)
    ns = min(n, args.loop_sample)
    for i in range(ns):  # untimed pass: the allocator holds every size class the sample needs
        C = bm.spgemm_dev(*dev[i])
        C.free()
    t0 = time.perf_counter()
    for i in range(ns):
        C = bm.spgemm_dev(*dev[i])
        C.free()
    loop_rate = ns / (time.perf_counter() - t0)
    print(f"[bench] one call per product: {loop_rate:.1f} products/s over {ns}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        C, ro = bm.spgemm_batched(dev)
        C.free()
    bm.set_param("profile", 0 if args.no_profile else 1)
    bm.prof_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        C, ro = bm.spgemm_batched(dev)
        nnz_c = C.nnz
        C.free()
    el = (time.perf_counter() - t0) / args.steps
    kernels = {k: {"launches": n_, "ms": round(ms / args.steps, 3)} for k, (n_, ms) in bm.prof_report().items()}
    bm.set_param("profile", 0)
    st = bm.last_stats()
    print(f"[bench] fused batch: {el * 1e3:.1f} ms per batch, nnz(C)={nnz_c}", file=sys.stderr, flush=True)
    plan_out = None
    if not args.no_plan:
        # the tuner's choice goes to a tune cache in a scratch directory; a
        # second plan over the same batch then takes its shape from the file
        import tempfile
        tune_dir = tempfile.mkdtemp(prefix="brm_tune_")
        tune_file = os.path.join(tune_dir, "tune.json")
        bm.set_tune_cache(tune_file)
        t0 = time.perf_counter()
        plan = bm.BatchPlan(dev, autotune=True)
        tcreate = time.perf_counter() - t0
        plan.execute()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.execute()
        pel = (time.perf_counter() - t0) / args.steps
        info = plan.info()
        plan.close()
        t0 = time.perf_counter()
        plan = bm.BatchPlan(dev, autotune=True)
        tcreate_cached = time.perf_counter() - t0
        info_cached = plan.info()
        plan.close()
        bm.set_tune_cache(None)
        os.remove(tune_file)
        os.rmdir(tune_dir)
        # like for like: a loop of replays of one-product plans (the same
        # sample of products as the one-call loop), each plan built once
        singles = [bm.BatchPlan([dev[i]], autotune=False) for i in range(ns)]
        for sp in singles:
            sp.execute()
        t0 = time.perf_counter()
        for sp in singles:
            sp.execute()
        single_rate = ns / (time.perf_counter() - t0)
        for sp in singles:
            sp.close()
        plan_out = {"products_per_s": round(n / pel, 2), "ms_per_execute": round(pel * 1e3, 3),
                    "create_s": round(tcreate, 3),
                    "create_s_from_tune_cache": round(tcreate_cached, 3),
                    "tune_key": info["tune_key"], "tune_source_second_plan": info_cached["tune_source"],
                    "gflops": round(2.0 * ip / pel / 1e9, 3),
                    "single_product_plans": {"products_per_s": round(single_rate, 2), "sample": ns},
                    "speedup_vs_single_plans": round(n / pel / single_rate, 2),
                    "tuned": {"lds_cfg": info["lds_cfg"], "med_cap": info["med_cap"], "wmed_cap": info["wmed_cap"],
                              "tuning_ms": info["tuning_ms"]}}
    for a, b in dev:
        a.free()
        b.free()
    rate = n / el
    return {
        "metric": f"Batched SpGEMM products/s ({n} independent pairs, BRMerge-Precise)",
        "value": round(rate, 2), "unit": "products/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched{n}", "pairs": n, "rows": rows, "input_nnz": nnz_in,
                   "intermediate_products": ip, "nnz_c": int(nnz_c), "inputs": "resident in HBM"},
        "gflops": round(2.0 * ip / el / 1e9, 3),
        "one_call_per_product": {"products_per_s": round(loop_rate, 2), "sample": ns},
        "speedup_vs_loop": round(rate / loop_rate, 2),
        "plan_speedup_vs_single_plans": plan_out["speedup_vs_single_plans"] if plan_out else None,
        "plan": plan_out, "kernels_per_step": kernels, "stats": st,
    }


def run_serve(args, bm):
    """Config: 8 resident B operands (R-MAT scale 18, seeds 0-7); requests
    A with round(lognormal(ln 4096, 1.5)) rows, 8-24 uniform columns per row
    over B's rows; Poisson arrivals; admission by the Upper bound against a
    64 GiB budget.  Reports p50/p99 latency at the offered rate, saturated
    throughput, and one call per request for comparison."""
    scale = 18
    Bs = [bm.gen.rmat(scale, seed=s) for s in range(8)]
    K = 1 << scale
    dBs = [bm.DeviceCSR.upload(B) for B in Bs]
    rng = np.random.default_rng(0)
    n = args.requests
    Ms = np.clip(np.round(np.exp(rng.normal(np.log(4096), 1.5, n))), 1, 1 << 20).astype(int)
    bsel = rng.integers(0, len(Bs), n)
    As = [bm.gen.er_var(int(M), K, 8, 24, 7919 + i) for i, M in enumerate(Ms)]
    ips = [_stats_ip(A, Bs[b]) for A, b in zip(As, bsel)]
    dAs = [bm.DeviceCSR.upload(A) for A in As]
    del As
    ip_tot = int(sum(ips))
    print(f"[bench] serve: {n} requests, median rows {int(np.median(Ms))}, mean ip {ip_tot / n:.3g}",
          file=sys.stderr, flush=True)
    # one call per request, sequential; an untimed pass first, so that the
    # allocator holds every size class the requests need (as the server's
    # warm-up drive does for the server)
    for i in range(n):
        C = bm.spgemm_dev(dAs[i], dBs[bsel[i]])
        C.free()
    t0 = time.perf_counter()
    for i in range(n):
        C = bm.spgemm_dev(dAs[i], dBs[bsel[i]])
        C.free()
    loop_s = time.perf_counter() - t0
    loop_rate = n / loop_s
    out = {}
    with bm.Server(budget_bytes=64 << 30, max_batch=args.max_batch, streams=args.streams) as srv:
        ids = [srv.add_b(d) for d in dBs]

        def drive(rate):
            """Submits all requests (Poisson at `rate`/s, or all at once if
            rate <= 0), waits for all; returns (span_s, latencies, batch sizes)."""
            gaps = rng.exponential(1.0 / rate, n) if rate > 0 else np.zeros(n)
            t_start = time.perf_counter()
            t_next = t_start
            rids = []
            for i in range(n):
                t_next += gaps[i]
                while True:
                    d = t_next - time.perf_counter()
                    if d <= 0:
                        break
                    time.sleep(min(d, 0.002))
                rids.append(srv.submit(ids[bsel[i]], dAs[i]))
            lat, bsz = [], []
            for r in rids:
                res = srv.wait(r)
                lat.append(res.t_done - res.t_submit)
                bsz.append(res.batch_size)
                srv.release(r)
            return time.perf_counter() - t_start, np.array(lat), np.array(bsz)

        drive(0)  # warm-up
        span, lat, bsz = drive(0)
        sat_rate = n / span
        offered = args.rate if args.rate > 0 else 0.5 * sat_rate
        span2, lat2, bsz2 = drive(offered)
    for d in dAs + dBs:
        d.free()
    return {
        "metric": "Served SpGEMM requests/s (8 resident R-MAT-18 B operands, continuous batching)",
        "value": round(sat_rate, 2), "unit": "requests/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(span * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "serve", "requests": n, "resident_b": len(Bs), "b": f"rmat{scale}",
                   "median_rows": int(np.median(Ms)), "intermediate_products": ip_tot,
                   "budget_bytes": 64 << 30, "max_batch": args.max_batch, "streams": args.streams},
        "gflops_saturated": round(2.0 * ip_tot / span / 1e9, 3),
        "saturated": {"requests_per_s": round(sat_rate, 2), "p50_ms": round(1e3 * float(np.median(lat)), 3),
                      "p99_ms": round(1e3 * float(np.quantile(lat, 0.99)), 3),
                      "mean_batch": round(float(bsz.mean()), 2)},
        "poisson": {"offered_per_s": round(offered, 2), "achieved_per_s": round(n / span2, 2),
                    "p50_ms": round(1e3 * float(np.median(lat2)), 3),
                    "p99_ms": round(1e3 * float(np.quantile(lat2, 0.99)), 3),
                    "mean_batch": round(float(bsz2.mean()), 2)},
        "one_call_per_request": {"requests_per_s": round(loop_rate, 2)},
        "speedup_vs_loop": round(sat_rate / loop_rate, 2),
    }


def run_amg(args, bm, n):
    """Config: A = 27-point stencil on an n^3 grid (26 diagonal, -1 off),
    P = 3x3x3 aggregation (one 1.0 per row); one step computes A*P and then
    P^T * (A*P) on the device (P^T formed once, outside the timing)."""
    S, P = bm.gen.stencil27(n), bm.gen.aggregation(n)
    PT = P.transpose()
    dS, dP, dPT = bm.DeviceCSR.upload(S), bm.DeviceCSR.upload(P), bm.DeviceCSR.upload(PT)
    ip1 = _stats_ip(S, P)

    def step():
        AP = bm.spgemm_dev(dS, dP)
        APo = AP.as_operand()
        C = bm.spgemm_dev(dPT, APo)
        return AP, APo, C

    AP, APo, C = step()
    APh = AP.download()
    ip2 = _stats_ip(PT, bm.CSR(np.asarray(APh.rpt, np.int32), APh.col, APh.val, APh.M, APh.N))
    nnz = (AP.nnz, C.nnz)
    for x in (AP, C):
        x.free()
    APo.rpt.free()
    for _ in range(args.warmup):
        AP, APo, C = step()
        AP.free(); C.free(); APo.rpt.free()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        AP, APo, C = step()
        AP.free(); C.free(); APo.rpt.free()
    el = (time.perf_counter() - t0) / args.steps
    for d in (dS, dP, dPT):
        d.free()
    return {
        "metric": "SpGEMM GFLOP/s (2*intermediate products/s), AMG Galerkin A*P then P^T*(A*P), BRMerge-Precise",
        "value": round(2.0 * (ip1 + ip2) / el / 1e9, 3), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"amg{n}", "grid": n, "rows": S.M, "coarse": P.N, "nnz_a": S.nnz,
                   "intermediate_products": [ip1, ip2], "nnz_c": list(nnz)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat20",
                    help="rmatS | er16k | amgN (stencil*aggregation) | galerkinN (A*P then P^T*(AP)) | batched | serve")
    ap.add_argument("--pairs", type=int, default=4096, help="batched: number of products")
    ap.add_argument("--loop-sample", type=int, default=1024, help="batched: products timed one call each")
    ap.add_argument("--no-plan", action="store_true", help="batched: skip the graph-captured plan")
    ap.add_argument("--requests", type=int, default=256, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=0.0, help="serve: offered requests/s (0: half of saturation)")
    ap.add_argument("--max-batch", type=int, default=64, help="serve: requests fused per batch")
    ap.add_argument("--streams", type=int, default=2, help="serve: worker streams")
    ap.add_argument("--no-upper", action="store_true", help="do not time BRMerge-Upper")
    ap.add_argument("--vendor", action="store_true",
                    help="also time rocSPARSE SpGEMM on the same product (int32 indices: nnz(C) < 2^31)")
    ap.add_argument("--hash", action="store_true", help="also time the Hash-SpGEMM comparator tier (hash_cap 1024)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-profile", action="store_true", help="do not bracket launches with hipEvents")
    ap.add_argument("--param", action="append", default=[], metavar="NAME=VALUE",
                    help="library tuning knob (brm_set_param), repeatable")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run the library's plain copy kernel over GIB GiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): the product would shard by rows "
              f"of C across GPUs, but only the one-GPU path is built", file=sys.stderr)
        sys.exit(2)

    import brmerge_amd as bm
    if bm.lib().brm_init(0) != 0:
        raise RuntimeError(bm._lib.last_error())
    if args.copy_calibration > 0:
        nbytes = int(args.copy_calibration * (1 << 30)) // 16 * 16
        src = bm.DeviceArray.alloc(nbytes // 8, np.float64)
        dst = bm.DeviceArray.alloc(nbytes // 8, np.float64)
        for _ in range(2):
            bm._lib.check(bm.lib().brm_bench_copy(dst.ptr, src.ptr, nbytes, None))
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 2}), flush=True)
        return
    params = {}
    for kv in args.param:
        k, v = kv.split("=", 1)
        bm.set_param(k, int(v))
        params[k] = int(v)

    if args.workload == "batched":
        print(json.dumps(run_batched(args, bm)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(args, bm)), flush=True)
        return
    if args.workload.startswith("galerkin"):
        print(json.dumps(run_amg(args, bm, int(args.workload[8:]))), flush=True)
        return

    t_gen = time.perf_counter()
    A, B = bm.gen.workload(args.workload)
    t_gen = time.perf_counter() - t_gen
    L, ip = row_stats(A, B)
    total_ip = int(ip.sum())
    dA = bm.DeviceCSR.upload(A)
    dB = dA if B is A else bm.DeviceCSR.upload(B)
    rows_chk, rows_what = panel(A, ip)

    print(f"[bench] {args.workload}: M={A.M} nnzA={A.nnz} ip={total_ip} (generated in {t_gen:.1f}s)",
          file=sys.stderr, flush=True)
    for w in range(args.warmup):
        tw = time.perf_counter()
        C = bm.spgemm_dev(dA, dB)
        C.free()
        print(f"[bench] warmup {w}: {time.perf_counter() - tw:.3f}s", file=sys.stderr, flush=True)

    bm.set_param("profile", 0 if args.no_profile else 1)
    bm.prof_reset()
    # The phase clock runs on the last step only, and progress is printed at
    # most once a second: both cost host time that a sub-millisecond product
    # would otherwise be charged with.
    t = bm.Precise_Timing()
    t0 = time.perf_counter()
    t_print = t0
    for _ in range(args.steps):
        lst = _ == args.steps - 1
        C = bm.spgemm_dev(dA, dB, timing=t if lst else None)
        nnz_c = C.nnz
        now = time.perf_counter()
        if lst or now - t_print > 1.0:
            print(f"[bench] step {_}: {now - t0:.3f}s since start, nnz(C)={nnz_c}", file=sys.stderr, flush=True)
            t_print = now
        last = C
        if not lst:
            C.free()
    elapsed = time.perf_counter() - t0
    timings = [t.as_dict()]
    prof = bm.prof_report()
    st = bm.last_stats()
    bm.set_param("profile", 0)
    crpt = last.rpt.to_host()
    got_precise = download_panel(last, rows_chk) if B is A else None
    last.free()
    nnz_row = np.diff(crpt)

    med_cap, wcap = bin_caps(bm, B.N)
    tier = tiers(L, ip, med_cap, wcap, force_long=bm.get_param("force_long") != 0)
    # the library's own row counts: empty, short, medium (LDS and register), long
    lib_rows = [int(x) for x in st["rows_by_tier"]]
    if lib_rows == [0, A.M, 0, 0]:  # the fused short-row path counts every row short
        tiers_match = bool(((tier == 0) | (tier == 1)).all())
    else:
        tiers_match = lib_rows == [int((tier == 0).sum()), int((tier == 1).sum()),
                                   int(((tier == 2) | (tier == 4)).sum()), int((tier == 3).sum())]
    if not tiers_match:
        print(f"[bench] row tiers differ from the library's {lib_rows}: kernel bytes are approximate",
              file=sys.stderr, flush=True)
    ms_per_step = 1000.0 * elapsed / args.steps
    gflops = 2.0 * total_ip * args.steps / elapsed / 1e9
    bytes_alg = 4 * (A.M + 1) + 20 * A.nnz + 12 * total_ip + 8 * (A.M + 1) + 12 * nnz_c
    kernels = {}
    for k, (n, ms) in prof.items():
        kernels[k] = {"launches": n, "ms": round(ms, 3)}
    dom = max(prof.items(), key=lambda kv: kv[1][1])[0] if prof else None
    roof = None
    if dom is not None:
        n, ms = prof[dom]
        by = kernel_bytes(dom, A, L, ip, nnz_row, tier)
        if by is not None:
            by_per_launch = by * args.steps / n
            avg_s = ms / n / 1e3
            ach = by_per_launch / avg_s / 1e9
            traffic = measured_traffic(dom, args.workload)
            roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "avg_launch_ms": round(ms / n, 4), "alg_bytes_per_launch": int(by_per_launch),
                    "launches_per_step": round(n / args.steps, 2),
                    "traffic_over_alg": round(traffic / by_per_launch, 3) if traffic else None,
                    "tiers_match_library": tiers_match}
            eb = engine_bytes(dom, st)
            if eb is not None:
                roof["engine_bytes_per_launch"] = int(eb / (n / args.steps))
                roof["workspace_bytes_per_step"] = int(eb - by)

    out = {
        "metric": "SpGEMM GFLOP/s (2*intermediate products/s), BRMerge-Precise, " + workload_label(args.workload),
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} C=A*A" if B is A else args.workload,
                   "rows": A.M, "nnz_a": A.nnz, "intermediate_products": total_ip, "nnz_c": int(nnz_c),
                   "variant": "precise", "inputs": "resident in HBM", "generator_s": round(t_gen, 2),
                   "params": params},
        "hbm_gbs_algorithmic": round(bytes_alg / (elapsed / args.steps) / 1e9, 2),
        "hbm_frac_algorithmic": round(bytes_alg / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roof,
        "phases_ms": {k: round(1000 * v, 3) for k, v in timings[-1].items()},
        "kernels": kernels,
        "stats": st,
    }
    got_upper = None
    if not args.no_upper:
        C = bm.spgemm_dev(dA, dB, upper=True)  # warm-up
        C.free()
        ts, tph = [], []
        for _ in range(max(1, args.steps)):
            t = bm.Upper_Timing()
            t0 = time.perf_counter()
            C = bm.spgemm_dev(dA, dB, upper=True, timing=t)
            ts.append(time.perf_counter() - t0)
            tph.append(t.as_dict())
            print(f"[bench] upper step {_}: {ts[-1]:.3f}s nnz(C)={C.nnz}", file=sys.stderr, flush=True)
            if _ == max(1, args.steps) - 1 and B is A:
                got_upper = download_panel(C, rows_chk)
            C.free()
        out["upper"] = {"value": round(2.0 * total_ip / np.mean(ts) / 1e9, 3), "unit": "GFLOP/s",
                        "ms_per_step": round(1000 * float(np.mean(ts)), 3), "steps": len(ts), "warmup": 1,
                        "vs_precise": round(float(np.mean(ts)) / (elapsed / args.steps), 3),
                        "phases_ms": {k: round(1000 * float(np.mean([p[k] for p in tph])), 3) for k in tph[0]}}
    if args.hash:
        # the Hash-SpGEMM comparator tier on the rows it takes (at most 1,024
        # products and lists), BRMerge's tiers on the others
        bm.set_param("hash_cap", 1024)
        C = bm.spgemm_dev(dA, dB)
        C.free()
        ts = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            C = bm.spgemm_dev(dA, dB)
            ts.append(time.perf_counter() - t0)
            C.free()
        bm.set_param("hash_cap", 0)
        out["hash_tier"] = {"value": round(2.0 * total_ip / np.mean(ts) / 1e9, 3), "unit": "GFLOP/s",
                            "ms_per_step": round(1000 * float(np.mean(ts)), 3),
                            "rows_in_tier": int(((ip > 0) & (ip <= 1024) & (L <= 1024)).sum()),
                            "vs_precise": round(float(np.mean(ts)) / (elapsed / args.steps), 3)}
    if args.vendor:
        if nnz_c >= 2 ** 31:
            out["rocsparse"] = {"skipped": "nnz(C) exceeds the int32 indices of the rocSPARSE path"}
        else:
            from tests import _rocsparse
            _rocsparse.spgemm(bm, dA, dB)  # warm-up
            ts, tc = [], []
            for _ in range(max(1, args.steps)):
                _, t_all, t_comp = _rocsparse.spgemm(bm, dA, dB)
                ts.append(t_all)
                tc.append(t_comp)
            out["rocsparse"] = {"gflops_total": round(2.0 * total_ip / float(np.mean(ts)) / 1e9, 3),
                                "gflops_compute_stage": round(2.0 * total_ip / float(np.mean(tc)) / 1e9, 3),
                                "ms_total": round(1e3 * float(np.mean(ts)), 3),
                                "ms_compute_stage": round(1e3 * float(np.mean(tc)), 3)}
    dA.free()
    if dB is not dA:
        dB.free()
    if not args.no_cpu and B is A:
        out["cpu_baseline"], R = cpu_baseline(A, rows_chk, rows_what)
        # The reference's driver checks every run's C (brmerge_precise.cpp:86-93); here the
        # GPU's C is compared bit for bit with the oracle's on the baseline's rows.
        out["verified"] = {"rows": rows_what, "products": int(ip[rows_chk].sum()),
                           "precise_bit_identical": bool(panel_matches(got_precise, R)),
                           "upper_bit_identical": (bool(panel_matches(got_upper, R))
                                                   if got_upper is not None else None)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
