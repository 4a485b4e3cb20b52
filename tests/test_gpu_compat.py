"""The drop-in C++ layer at run time.

tests/compat/compat_caller.cpp is written against the reference's own C++
interface (CSR, Precise_Timing, Upper_Timing, both overloads of
brmerge_precise / brmerge_upper, compute_flop: BRMerge-SpGEMM/inc/spgemm.h
:59-66, :220-223, :460-471, :737-741; CSR.cpp:102-253; Timings.cpp:4-165)
and compiled against include/brmerge_compat.hpp.  Its C must equal the
oracle's bit for bit, its phase tables must follow the reference's layout
(sum_total excludes cleanup for Precise, Timings.cpp:135, and includes every
phase for Upper, Timings.cpp:54), and brm_driver (the reference driver's
protocol, brmerge_precise.cpp:38-93) must report Precise and Upper equal.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from tests._util import as_tuple, assert_csr_equal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "compat", "compat_caller")
DRIVER = os.path.join(ROOT, "brmerge_amd", "brm_driver")


def phase_table(text, header):
    """{phase: ms} of the 'time(ms):' block after `header`."""
    block = text.split(header, 1)[1].split("perf(GFLOPS):", 1)[0]
    return {m.group(1): float(m.group(2)) for m in re.finditer(r"^\s+(\w+)\s+(-?[\d.]+)ms", block, re.M)}


@pytest.mark.parametrize("case", ["rmat10_rect", "amg9"])
def test_compat_caller(gpu, tmp_path, case):
    if case == "rmat10_rect":
        A = gpu.gen.rmat(10, seed=2)
        B = gpu.gen.er(1024, 700, 6, 3)
    else:
        A, B = gpu.gen.stencil27(9), gpu.gen.aggregation(9)
    a, b, prefix = tmp_path / "a.mtx", tmp_path / "b.mtx", str(tmp_path / "c")
    A.write_mtx(a)
    B.write_mtx(b)
    p = subprocess.run([CALLER, str(a), str(b), prefix], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    out = p.stdout
    flop = int(re.search(r"^flop (\d+)", out, re.M).group(1))
    assert flop == int(np.diff(np.asarray(B.rpt, np.int64))[A.col].sum())
    tp = phase_table(out, "== precise")
    names = ["pre_allocate", "compute_flop", "prefix_sum_flop", "load_balance", "symbolic", "prefix_sum_nnz",
             "allocate_C", "compute", "cleanup", "sum_total"]
    assert list(tp) == names
    assert abs(tp["sum_total"] - sum(tp[k] for k in names[:8])) <= 0.006
    tu = phase_table(out, "== upper")
    unames = ["thread", "pre_allocate", "compute_flop", "prefix_sum_flop", "load_balance", "allocate_hook",
              "compute", "prefix_sum_nnz", "allocate_C", "copy", "cleanup", "sum_total"]
    assert list(tu) == unames
    assert abs(tu["sum_total"] - sum(tu[k] for k in unames[:11])) <= 0.007
    assert re.search(r"checks_failed 0$", out, re.M)
    rpt = np.fromfile(prefix + ".rpt", np.int32)
    col = np.fromfile(prefix + ".col", np.int32)
    val = np.fromfile(prefix + ".val", np.float64)
    R = oracle.brmerge_precise(as_tuple(A), as_tuple(B))
    assert_csr_equal(gpu.CSR(rpt, col, val, A.M, B.N), R)


def test_brm_driver_reports_precise_equals_upper(gpu):
    for extra in ([], ["--upper"]):
        p = subprocess.run([DRIVER, *extra, "--reps", "2", "gen:rmat:12"], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        assert "Precise and Upper agree: pass" in p.stdout
        assert p.stdout.count("time(ms):") == 2
