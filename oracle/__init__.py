"""CPU oracle for BRMerge SpGEMM -- test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package, and only as the checker.  The product library
(``brmerge_amd``) never imports it.
"""
from .oracle import (  # noqa: F401
    OracleCSR, load, compute_flop, upper_log, prefix_sum, binary_search_approx,
    brmerge_precise, brmerge_upper, hash_spgemm, max_threads,
)
