"""brmerge_amd -- BRMerge SpGEMM for AMD Instinct MI355X (gfx950).

Host-side mirror of the reference's SpGEMM interface over libbrmerge.so's
C ABI (include/brmerge.h).  The numeric work runs in hand-written HIP
kernels; importing this package never falls back to a CPU implementation.
"""
from ._lib import BrmError, lib, LIB_PATH  # noqa: F401
from .csr import CSR, Precise_Timing, Upper_Timing  # noqa: F401
from .spgemm import (  # noqa: F401
    brmerge_precise, brmerge_upper, compute_flop, DeviceArray, DeviceCSR, spgemm_dev,
    set_param, get_param, prof_reset, prof_report, last_stats,
)
from .batch import spgemm_batched, split_products, BatchPlan, set_tune_cache  # noqa: F401
from .serve import Server  # noqa: F401
from . import gen  # noqa: F401

__version__ = "0.1.0"
