"""ctypes binding of libbrmerge.so (include/brmerge.h, include/brmerge_gen.h).

The library is built in-tree by ``brmerge_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for any product entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# BRM_LIB selects another build of the library (e.g. a kernel-shape variant)
LIB_PATH = os.environ.get("BRM_LIB") or os.path.join(HERE, "libbrmerge.so")

BRM_OK = 0
BRM_E_INVALID = 1
BRM_E_NOMEM = 2
BRM_E_OVERFLOW = 3
BRM_E_DEVICE = 4
BRM_E_INTERNAL = 5

STATUS_NAMES = {0: "BRM_OK", 1: "BRM_E_INVALID", 2: "BRM_E_NOMEM", 3: "BRM_E_OVERFLOW",
                4: "BRM_E_DEVICE", 5: "BRM_E_INTERNAL"}


class BrmError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class PreciseTimingC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "pre_allocate", "compute_flop", "prefix_sum_flop", "load_balance", "symbolic",
        "prefix_sum_nnz", "allocate_C", "compute", "cleanup", "total", "h2d", "d2h")]


class UpperTimingC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "thread", "pre_allocate", "compute_flop", "prefix_sum_flop", "load_balance", "allocate_hook",
        "compute", "prefix_sum_nnz", "allocate_C", "copy", "cleanup", "total", "h2d", "d2h")]


class StatsC(ctypes.Structure):
    _fields_ = [("rows_by_tier", ctypes.c_int64 * 4)] + [(n, ctypes.c_int64) for n in (
        "total_ip", "nnz_c", "long_waves", "leaf_tasks", "leaf_ip", "leaf_out", "passes", "group_tasks",
        "group_rounds", "group_in", "group_out", "upper_waves")]


class PairC(ctypes.Structure):
    _fields_ = [("arpt", ctypes.c_void_p), ("acol", ctypes.c_void_p), ("aval", ctypes.c_void_p),
                ("nnzA", ctypes.c_int64), ("brpt", ctypes.c_void_p), ("bcol", ctypes.c_void_p),
                ("bval", ctypes.c_void_p), ("nnzB", ctypes.c_int64), ("M", ctypes.c_int), ("K", ctypes.c_int),
                ("N", ctypes.c_int)]


class ServerResultC(ctypes.Structure):
    _fields_ = [("crpt", ctypes.c_void_p), ("ccol", ctypes.c_void_p), ("cval", ctypes.c_void_p),
                ("nnz", ctypes.c_int64), ("ip", ctypes.c_int64), ("M", ctypes.c_int), ("N", ctypes.c_int),
                ("status", ctypes.c_int), ("batch_size", ctypes.c_int), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_done", ctypes.c_double)]


# Every symbol include/brmerge.h and include/brmerge_gen.h declare.
EXPORTED = [
    "brm_init", "brm_last_error", "brm_version",
    "brm_spgemm_precise_i32", "brm_spgemm_upper_i32", "brm_spgemm_precise_i64", "brm_spgemm_upper_i64",
    "brm_host_free_i32", "brm_host_free_i64", "brm_host_free_f64",
    "brm_spgemm_precise_dev", "brm_spgemm_upper_dev",
    "brm_dev_malloc", "brm_dev_free", "brm_dev_trim", "brm_rpt_to_i32", "brm_csr_transpose_dev",
    "brm_flop_count", "brm_flop_count_dev",
    "brm_set_param", "brm_get_param", "brm_debug_fail_alloc_after",
    "brm_memcpy_h2d", "brm_memcpy_d2h", "brm_prof_reset", "brm_prof_report", "brm_last_stats", "brm_bench_copy",
    "brm_checksum_dev", "brm_spgemm_precise_dev_noalloc", "brm_noalloc_scratch_bytes",
    "brm_spgemm_batched_dev", "brm_batch_plan_create", "brm_batch_plan_execute", "brm_batch_plan_result",
    "brm_batch_plan_info", "brm_batch_plan_destroy", "brm_set_tune_cache",
    "brm_server_create", "brm_server_add_b", "brm_server_submit", "brm_server_poll", "brm_server_wait",
    "brm_server_release", "brm_server_stats", "brm_server_destroy",
    "brm_gen_er", "brm_gen_er_var", "brm_gen_rmat", "brm_gen_stencil27", "brm_gen_aggregation",
    "brm_gen_banded", "brm_csr_transpose", "brm_read_mtx", "brm_write_mtx",
]

_LIB = None

P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
U64 = ctypes.c_uint64
PP = ctypes.POINTER(ctypes.c_void_p)


def lib():
    """Load libbrmerge.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m brmerge_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "brm_init": ([I], I),
        "brm_last_error": ([], ctypes.c_char_p),
        "brm_version": ([], ctypes.c_char_p),
        "brm_spgemm_precise_i32": ([P, P, P, P, P, P, PP, PP, PP, I, I, I, ctypes.POINTER(I), P], I),
        "brm_spgemm_upper_i32": ([P, P, P, P, P, P, PP, PP, PP, I, I, I, ctypes.POINTER(I), P], I),
        "brm_spgemm_precise_i64": ([P, P, P, P, P, P, PP, PP, PP, I, I, I, ctypes.POINTER(I64), P], I),
        "brm_spgemm_upper_i64": ([P, P, P, P, P, P, PP, PP, PP, I, I, I, ctypes.POINTER(I64), P], I),
        "brm_host_free_i32": ([P], None),
        "brm_host_free_i64": ([P], None),
        "brm_host_free_f64": ([P], None),
        "brm_spgemm_precise_dev": ([P, P, P, I64, P, P, P, I64, I, I, I, PP, PP, PP, ctypes.POINTER(I64), P, P], I),
        "brm_spgemm_upper_dev": ([P, P, P, I64, P, P, P, I64, I, I, I, PP, PP, PP, ctypes.POINTER(I64), P, P], I),
        "brm_dev_malloc": ([U64], P),
        "brm_dev_free": ([P], None),
        "brm_dev_trim": ([], None),
        "brm_rpt_to_i32": ([P, I64, P, P], I),
        "brm_csr_transpose_dev": ([P, P, P, I64, I, I, P, P, P, P], I),
        "brm_flop_count": ([P, P, P, I, I, ctypes.POINTER(I64)], I),
        "brm_flop_count_dev": ([P, P, I64, P, I, I, P, ctypes.POINTER(I64), P], I),
        "brm_set_param": ([ctypes.c_char_p, I64], I),
        "brm_get_param": ([ctypes.c_char_p], I64),
        "brm_debug_fail_alloc_after": ([I64], None),
        "brm_memcpy_h2d": ([P, P, U64], I),
        "brm_memcpy_d2h": ([P, P, U64], I),
        "brm_prof_reset": ([], None),
        "brm_prof_report": ([ctypes.c_char_p, U64], I),
        "brm_last_stats": ([P], I),
        "brm_bench_copy": ([P, P, U64, P], I),
        "brm_checksum_dev": ([P, U64, ctypes.POINTER(U64), P], I),
        "brm_spgemm_precise_dev_noalloc": ([P, P, P, I64, P, P, P, I64, I, I, I, P, P, P, I64, ctypes.POINTER(I64),
                                            P], I),
        "brm_noalloc_scratch_bytes": ([], U64),
        "brm_spgemm_batched_dev": ([P, I, PP, PP, PP, ctypes.POINTER(I64), P, P, P], I),
        "brm_batch_plan_create": ([P, I, I, PP, P], I),
        "brm_batch_plan_execute": ([P, P], I),
        "brm_batch_plan_result": ([P, PP, PP, PP, ctypes.POINTER(I64), PP], I),
        "brm_batch_plan_info": ([P, ctypes.c_char_p, U64], I),
        "brm_batch_plan_destroy": ([P], None),
        "brm_set_tune_cache": ([ctypes.c_char_p], I),
        "brm_server_create": ([I64, I, I, PP], I),
        "brm_server_add_b": ([P, P, P, P, I64, I, I, ctypes.POINTER(I)], I),
        "brm_server_submit": ([P, I, P, P, P, I64, I, ctypes.POINTER(I64)], I),
        "brm_server_poll": ([P, I64, ctypes.POINTER(I)], I),
        "brm_server_wait": ([P, I64, P], I),
        "brm_server_release": ([P, I64], I),
        "brm_server_stats": ([P, ctypes.c_char_p, U64], I),
        "brm_server_destroy": ([P], None),
        "brm_gen_er": ([I, I, I, U64, PP, PP, PP], I),
        "brm_gen_er_var": ([I, I, I, I, U64, PP, PP, PP], I),
        "brm_gen_rmat": ([I, I, ctypes.c_double, ctypes.c_double, ctypes.c_double, U64, PP, PP, PP], I),
        "brm_gen_stencil27": ([I, PP, PP, PP], I),
        "brm_gen_aggregation": ([I, I, PP, PP, PP], I),
        "brm_gen_banded": ([I, I, I, I, U64, PP, PP, PP], I),
        "brm_csr_transpose": ([I, I, P, P, P, PP, PP, PP], I),
        "brm_read_mtx": ([ctypes.c_char_p, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I64), PP, PP, PP], I),
        "brm_write_mtx": ([ctypes.c_char_p, I, I, P, P, P], I),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


def last_error():
    return lib().brm_last_error().decode(errors="replace")


def check(status):
    if status != BRM_OK:
        raise BrmError(status, last_error())
    return status
