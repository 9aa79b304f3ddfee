"""ctypes binding of libeiger_groupby.so (the C ABI in include/eiger_groupby.h).

This is the binding a Python caller of the reference's operator API would add:
plain pointers and sizes, no torch types.  Loading fails loudly when the
library has not been built; nothing here falls back to a CPU path.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libeiger_groupby.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "eiger_groupby.h")


class ColumnView(C.Structure):
    """eg_column_view = gamax::column_view (memory/column.hpp:13-58)."""
    _fields_ = [("data", C.c_void_p), ("size", C.c_int64), ("type", C.c_int32),
                ("reserved", C.c_int32), ("offset", C.c_int64)]


class AggRequest(C.Structure):
    """eg_agg_request = gamax::groupby::aggregation_request (aggregation.hpp:24-31)."""
    _fields_ = [("column_index", C.c_int64), ("agg_kind", C.c_int32), ("reserved", C.c_int32)]


class Hint(C.Structure):
    """eg_groupby_hint = gamax::groupby::groupby_hint (groupby.hpp:27-42)."""
    _fields_ = [("groups_upper_bound", C.c_int64), ("zero_based_ids", C.c_int32),
                ("optimize_sort", C.c_int32), ("groupby_algo", C.c_int32),
                ("agg_count", C.c_int32), ("enable_skf", C.c_int32),
                ("allow_hash_fallback", C.c_int32)]


BATCH_MAX_AGGS = 4


class BatchQuery(C.Structure):
    """eg_batch_query: one query of a batched group-by."""
    _fields_ = [("keys", C.c_void_p), ("vals", C.c_void_p), ("rows", C.c_int64),
                ("group_bound", C.c_int64), ("capacity", C.c_int64), ("out_keys", C.c_void_p),
                ("out_aggs", C.c_void_p * BATCH_MAX_AGGS)]


class BatchTuning(C.Structure):
    """eg_batch_tuning: execution parameters of a batched group-by."""
    _fields_ = [("tile_rows_log2", C.c_int32), ("lds_slots_log2", C.c_int32),
                ("lds_max_bound_log2", C.c_int32), ("blocks_per_cu", C.c_int32), ("radix_bits", C.c_int32)]


MLP_PARAMS = 5508


class MlpQuery(C.Structure):
    """eg_mlp_query: one key column to featurise."""
    _fields_ = [("keys", C.c_void_p), ("rows", C.c_int64), ("group_bound", C.c_int64), ("key_type", C.c_int32),
                ("nops", C.c_int32), ("value_bytes", C.c_int32), ("reserved", C.c_int32)]


class MlpTrainConfig(C.Structure):
    """eg_mlp_train_config."""
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("lam", C.c_float),
                ("weight_decay", C.c_float), ("seed", C.c_uint64)]


class LibraryNotBuilt(RuntimeError):
    pass


_lib = None


def _proto(L):
    vp, i32, i64, u32, u64, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_size_t
    CV, PCV = ColumnView, C.POINTER(ColumnView)
    PAR, PH = C.POINTER(AggRequest), C.POINTER(Hint)
    PT = C.POINTER(vp)

    def f(name, res, args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args

    f("eg_last_error", C.c_char_p, [])
    f("eg_version", C.c_char_p, [])
    f("eg_hint_init", None, [PH])
    gb = [PCV, i64, PCV, i64, PAR, i64, PH, vp, PT]
    for name in ("eg_groupby", "eg_groupby_hash", "eg_groupby_sort", "eg_groupby_sort_gfur",
                 "eg_groupby_partition", "eg_groupby_partition_gfur"):
        f(name, C.c_int, gb)
    f("eg_groupby_hash_noalloc", C.c_int, [PCV, i64, PCV, i64, PAR, i64, i64, PCV, i64, C.POINTER(i64), vp])
    f("eg_distinct_hash", C.c_int, [PCV, i64, PH, vp, PT])
    f("eg_distinct_sort", C.c_int, [PCV, i64, PH, vp, PT])
    f("eg_aggregate", C.c_int, [PCV, i64, PAR, i64, vp, PT])
    f("eg_agg_output_type", C.c_int, [i32, i32, C.POINTER(i32)])
    f("eg_table_num_columns", i64, [vp])
    f("eg_table_num_rows", i64, [vp])
    f("eg_table_get_column", C.c_int, [vp, i64, PCV])
    f("eg_table_release_column", C.c_int, [vp, i64, PCV])
    f("eg_table_free", None, [vp])
    f("eg_malloc", C.c_int, [PT, i64, i32, vp])
    f("eg_free", C.c_int, [vp])
    f("eg_memory_usage", C.c_int, [C.POINTER(i64), C.POINTER(i64)])
    f("eg_set_memory_mode", C.c_int, [i32, i64])
    f("eg_make_column_from_host", C.c_int, [vp, i64, i32, vp, PCV])
    f("eg_copy_column_to_host", C.c_int, [CV, vp, vp])
    f("eg_sort_column_pair", C.c_int, [CV, CV, i32, i32, i32, vp, sz, vp, PCV, PCV])
    f("eg_sort_column", C.c_int, [CV, i32, i32, i32, vp, sz, vp, PCV])
    f("eg_argsort", C.c_int, [CV, i32, i32, i32, vp, sz, vp, PCV, PCV])
    f("eg_partition_column_pair", C.c_int, [CV, CV, i32, i32, vp, sz, vp, PCV, PCV])
    f("eg_partition_column", C.c_int, [CV, i32, i32, vp, sz, vp, PCV])
    f("eg_arg_partition", C.c_int, [CV, i32, i32, vp, sz, vp, PCV, PCV])
    f("eg_compute_partition_offsets", C.c_int, [CV, i32, i32, vp, PCV])
    f("eg_dictionary_encode_with_dict", C.c_int, [CV, CV, i32, vp, PCV])
    f("eg_dictionary_encode", C.c_int, [CV, i32, i64, i32, vp, PCV, C.POINTER(i64), PCV])
    f("eg_compute_ndv_approx", C.c_int, [CV, vp, C.POINTER(i64)])
    f("eg_auto_select_algo", C.c_int, [PCV, i64, PH, vp])
    f("eg_kernel_timing_enable", C.c_int, [i32])
    f("eg_kernel_timing_reset", C.c_int, [])
    f("eg_kernel_timing_query", C.c_int, [C.c_char_p, C.POINTER(i64), C.POINTER(C.c_double)])
    f("eg_probe_copy", C.c_int, [i64, i32, vp, C.POINTER(C.c_double)])
    f("eg_jit_configure", C.c_int, [i32, i64])
    f("eg_jit_wait", C.c_int, [])
    f("eg_jit_stats", C.c_int, [C.POINTER(i64)])
    f("eg_jit_selftest", C.c_int, [C.c_char_p, C.c_char_p, i64])
    f("eg_jit_prefetch", C.c_int, [i32, C.POINTER(i32), i64, PAR, i64, C.c_char_p, C.POINTER(i32)])
    f("eg_gen_keys_uniform", C.c_int, [vp, i32, i64, i64, i64, i64, u64, vp])
    f("eg_zipf_cdf", C.c_int, [vp, i64, C.c_double])
    f("eg_gen_keys_zipf", C.c_int, [vp, i32, i64, i64, vp, i64, u64, vp])
    f("eg_gen_values", C.c_int, [vp, i32, i32, i64, i64, u64, u32, vp])
    PMQ, PF = C.POINTER(MlpQuery), C.POINTER(C.c_float)
    f("eg_mlp_features", C.c_int, [PMQ, i64, vp, vp])
    f("eg_mlp_forward", C.c_int, [vp, vp, i64, vp, vp])
    f("eg_mlp_loss_grad", C.c_int, [vp, vp, vp, vp, i64, C.c_float, vp, PF, vp])
    f("eg_mlp_train", C.c_int, [vp, vp, vp, vp, i64, C.POINTER(MlpTrainConfig), PF, vp])
    f("eg_mlp_load", C.c_int, [vp, i64])
    f("eg_mlp_loaded", C.c_int, [])
    f("eg_mlp_select_batch", C.c_int, [PMQ, i64, vp, vp, vp])
    f("eg_auto_select_algo_batch", C.c_int, [PCV, i64, vp, vp, vp])
    f("eg_mlp_select", C.c_int, [ColumnView, i64, i32, i32, vp, C.POINTER(i32), C.POINTER(i64)])
    PBQ, PBT = C.POINTER(BatchQuery), C.POINTER(BatchTuning)
    f("eg_batch_tuning_default", None, [PBT])
    f("eg_batch_plan_create", C.c_int, [PBQ, i64, i32, i32, PAR, i64, PBT, vp, PT])
    f("eg_batch_plan_tuning", C.c_int, [vp, PBT])
    f("eg_batch_run", C.c_int, [vp, vp])
    f("eg_batch_graph_launch", C.c_int, [vp, vp])
    f("eg_batch_results", C.c_int, [vp, vp, vp, vp])
    f("eg_batch_plan_free", None, [vp])
    f("eg_batch_autotune", C.c_int, [PBQ, i64, i32, i32, PAR, i64, PBT, i64, i32, vp, PBT,
                                     C.POINTER(C.c_double)])


def lib():
    """The loaded library (raises LibraryNotBuilt if the .so is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(
                f"{LIB_PATH} is missing: build it with `make -C eiger_amd/csrc` "
                "(or __graft_entry__.build()) before using eiger_amd")
        L = C.CDLL(LIB_PATH)
        _proto(L)
        _lib = L
        atexit.register(_drain_compiles)
    return _lib


def _drain_compiles():
    """Interpreter exit: lets queued run-time compiles finish while the
    process is still whole, so none is running once native teardown starts."""
    if _lib is not None:
        _lib.eg_jit_wait()


def header_functions(path: str = HEADER):
    """Names of every function the C header declares."""
    import re
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(eg_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))
