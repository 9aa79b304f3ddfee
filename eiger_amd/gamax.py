"""Python mirror of the reference's gamax operator API over the C ABI.

Names and argument meaning follow the reference (src/cuda/ paths):
  gamax::type_id                     types/data_type.hpp:12-28
  gamax::column / table              memory/column.hpp:96-172, memory/table.hpp:143-222
  gamax::groupby::groupby            operators/groupby/groupby.hpp:44-49
  gamax::groupby::{hash,sort,partition}::groupby   groupby_{hash,sort,partition}.hpp
  gamax::groupby::aggregate::aggregate             aggregation.hpp:34-39
  gamax::sort / gamax::partition sub-operators     operators/sort/sort.hpp, operators/partition/partition.hpp
Errors: the reference throws std::runtime_error (GamaxError here) and
std::bad_alloc (MemoryError here).  Every call runs the HIP library; there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from enum import IntEnum
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _lib as _lib_mod
from ._lib import AggRequest, ColumnView, Hint, lib

MAXIMUM_SIZE = (1 << 63) - 1

EG_OK, EG_ERR_INVALID, EG_ERR_UNSUPPORTED, EG_ERR_CAPACITY, EG_ERR_OOM, EG_ERR_DEVICE = range(6)


class type_id(IntEnum):
    EMPTY = 0
    INT8 = 1
    INT16 = 2
    INT32 = 3
    INT64 = 4
    BOOL8 = 5
    UINT8 = 6
    UINT16 = 7
    UINT32 = 8
    UINT64 = 9
    FLOAT32 = 10
    FLOAT64 = 11
    STRING = 12


class aggregation_id(IntEnum):
    SUM = 0
    MAX = 1
    MIN = 2
    AVG = 3
    ANY = 4
    COUNT_ALL = 5


class groupby_algo(IntEnum):
    AUTO_ALGO = 0
    HASH = 1
    SORT = 2
    SORT_OPT = 3
    SORT_GFUR = 4
    PARTITION = 5
    PARTITION_GFUR = 6


class sort_order(IntEnum):
    ASCENDING = 0
    DESCENDING = 1


class val_dist(IntEnum):
    UNIFORM_I64 = 0
    UNIFORM_F32 = 1
    UNIFORM_I32_1000 = 2
    ALL_ONES = 3
    SEQ = 4


_NP = {
    type_id.INT8: np.int8, type_id.INT16: np.int16, type_id.INT32: np.int32,
    type_id.INT64: np.int64, type_id.BOOL8: np.uint8, type_id.UINT8: np.uint8,
    type_id.UINT16: np.uint16, type_id.UINT32: np.uint32, type_id.UINT64: np.uint64,
    type_id.FLOAT32: np.float32, type_id.FLOAT64: np.float64,
}
_FROM_NP = {np.dtype(v): k for k, v in _NP.items() if k != type_id.BOOL8}


def to_type_id(dtype) -> type_id:
    return _FROM_NP[np.dtype(dtype)]


def to_numpy_dtype(t: int):
    return np.dtype(_NP[type_id(t)])


def size_of(t: int) -> int:
    return to_numpy_dtype(t).itemsize


class GamaxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


def _check(rc: int):
    if rc == EG_OK:
        return
    msg = lib().eg_last_error().decode()
    if rc == EG_ERR_OOM:
        raise MemoryError(msg)
    raise GamaxError(rc, msg)


def _stream(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(getattr(stream, "cuda_stream", stream))


@dataclass
class groupby_hint:
    groups_upper_bound: int = MAXIMUM_SIZE
    zero_based_ids: bool = False
    optimize_sort: bool = False
    groupby_algo: int = groupby_algo.AUTO_ALGO
    agg_count: bool = False
    enable_skf: bool = False
    allow_hash_fallback: bool = False

    def _to_c(self) -> Hint:
        h = Hint()
        h.groups_upper_bound = int(self.groups_upper_bound)
        h.zero_based_ids = int(self.zero_based_ids)
        h.optimize_sort = int(self.optimize_sort)
        h.groupby_algo = int(self.groupby_algo)
        h.agg_count = int(self.agg_count)
        h.enable_skf = int(self.enable_skf)
        h.allow_hash_fallback = int(self.allow_hash_fallback)
        return h

    def _from_c(self, h: Hint):
        # the hint is in/out (groupby.cu:73-118, 304-306)
        self.groupby_algo = groupby_algo(h.groupby_algo)
        self.optimize_sort = bool(h.optimize_sort)


@dataclass
class aggregation_request:
    column_index: int
    agg_kind: int


class column_view:
    """Non-owning view of device data (gamax::column_view)."""

    def __init__(self, data: int, size: int, type: int, offset: int = 0):
        self.data, self.size, self.type, self.offset = int(data or 0), int(size), int(type), int(offset)

    def _to_c(self) -> ColumnView:
        v = ColumnView()
        v.data = self.data or None
        v.size = self.size
        v.type = self.type
        v.offset = self.offset
        return v

    def view(self, offset: int = 0) -> "column_view":
        return column_view(self.data, self.size - offset, self.type, self.offset + offset)

    def to_numpy(self, stream=None) -> np.ndarray:
        return copy_column_to_host(self, stream)


class column(column_view):
    """Owning device column (gamax::column); memory from the library's manager."""

    def __init__(self, data: int, size: int, type: int):
        super().__init__(data, size, type, 0)
        self._owned = bool(data)

    def free(self):
        if self._owned and self.data:
            _check(lib().eg_free(C.c_void_p(self.data)))
        self._owned = False
        self.data = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def view(self, offset: int = 0) -> column_view:
        return column_view(self.data, self.size - offset, self.type, offset)

    def type_id(self) -> type_id:
        return type_id(self.type)


class table:
    """Owning result table (gamax::table)."""

    def __init__(self, columns: List[column]):
        self._columns = columns

    def num_columns(self) -> int:
        return len(self._columns)

    def num_rows(self) -> int:
        return self._columns[0].size if self._columns else 0

    def get_column(self, i: int) -> column:
        return self._columns[i]

    def __getitem__(self, i: int) -> column:
        return self._columns[i]

    def columns(self) -> List[column]:
        return list(self._columns)

    def to_numpy(self, stream=None) -> List[np.ndarray]:
        return [c.to_numpy(stream) for c in self._columns]

    def free(self):
        for c in self._columns:
            c.free()


def _take_table(handle: C.c_void_p) -> table:
    L = lib()
    cols = []
    try:
        for i in range(L.eg_table_num_columns(handle)):
            v = ColumnView()
            _check(L.eg_table_release_column(handle, i, C.byref(v)))
            cols.append(column(v.data or 0, v.size, v.type))
    finally:
        L.eg_table_free(handle)
    return table(cols)


def _views(cols: Sequence[column_view]):
    n = max(len(cols), 1)
    arr = (ColumnView * n)()
    for i, c in enumerate(cols):
        arr[i] = c._to_c()
    return arr


def _reqs(requests: Sequence):
    n = max(len(requests), 1)
    arr = (AggRequest * n)()
    for i, r in enumerate(requests):
        ci, kind = (r.column_index, r.agg_kind) if isinstance(r, aggregation_request) else r
        arr[i].column_index = int(ci)
        arr[i].agg_kind = int(kind)
    return arr


def _tv(x) -> List[column_view]:
    if isinstance(x, table):
        return x.columns()
    if isinstance(x, column_view):
        return [x]
    return list(x)


def _groupby_call(fn_name, key, val, requests, hint, stream):
    hint = hint if hint is not None else groupby_hint()
    keys, vals = _tv(key), _tv(val)
    h = hint._to_c()
    out = C.c_void_p()
    rc = getattr(lib(), fn_name)(_views(keys), len(keys), _views(vals), len(vals), _reqs(requests),
                                 len(requests), C.byref(h), _stream(stream), C.byref(out))
    hint._from_c(h)
    _check(rc)
    return _take_table(out)


def _distinct_call(fn_name, key, hint, stream):
    hint = hint if hint is not None else groupby_hint()
    keys = _tv(key)
    h = hint._to_c()
    out = C.c_void_p()
    rc = getattr(lib(), fn_name)(_views(keys), len(keys), C.byref(h), _stream(stream), C.byref(out))
    hint._from_c(h)
    _check(rc)
    return _take_table(out)


def hash_distinct(key, hint: Optional[groupby_hint] = None, stream=None) -> table:
    """gamax::groupby::hash::distinct (groupby_hash.hpp:24-32): one row per
    unique key (tuple); row order unspecified."""
    return _distinct_call("eg_distinct_hash", key, hint, stream)


def sort_distinct(key, hint: Optional[groupby_hint] = None, stream=None) -> table:
    """gamax::groupby::sort::distinct (groupby_sort.hpp:31-48): one key
    column, or several with hint.enable_skf; ascending key order."""
    return _distinct_call("eg_distinct_sort", key, hint, stream)


def groupby(key, val, agg_requests, hint: Optional[groupby_hint] = None, stream=None) -> table:
    """gamax::groupby::groupby (operators/groupby/groupby.hpp:44-49)."""
    return _groupby_call("eg_groupby", key, val, agg_requests, hint, stream)


def hash_groupby(key, val, agg_requests, hint=None, stream=None) -> table:
    """gamax::groupby::hash::groupby (groupby_hash.hpp:17-22)."""
    return _groupby_call("eg_groupby_hash", key, val, agg_requests, hint, stream)


def hash_groupby_noalloc(key, val, agg_requests, out_columns, groups_upper_bound: int = 0, stream=None) -> int:
    """The hash strategy into caller-owned columns with nothing allocated
    (eg_groupby_hash_noalloc; no counterpart in the reference).  out_columns:
    [key, one per request] views typed as hash_groupby's result, each view's
    size its capacity in rows.  Returns the number of groups written."""
    keys, vals, outs = _tv(key), _tv(val), _tv(out_columns)
    ng = C.c_int64(0)
    _check(lib().eg_groupby_hash_noalloc(_views(keys), len(keys), _views(vals), len(vals), _reqs(agg_requests),
                                         len(agg_requests), int(groups_upper_bound), _views(outs), len(outs),
                                         C.byref(ng), _stream(stream)))
    return int(ng.value)


def sort_groupby(key, val, agg_requests, hint=None, stream=None) -> table:
    """gamax::groupby::sort::groupby (groupby_sort.hpp:17-22)."""
    return _groupby_call("eg_groupby_sort", key, val, agg_requests, hint, stream)


def sort_groupby_gfur(key, val, agg_requests, hint=None, stream=None) -> table:
    """gamax::groupby::sort::groupby_gfur (groupby_sort.hpp:24-29)."""
    return _groupby_call("eg_groupby_sort_gfur", key, val, agg_requests, hint, stream)


def partition_groupby(key, val, agg_requests, hint=None, stream=None) -> table:
    """gamax::groupby::partition::groupby (groupby_partition.hpp:52-57)."""
    return _groupby_call("eg_groupby_partition", key, val, agg_requests, hint, stream)


def partition_groupby_gfur(key, val, agg_requests, hint=None, stream=None) -> table:
    """gamax::groupby::partition::groupby_gfur (groupby_partition.hpp:59-64)."""
    return _groupby_call("eg_groupby_partition_gfur", key, val, agg_requests, hint, stream)


def aggregate(input, agg_requests, stream=None) -> table:
    """gamax::groupby::aggregate::aggregate (aggregation.hpp:34-39)."""
    cols = _tv(input)
    out = C.c_void_p()
    _check(lib().eg_aggregate(_views(cols), len(cols), _reqs(agg_requests), len(agg_requests),
                              _stream(stream), C.byref(out)))
    return _take_table(out)


def agg_output_data_type(agg_kind: int, t: int) -> type_id:
    """gamax::groupby::agg_output_data_type (aggregation.hpp:33)."""
    o = C.c_int32()
    _check(lib().eg_agg_output_type(int(agg_kind), int(t), C.byref(o)))
    return type_id(o.value)


# ---------------------------------------------------------------- columns
def make_column_from_host(arr: np.ndarray, stream=None, type: Optional[int] = None) -> column:
    """gamax::make_column_from_host (memory/column.hpp:185-191)."""
    arr = np.ascontiguousarray(arr)
    t = int(type) if type is not None else int(to_type_id(arr.dtype))
    v = ColumnView()
    _check(lib().eg_make_column_from_host(arr.ctypes.data_as(C.c_void_p) if arr.size else None,
                                          arr.size, t, _stream(stream), C.byref(v)))
    return column(v.data or 0, v.size, v.type)


def copy_column_to_host(col: column_view, stream=None) -> np.ndarray:
    """gamax::copy_column_to_host (memory/column.hpp:199-213)."""
    out = np.empty(col.size, dtype=to_numpy_dtype(col.type))
    if col.size:
        _check(lib().eg_copy_column_to_host(col._to_c(), out.ctypes.data_as(C.c_void_p), _stream(stream)))
    return out


def make_column(t: int, size: int, clear: bool = False, stream=None) -> column:
    """gamax::make_column (memory/column.hpp:176)."""
    p = C.c_void_p()
    _check(lib().eg_malloc(C.byref(p), size * size_of(t), int(clear), _stream(stream)))
    return column(p.value or 0, size, t)


def memory_usage():
    cur, peak = C.c_int64(), C.c_int64()
    _check(lib().eg_memory_usage(C.byref(cur), C.byref(peak)))
    return cur.value, peak.value


def set_memory_mode(mode: int, pool_bytes: int = 0):
    _check(lib().eg_set_memory_mode(int(mode), int(pool_bytes)))


# ---------------------------------------------------------------- sort / partition
def _own(v: ColumnView) -> column:
    return column(v.data or 0, v.size, v.type)


def sort_column_pair(key_col, val_col, begin_bit=0, end_bit=-1, order=sort_order.ASCENDING, stream=None):
    """gamax::sort::sort_column_pair (operators/sort/sort.hpp:25-34)."""
    a, b = ColumnView(), ColumnView()
    _check(lib().eg_sort_column_pair(key_col._to_c(), val_col._to_c(), begin_bit, end_bit, int(order),
                                     None, 0, _stream(stream), C.byref(a), C.byref(b)))
    return _own(a), _own(b)


def sort_column(col, begin_bit=0, end_bit=-1, order=sort_order.ASCENDING, stream=None) -> column:
    """gamax::sort::sort_column (sort.hpp:46-54)."""
    a = ColumnView()
    _check(lib().eg_sort_column(col._to_c(), begin_bit, end_bit, int(order), None, 0, _stream(stream), C.byref(a)))
    return _own(a)


def argsort(col, begin_bit=0, end_bit=-1, order=sort_order.ASCENDING, stream=None):
    """gamax::sort::argsort (sort.hpp:66-75): (sorted column, int64 row ids)."""
    a, b = ColumnView(), ColumnView()
    _check(lib().eg_argsort(col._to_c(), begin_bit, end_bit, int(order), None, 0, _stream(stream),
                            C.byref(a), C.byref(b)))
    return _own(a), _own(b)


def partition_column_pair(key_col, val_col, begin_bit, end_bit, stream=None):
    """gamax::partition::partition_column_pair (operators/partition/partition.hpp:46-54)."""
    a, b = ColumnView(), ColumnView()
    _check(lib().eg_partition_column_pair(key_col._to_c(), val_col._to_c(), begin_bit, end_bit, None, 0,
                                          _stream(stream), C.byref(a), C.byref(b)))
    return _own(a), _own(b)


def partition_column(col, begin_bit, end_bit, stream=None) -> column:
    """gamax::partition::partition_column (partition.hpp:56-63)."""
    a = ColumnView()
    _check(lib().eg_partition_column(col._to_c(), begin_bit, end_bit, None, 0, _stream(stream), C.byref(a)))
    return _own(a)


def arg_partition(col, begin_bit, end_bit, stream=None):
    """gamax::partition::arg_partition (partition.hpp:65-72)."""
    a, b = ColumnView(), ColumnView()
    _check(lib().eg_arg_partition(col._to_c(), begin_bit, end_bit, None, 0, _stream(stream), C.byref(a), C.byref(b)))
    return _own(a), _own(b)


def compute_partition_offsets(partitioned_key_col, begin_bit, end_bit, stream=None) -> column:
    """gamax::partition::compute_partition_offsets (partition.hpp:74-79)."""
    a = ColumnView()
    _check(lib().eg_compute_partition_offsets(partitioned_key_col._to_c(), begin_bit, end_bit, _stream(stream),
                                              C.byref(a)))
    return _own(a)


# ---------------------------------------------------------------- dictionary encoding
def dictionary_encode_with_dict(col, dict_col, encode_type=type_id.INT32, stream=None) -> column:
    """gamax::compress::dictionary_encode(column, dict, encode_type)
    (compress/dictionary_encode.hpp:26-30, dictionary_encode_dict_ready.cu:15-80):
    the position of each key in dict_col, INT32_MIN where absent."""
    a = ColumnView()
    _check(lib().eg_dictionary_encode_with_dict(col._to_c(), dict_col._to_c(), int(encode_type), _stream(stream),
                                                C.byref(a)))
    return _own(a)


def dictionary_encode(col, encode_type=type_id.INT32, dictionary_size_hint=0, with_order=False,
                      return_dictionary=False, stream=None):
    """gamax::compress::dictionary_encode / dictionary_encode_with_order
    (compress/dictionary_encode.hpp:12-24): (encoded column, dictionary size),
    dense ids in [0, size); with_order numbers keys in ascending order.  With
    return_dictionary the distinct keys (id i at row i) come third."""
    a, d = ColumnView(), ColumnView()
    size = C.c_int64()
    _check(lib().eg_dictionary_encode(col._to_c(), int(encode_type), int(dictionary_size_hint), int(with_order),
                                      _stream(stream), C.byref(a), C.byref(size),
                                      C.byref(d) if return_dictionary else None))
    if return_dictionary:
        return _own(a), size.value, _own(d)
    return _own(a), size.value


# ---------------------------------------------------------------- statistics / dispatch
def compute_ndv_approx(col, stream=None) -> int:
    """gamax::stats::compute_ndv_approx (stats/stats_compute.hpp:72)."""
    o = C.c_int64()
    _check(lib().eg_compute_ndv_approx(col._to_c(), _stream(stream), C.byref(o)))
    return o.value


def auto_select_groupby_algo(key, hint: groupby_hint, stream=None) -> groupby_hint:
    """auto_select_groupby_algo (operators/groupby/groupby.cu:68-119)."""
    keys = _tv(key)
    h = hint._to_c()
    _check(lib().eg_auto_select_algo(_views(keys), len(keys), C.byref(h), _stream(stream)))
    hint._from_c(h)
    return hint


# ---------------------------------------------------------------- timing
def kernel_timing(enable: bool):
    _check(lib().eg_kernel_timing_enable(int(enable)))


def kernel_timing_reset():
    _check(lib().eg_kernel_timing_reset())


def kernel_time(name: str):
    """(launches, total milliseconds) of kernel `name` since the last reset."""
    n, ms = C.c_int64(), C.c_double()
    _check(lib().eg_kernel_timing_query(name.encode(), C.byref(n), C.byref(ms)))
    return n.value, ms.value


def probe_copy(nbytes: int, iters: int = 10, stream=None) -> float:
    """Mean milliseconds of one device copy of `nbytes` (kernel "probe_copy")."""
    ms = C.c_double()
    _check(lib().eg_probe_copy(int(nbytes), int(iters), _stream(stream), C.byref(ms)))
    return ms.value


# ---------------------------------------------------------------- run-time specialisation
class jit_mode(IntEnum):
    OFF = 0
    ASYNC = 1
    SYNC = 2


def jit_configure(mode: Optional[int] = None, min_rows: Optional[int] = None):
    """Mode of the run-time kernel specialisation (OFF / ASYNC / SYNC) and the
    smallest input it is used for; None keeps the current value."""
    _check(lib().eg_jit_configure(-1 if mode is None else int(mode), -1 if min_rows is None else int(min_rows)))


def jit_wait():
    """Blocks until no specialised kernel is being compiled."""
    _check(lib().eg_jit_wait())


def jit_stats() -> dict:
    out = (C.c_int64 * 6)()
    _check(lib().eg_jit_stats(out))
    return {"compiled": out[0], "failed": out[1], "pending": out[2], "launches": out[3],
            "mode": out[4], "min_rows": out[5]}


def jit_prefetch(key_type: int, value_types: Sequence[int], requests, arch: Optional[str] = None) -> int:
    """Queues the hash and partition kernels specialised for this request list
    on the background compiler (arch None = the current device's); returns the
    number of new compiles.  Needs no device when `arch` is given."""
    vt = (C.c_int32 * max(len(value_types), 1))(*[int(t) for t in value_types])
    q = C.c_int32(0)
    _check(lib().eg_jit_prefetch(int(key_type), vt, len(value_types), _reqs(requests), len(requests),
                                 arch.encode() if arch else None, C.byref(q)))
    return q.value


def jit_selftest(arch: str = "gfx950") -> None:
    """Compiles a fixed set of specialised kernels for `arch` (no device
    needed); raises with the compiler log on failure."""
    log = C.create_string_buffer(1 << 16)
    st = lib().eg_jit_selftest(arch.encode(), log, len(log))
    if st != EG_OK:
        raise GamaxError(st, lib().eg_last_error().decode() + "\n" + log.value.decode(errors="replace"))


# ---------------------------------------------------------------- synthetic data
def gen_keys_uniform(n: int, distinct: int, seed: int = 42, t: int = type_id.INT32,
                     begin: int = 0, count: Optional[int] = None, stream=None) -> column:
    count = n - begin if count is None else count
    col = make_column(t, count, stream=stream)
    _check(lib().eg_gen_keys_uniform(C.c_void_p(col.data), int(t), begin, count, n, distinct, seed, _stream(stream)))
    return col


def zipf_cdf(distinct: int, s: float) -> np.ndarray:
    cdf = np.empty(distinct, dtype=np.float64)
    _check(lib().eg_zipf_cdf(cdf.ctypes.data_as(C.c_void_p), distinct, float(s)))
    return cdf


def gen_keys_zipf(n: int, distinct: int, s: float, seed: int = 42, t: int = type_id.INT32,
                  begin: int = 0, stream=None, cdf_col: Optional[column] = None) -> column:
    own_cdf = cdf_col is None
    if own_cdf:
        cdf_col = make_column_from_host(zipf_cdf(distinct, s), stream)
    col = make_column(t, n, stream=stream)
    _check(lib().eg_gen_keys_zipf(C.c_void_p(col.data), int(t), begin, n, C.c_void_p(cdf_col.data), distinct,
                                  seed, _stream(stream)))
    if own_cdf:
        cdf_col.free()
    return col


def gen_values(n: int, dist: int, t: int, seed: int = 42, stream_id: int = 2, begin: int = 0,
               stream=None) -> column:
    col = make_column(t, n, stream=stream)
    _check(lib().eg_gen_values(C.c_void_p(col.data), int(t), int(dist), begin, n, seed, stream_id, _stream(stream)))
    return col


# ---------------------------------------------------------------- batched group-by
class batch_query:
    """One query of a batched group-by: key column, optional value column,
    group upper bound, and caller-owned output columns of `capacity` rows."""

    def __init__(self, keys: column_view, vals: Optional[column_view], group_bound: int,
                 out_keys: column_view, out_aggs: Sequence[column_view], capacity: Optional[int] = None):
        self.keys, self.vals, self.group_bound = keys, vals, int(group_bound)
        self.out_keys, self.out_aggs = out_keys, list(out_aggs)
        self.capacity = int(out_keys.size if capacity is None else capacity)


def batch_tuning(tile_rows_log2=None, lds_slots_log2=None, lds_max_bound_log2=None, blocks_per_cu=None,
                 radix_bits=None):
    t = _lib_mod.BatchTuning()
    lib().eg_batch_tuning_default(C.byref(t))
    for k, v in (("tile_rows_log2", tile_rows_log2), ("lds_slots_log2", lds_slots_log2),
                 ("lds_max_bound_log2", lds_max_bound_log2), ("blocks_per_cu", blocks_per_cu),
                 ("radix_bits", radix_bits)):
        if v is not None:
            setattr(t, k, int(v))
    return t


def _batch_array(queries: Sequence[batch_query]):
    arr = (_lib_mod.BatchQuery * max(len(queries), 1))()
    for i, q in enumerate(queries):
        a = arr[i]
        a.keys = q.keys.data + q.keys.offset * size_of(q.keys.type) if q.keys.data else None
        a.vals = (q.vals.data + q.vals.offset * size_of(q.vals.type)) if (q.vals is not None and q.vals.data) else None
        a.rows = q.keys.size
        a.group_bound = q.group_bound
        a.capacity = q.capacity
        a.out_keys = q.out_keys.data or None
        for j, c in enumerate(q.out_aggs):
            a.out_aggs[j] = c.data or None
    return arr


class batch_plan:
    """A planned batch: run() / graph_launch() execute it asynchronously on
    the stream; results() synchronises and returns (num_groups, status)."""

    def __init__(self, queries: Sequence[batch_query], key_type: int, val_type: int, requests, tuning=None,
                 stream=None):
        self._nq = len(queries)
        self._arr = _batch_array(queries)
        rq = _reqs(requests)
        h = C.c_void_p()
        _check(lib().eg_batch_plan_create(self._arr, self._nq, int(key_type), int(val_type), rq, len(requests),
                                          C.byref(tuning) if tuning is not None else None, _stream(stream),
                                          C.byref(h)))
        self._h = h

    def tuning(self):
        t = _lib_mod.BatchTuning()
        _check(lib().eg_batch_plan_tuning(self._h, C.byref(t)))
        return t

    def run(self, stream=None):
        _check(lib().eg_batch_run(self._h, _stream(stream)))

    def graph_launch(self, stream=None):
        _check(lib().eg_batch_graph_launch(self._h, _stream(stream)))

    def results(self, stream=None):
        ng = np.zeros(self._nq, dtype=np.int64)
        st = np.zeros(self._nq, dtype=np.int32)
        _check(lib().eg_batch_results(self._h, ng.ctypes.data_as(C.c_void_p), st.ctypes.data_as(C.c_void_p),
                                      _stream(stream)))
        return ng, st

    def free(self):
        if self._h:
            lib().eg_batch_plan_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def batch_autotune(queries: Sequence[batch_query], key_type: int, val_type: int, requests, candidates,
                   reps: int = 2, stream=None):
    """Times every candidate tuning on this batch; returns (best, times_ms)."""
    arr = _batch_array(queries)
    cand = (_lib_mod.BatchTuning * len(candidates))(*candidates)
    times = np.zeros(len(candidates), dtype=np.float64)
    best = _lib_mod.BatchTuning()
    _check(lib().eg_batch_autotune(arr, len(queries), int(key_type), int(val_type), _reqs(requests), len(requests),
                                   cand, len(candidates), int(reps), _stream(stream), C.byref(best),
                                   times.ctypes.data_as(C.POINTER(C.c_double))))
    return best, times


# ---------------------------------------------------------------- learned dispatch
MLP_PARAMS = _lib_mod.MLP_PARAMS
MLP_FEATURES = 16


def _dev_f32(a) -> column:
    return make_column_from_host(np.ascontiguousarray(a, dtype=np.float32).ravel())


def mlp_features(keys: Sequence[column_view], group_bounds=None, nops=2, value_bytes=8, stream=None) -> np.ndarray:
    """The 16 dispatch features of each key column (computed on the device)."""
    n = len(keys)
    arr = (_lib_mod.MlpQuery * max(n, 1))()
    for i, k in enumerate(keys):
        arr[i].keys = (k.data + k.offset * size_of(k.type)) if k.data else None
        arr[i].rows = k.size
        arr[i].group_bound = int(group_bounds[i]) if group_bounds is not None else 0
        arr[i].key_type = k.type
        arr[i].nops = int(nops[i]) if hasattr(nops, "__len__") else int(nops)
        arr[i].value_bytes = int(value_bytes[i]) if hasattr(value_bytes, "__len__") else int(value_bytes)
    out = make_column(type_id.FLOAT32, max(n, 1) * MLP_FEATURES, stream=stream)
    _check(lib().eg_mlp_features(arr, n, C.c_void_p(out.data), _stream(stream)))
    f = out.to_numpy()[: n * MLP_FEATURES].reshape(n, MLP_FEATURES)
    out.free()
    return f


def mlp_forward(params: np.ndarray, x: np.ndarray, stream=None) -> np.ndarray:
    P, X = _dev_f32(params), _dev_f32(x)
    n = x.shape[0]
    Y = make_column(type_id.FLOAT32, max(n, 1) * 4, stream=stream)
    _check(lib().eg_mlp_forward(C.c_void_p(P.data), C.c_void_p(X.data), n, C.c_void_p(Y.data), _stream(stream)))
    y = Y.to_numpy()[: n * 4].reshape(n, 4)
    for c in (P, X, Y):
        c.free()
    return y


def mlp_loss_grad(params, x, target, card, lam, stream=None):
    P, X, T, Cc = _dev_f32(params), _dev_f32(x), _dev_f32(target), _dev_f32(card)
    Gd = make_column(type_id.FLOAT32, MLP_PARAMS, stream=stream)
    loss = C.c_float()
    _check(lib().eg_mlp_loss_grad(C.c_void_p(P.data), C.c_void_p(X.data), C.c_void_p(T.data), C.c_void_p(Cc.data),
                                  x.shape[0], float(lam), C.c_void_p(Gd.data), C.byref(loss), _stream(stream)))
    g = Gd.to_numpy()
    for c in (P, X, T, Cc, Gd):
        c.free()
    return loss.value, g


def mlp_train(params, x, target, card, epochs=60, batch=1024, lr=2e-3, lam=1.0, weight_decay=0.0, seed=1,
              stream=None):
    """Adam training on the device; returns (params, loss per epoch)."""
    P, X, T, Cc = _dev_f32(params), _dev_f32(x), _dev_f32(target), _dev_f32(card)
    cfg = _lib_mod.MlpTrainConfig(int(epochs), int(batch), float(lr), float(lam), float(weight_decay), int(seed))
    losses = np.zeros(int(epochs), dtype=np.float32)
    _check(lib().eg_mlp_train(C.c_void_p(P.data), C.c_void_p(X.data), C.c_void_p(T.data), C.c_void_p(Cc.data),
                              x.shape[0], C.byref(cfg), losses.ctypes.data_as(C.POINTER(C.c_float)), _stream(stream)))
    out = P.to_numpy().astype(np.float32)
    for c in (P, X, T, Cc):
        c.free()
    return out, losses


def mlp_init(seed=0) -> np.ndarray:
    """He-initialised parameters (host, float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_PARAMS, dtype=np.float32)
    p[:1024] = rng.standard_normal(1024) * np.sqrt(2 / 16)
    p[1088:5184] = rng.standard_normal(4096) * np.sqrt(2 / 64)
    p[5248:5504] = rng.standard_normal(256) * np.sqrt(1 / 64)
    return p


def mlp_load(params: Optional[np.ndarray]):
    """Loads a dispatch model (None unloads); eg_groupby AUTO then uses it."""
    if params is None:
        _check(lib().eg_mlp_load(None, 0))
        return
    p = np.ascontiguousarray(params, dtype=np.float32)
    _check(lib().eg_mlp_load(p.ctypes.data_as(C.c_void_p), p.size))


def mlp_loaded() -> bool:
    return bool(lib().eg_mlp_loaded())


def mlp_select(key: column_view, group_bound=0, nops=2, value_bytes=8, stream=None):
    a, card = C.c_int32(), C.c_int64()
    _check(lib().eg_mlp_select(key._to_c(), int(group_bound), int(nops), int(value_bytes), _stream(stream),
                               C.byref(a), C.byref(card)))
    return groupby_algo(a.value), card.value


def _mlp_queries(keys, group_bounds, nops, value_bytes):
    n = len(keys)
    arr = (_lib_mod.MlpQuery * max(n, 1))()
    for i, k in enumerate(keys):
        arr[i].keys = (k.data + k.offset * size_of(k.type)) if k.data else None
        arr[i].rows = k.size
        arr[i].group_bound = int(group_bounds[i]) if group_bounds is not None else 0
        arr[i].key_type = k.type
        arr[i].nops = int(nops[i]) if hasattr(nops, "__len__") else int(nops)
        arr[i].value_bytes = int(value_bytes[i]) if hasattr(value_bytes, "__len__") else int(value_bytes)
    return arr


def mlp_select_batch(keys: Sequence[column_view], group_bounds=None, nops=2, value_bytes=8, stream=None):
    """The loaded model's strategies and group-count estimates for many keys
    (one feature launch, one forward launch, one synchronisation)."""
    n = len(keys)
    algos = np.zeros(max(n, 1), np.int32)
    card = np.zeros(max(n, 1), np.int64)
    _check(lib().eg_mlp_select_batch(_mlp_queries(keys, group_bounds, nops, value_bytes), n,
                                     algos.ctypes.data_as(C.c_void_p), card.ctypes.data_as(C.c_void_p),
                                     _stream(stream)))
    return algos[:n], card[:n]


def auto_select_groupby_algo_batch(keys: Sequence[column_view], stream=None):
    """The reference rule for many keys with one synchronisation."""
    n = len(keys)
    algos = np.zeros(max(n, 1), np.int32)
    ndv = np.zeros(max(n, 1), np.int64)
    _check(lib().eg_auto_select_algo_batch(_views(keys), n, algos.ctypes.data_as(C.c_void_p),
                                           ndv.ctypes.data_as(C.c_void_p), _stream(stream)))
    return algos[:n], ndv[:n]


DEFAULT_DISPATCH_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "dispatch_mlp.f32")


def load_dispatch_model(path: Optional[str] = None) -> bool:
    """Loads the trained dispatch model (default: the one shipped with the
    package) so that groupby() with AUTO uses it; False if the file is absent."""
    path = path or DEFAULT_DISPATCH_MODEL
    if not os.path.exists(path):
        return False
    p = np.fromfile(path, dtype="<f4")
    if p.size != MLP_PARAMS:
        raise GamaxError(EG_ERR_INVALID, f"{path}: expected {MLP_PARAMS} float32 parameters, found {p.size}")
    mlp_load(p)
    return True
