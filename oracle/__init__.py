"""CPU oracle for the single-key group-by path (ctypes over oracle/_build/libeg_oracle.so).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline.  The product (``eiger_amd``,
``libeiger_groupby.so``, the gamax C++ headers) never imports or links it.

The C sources (``eg_oracle.c``) restate the reference's semantics and cite
the reference file:line for each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libeg_oracle.so")

# gamax::type_id numbering (types/data_type.hpp:12-28)
INT8, INT16, INT32, INT64, BOOL8 = 1, 2, 3, 4, 5
UINT8, UINT16, UINT32, UINT64, FLOAT32, FLOAT64 = 6, 7, 8, 9, 10, 11

# gamax::groupby::aggregation_id (aggregation.hpp:9-22)
SUM, MAX, MIN, AVG, ANY, COUNT_ALL = 0, 1, 2, 3, 4, 5

# gamax::groupby::groupby_algo (groupby.hpp:16-25)
AUTO, HASH, SORT, SORT_OPT, SORT_GFUR, PARTITION, PARTITION_GFUR = range(7)

VAL_UNIFORM_I64, VAL_UNIFORM_F32, VAL_UNIFORM_I32_1000, VAL_ALL_ONES, VAL_SEQ = range(5)

_DT2ID = {
    np.dtype(np.int8): INT8, np.dtype(np.int16): INT16, np.dtype(np.int32): INT32,
    np.dtype(np.int64): INT64, np.dtype(np.uint8): UINT8, np.dtype(np.uint16): UINT16,
    np.dtype(np.uint32): UINT32, np.dtype(np.uint64): UINT64,
    np.dtype(np.float32): FLOAT32, np.dtype(np.float64): FLOAT64,
}
_ID2DT = {v: k for k, v in _DT2ID.items()}

MAX_COLS = 64


class _Col(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_int64), ("type", C.c_int32), ("pad_", C.c_int32)]


class _Req(C.Structure):
    _fields_ = [("column_index", C.c_int64), ("agg_kind", C.c_int32), ("pad_", C.c_int32)]


class _Result(C.Structure):
    _fields_ = [("num_rows", C.c_int64), ("num_cols", C.c_int32),
                ("types", C.c_int32 * MAX_COLS), ("cols", C.c_void_p * MAX_COLS)]


_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, u64, u32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_uint32
        L.orc_perm.restype = u64
        L.orc_perm.argtypes = [u64, u64, u64]
        L.orc_gen_keys_uniform.argtypes = [vp, i32, i64, i64, i64, i64, u64]
        L.orc_zipf_cdf.argtypes = [vp, i64, C.c_double]
        L.orc_gen_keys_zipf.argtypes = [vp, i32, i64, i64, vp, i64, u64]
        L.orc_gen_values.argtypes = [vp, i32, i32, i64, i64, u64, u32]
        L.orc_fmix32.restype = u32
        L.orc_fmix32.argtypes = [u32]
        L.orc_fmix64.restype = u64
        L.orc_fmix64.argtypes = [u64]
        L.orc_partition_ids.argtypes = [vp, i32, i64, u32, vp]
        L.orc_dict_encode.argtypes = [vp, vp, i32, i64, i64, i64, vp]
        L.orc_dict_encode.restype = C.c_int
        L.orc_dict_encode_ordered.argtypes = [vp, i32, i64, vp]
        L.orc_dict_encode_ordered.restype = i64
        L.orc_agg_output_type.restype = i32
        L.orc_agg_output_type.argtypes = [i32, i32]
        L.orc_groupby.restype = C.c_int
        L.orc_groupby.argtypes = [C.POINTER(_Col), C.POINTER(_Col), i32, C.POINTER(_Req), i32, i32, C.POINTER(_Result)]
        L.orc_groupby_parallel.restype = C.c_int
        L.orc_groupby_parallel.argtypes = [C.POINTER(_Col), C.POINTER(_Col), i32, C.POINTER(_Req), i32, i32, i32, C.POINTER(_Result)]
        L.orc_aggregate.restype = C.c_int
        L.orc_aggregate.argtypes = [C.POINTER(_Col), i32, C.POINTER(_Req), i32, C.POINTER(_Result)]
        L.orc_result_free.argtypes = [C.POINTER(_Result)]
        L.orc_last_error.restype = C.c_char_p
        L.orc_sort_pairs.restype = C.c_int
        L.orc_sort_pairs.argtypes = [vp, i32, vp, i32, i64, i32, i32, i32, vp, vp]
        L.orc_partition_offsets.restype = C.c_int
        L.orc_partition_offsets.argtypes = [vp, i32, i64, i32, i32, vp]
        L.orc_hash_partition.restype = C.c_int
        L.orc_hash_partition.argtypes = [vp, i32, i64, i32, i32, vp, vp]
        L.orc_rule_select.restype = i32
        L.orc_rule_select.argtypes = [i32, i64]
        L.orc_count_distinct.restype = i64
        L.orc_count_distinct.argtypes = [vp, i32, i64]
        L.orc_mlp_forward.restype = C.c_int
        L.orc_mlp_forward.argtypes = [vp, vp, i64, vp]
        L.orc_mlp_loss_grad.restype = C.c_int
        L.orc_mlp_loss_grad.argtypes = [vp, vp, vp, vp, i64, C.c_double, vp, vp]
        L.orc_argsort_keys.restype = C.c_int
        L.orc_argsort_keys.argtypes = [vp, i32, i64, i32, vp]
        L.orc_gather.restype = C.c_int
        L.orc_gather.argtypes = [vp, i32, vp, i64, vp, i32]
        L.orc_sample_features.restype = C.c_int
        L.orc_sample_features.argtypes = [vp, i32, i64, i64, i32, i32, vp]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def type_id(dtype) -> int:
    return _DT2ID[np.dtype(dtype)]


def dtype_of(tid: int):
    return _ID2DT[tid]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- generator
def perm(i: int, n: int, seed: int) -> int:
    return int(lib().orc_perm(i, n, seed))


def gen_keys_uniform(n, distinct, seed=42, dtype=np.int32, begin=0, count=None):
    count = n - begin if count is None else count
    out = np.empty(count, dtype=dtype)
    lib().orc_gen_keys_uniform(_ptr(out), type_id(dtype), begin, count, n, distinct, seed)
    return out


def zipf_cdf(distinct, s):
    cdf = np.empty(distinct, dtype=np.float64)
    lib().orc_zipf_cdf(_ptr(cdf), distinct, float(s))
    return cdf


def gen_keys_zipf(n, cdf, seed=42, dtype=np.int32, begin=0):
    cdf = np.ascontiguousarray(cdf, dtype=np.float64)
    out = np.empty(n, dtype=dtype)
    lib().orc_gen_keys_zipf(_ptr(out), type_id(dtype), begin, n, _ptr(cdf), len(cdf), seed)
    return out


def gen_values(n, dist, dtype, seed=42, stream_id=2, begin=0):
    out = np.empty(n, dtype=dtype)
    lib().orc_gen_values(_ptr(out), type_id(dtype), dist, begin, n, seed, stream_id)
    return out


# ---------------------------------------------------------------- hashing
def fmix32(x: int) -> int:
    return int(lib().orc_fmix32(x & 0xFFFFFFFF))


def fmix64(x: int) -> int:
    return int(lib().orc_fmix64(x & 0xFFFFFFFFFFFFFFFF))


def partition_ids(keys: np.ndarray, mask: int) -> np.ndarray:
    keys = np.ascontiguousarray(keys)
    out = np.empty(len(keys), dtype=np.int32)
    lib().orc_partition_ids(_ptr(keys), type_id(keys.dtype), len(keys), mask, _ptr(out))
    return out


def agg_output_type(agg_kind: int, tid: int) -> int:
    return int(lib().orc_agg_output_type(agg_kind, tid))


# ---------------------------------------------------------------- group-by
def _cols(arrays):
    arrays = [np.ascontiguousarray(a) for a in arrays]
    n = max(len(arrays), 1)
    cols = (_Col * n)()
    for i, a in enumerate(arrays):
        cols[i].data = a.ctypes.data
        cols[i].size = len(a)
        cols[i].type = type_id(a.dtype)
    return cols, arrays


def _reqs(requests):
    n = max(len(requests), 1)
    r = (_Req * n)()
    for i, (c, k) in enumerate(requests):
        r[i].column_index = c
        r[i].agg_kind = k
    return r


def _take(res: _Result):
    out = []
    for j in range(res.num_cols):
        dt = _ID2DT[res.types[j]]
        buf = (C.c_char * (max(res.num_rows, 0) * dt.itemsize)).from_address(res.cols[j]) if res.num_rows else b""
        out.append(np.frombuffer(bytes(buf), dtype=dt).copy() if res.num_rows else np.empty(0, dtype=dt))
    lib().orc_result_free(C.byref(res))
    return out


def groupby(key: np.ndarray, vals, requests, agg_count=False, parallel=False, nthreads=None):
    """Brute-force (or OpenMP) group-by; returns [key, agg columns...].

    The brute-force result is sorted by key; the parallel one is not.
    """
    key = np.ascontiguousarray(key)
    kc, _k = _cols([key])
    vc, _v = _cols(list(vals))
    rq = _reqs(requests)
    res = _Result()
    if parallel:
        nt = nthreads or os.cpu_count() or 1
        rc = lib().orc_groupby_parallel(kc, vc, len(vals), rq, len(requests), int(agg_count), nt, C.byref(res))
    else:
        rc = lib().orc_groupby(kc, vc, len(vals), rq, len(requests), int(agg_count), C.byref(res))
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())
    return _take(res)


def aggregate(cols, requests):
    cc, _c = _cols(list(cols))
    rq = _reqs(requests)
    res = _Result()
    rc = lib().orc_aggregate(cc, len(cols), rq, len(requests), C.byref(res))
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())
    return _take(res)


def sort_by_key(cols):
    """Rows of a group-by result ordered by key (column 0), stable."""
    order = np.argsort(cols[0], kind="stable")
    return [c[order] for c in cols]


def _threads(nthreads):
    if nthreads:
        return int(nthreads)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def argsort_keys(keys, nthreads=None) -> np.ndarray:
    """Stable ascending order of an integer key column (parallel LSD radix)."""
    keys = np.ascontiguousarray(keys)
    perm_out = np.empty(len(keys), dtype=np.int64)
    if lib().orc_argsort_keys(_ptr(keys), type_id(keys.dtype), len(keys), _threads(nthreads), _ptr(perm_out)) != 0:
        raise OracleError(lib().orc_last_error().decode())
    return perm_out


def take(a, idx, nthreads=None) -> np.ndarray:
    """a[idx] (parallel gather)."""
    a = np.ascontiguousarray(a)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.empty(len(idx), dtype=a.dtype)
    if lib().orc_gather(_ptr(a), a.dtype.itemsize, _ptr(idx), len(idx), _ptr(out), _threads(nthreads)) != 0:
        raise OracleError(lib().orc_last_error().decode())
    return out


def sort_by_key_fast(cols, nthreads=None):
    """sort_by_key for large results (parallel radix order + gathers)."""
    order = argsort_keys(cols[0], nthreads)
    return [take(c, order, nthreads) for c in cols]


# ---------------------------------------------------------------- sort / partition
def sort_pairs(keys, vals=None, begin_bit=0, end_bit=-1, descending=False):
    keys = np.ascontiguousarray(keys)
    n = len(keys)
    ok = np.empty_like(keys)
    if vals is not None:
        vals = np.ascontiguousarray(vals)
        ov = np.empty_like(vals)
        vp, vb, ovp = _ptr(vals), vals.dtype.itemsize, _ptr(ov)
    else:
        ov, vp, vb, ovp = None, None, 0, None
    rc = lib().orc_sort_pairs(_ptr(keys), type_id(keys.dtype), vp, vb, n, begin_bit, end_bit, int(descending), _ptr(ok), ovp)
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())
    return ok, ov


def partition_offsets(keys, begin_bit, end_bit):
    keys = np.ascontiguousarray(keys)
    out = np.empty((1 << (end_bit - begin_bit)) + 1, dtype=np.int64)
    rc = lib().orc_partition_offsets(_ptr(keys), type_id(keys.dtype), len(keys), begin_bit, end_bit, _ptr(out))
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())
    return out


def hash_partition(keys, bits, shift=0):
    keys = np.ascontiguousarray(keys)
    perm_out = np.empty(len(keys), dtype=np.int64)
    offs = np.empty((1 << bits) + 1, dtype=np.int64)
    rc = lib().orc_hash_partition(_ptr(keys), type_id(keys.dtype), len(keys), bits, shift, _ptr(perm_out), _ptr(offs))
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())
    return perm_out, offs


def rule_select(key_bytes: int, ndv: int) -> int:
    return int(lib().orc_rule_select(key_bytes, ndv))


def count_distinct(keys) -> int:
    keys = np.ascontiguousarray(keys)
    return int(lib().orc_count_distinct(_ptr(keys), type_id(keys.dtype), len(keys)))


INT32_MIN = -(1 << 31)


def dict_encode(keys, dict_keys, miss=INT32_MIN):
    """Position of each key in dict_keys (first one if repeated), `miss` when absent."""
    keys = np.ascontiguousarray(keys)
    dict_keys = np.ascontiguousarray(dict_keys)
    if keys.dtype != dict_keys.dtype:
        raise OracleError("dictionary and column types differ")
    out = np.empty(len(keys), dtype=np.int64)
    rc = lib().orc_dict_encode(_ptr(keys), _ptr(dict_keys), type_id(keys.dtype), len(keys), len(dict_keys), miss,
                               _ptr(out))
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())
    return out


def dict_encode_ordered(keys):
    """(rank of each key among the distinct keys in ascending order, number of distinct keys)."""
    keys = np.ascontiguousarray(keys)
    out = np.empty(len(keys), dtype=np.int64)
    d = int(lib().orc_dict_encode_ordered(_ptr(keys), type_id(keys.dtype), len(keys), _ptr(out)))
    if d < 0:
        raise OracleError("out of memory")
    return out, d


# ---------------------------------------------------------------- learned dispatch
MLP_PARAMS = 5508


def mlp_forward(params, x):
    """fp64 forward pass: params (5508,), x (n, 16) float32 -> (n, 4) float64."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty((x.shape[0], 4), dtype=np.float64)
    lib().orc_mlp_forward(_ptr(p), _ptr(x), x.shape[0], _ptr(y))
    return y


def mlp_loss_grad(params, x, target, card, lam):
    """fp64 mean loss and gradient (see eg_mlp_loss_grad)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(target, dtype=np.float32)
    c = np.ascontiguousarray(card, dtype=np.float32)
    g = np.empty(MLP_PARAMS, dtype=np.float64)
    loss = C.c_double()
    lib().orc_mlp_loss_grad(_ptr(p), _ptr(x), _ptr(t), _ptr(c), x.shape[0], float(lam), _ptr(g), C.byref(loss))
    return loss.value, g


def sample_features(keys, group_bound=0, nops=2, value_bytes=8):
    """The 16 dispatch features of a key column (float32)."""
    keys = np.ascontiguousarray(keys)
    out = np.empty(16, dtype=np.float32)
    rc = lib().orc_sample_features(_ptr(keys), type_id(keys.dtype), len(keys), int(group_bound), int(nops),
                                   int(value_bytes), _ptr(out))
    if rc != 0:
        raise OracleError("sample_features failed")
    return out
