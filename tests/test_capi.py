"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-only logic (output typing, hint defaults, Zipf CDF, argument
validation that runs before any device call) matches the oracle and the
reference's error behaviour.  No kernel is launched here."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from eiger_amd import _lib
from eiger_amd import gamax as G


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = _lib.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_hint_defaults():
    L = _lib.lib()
    assert b"gfx950" in L.eg_version()
    h = _lib.Hint()
    L.eg_hint_init(C.byref(h))
    # groupby.hpp:27-42 defaults
    assert h.groups_upper_bound == (1 << 63) - 1
    assert (h.zero_based_ids, h.optimize_sort, h.groupby_algo, h.agg_count, h.enable_skf,
            h.allow_hash_fallback) == (0, 0, 0, 0, 0, 0)


@pytest.mark.parametrize("kind", [O.SUM, O.MAX, O.MIN, O.AVG, O.ANY, O.COUNT_ALL])
@pytest.mark.parametrize("t", [O.INT8, O.INT16, O.INT32, O.INT64, O.UINT8, O.UINT16, O.UINT32,
                               O.UINT64, O.FLOAT32, O.FLOAT64])
def test_agg_output_type_matches_oracle(kind, t):
    assert int(G.agg_output_data_type(kind, t)) == O.agg_output_type(kind, t)


@pytest.mark.parametrize("d,s", [(16, 0.5), (1024, 1.0), (4096, 1.5), (1, 1.0)])
def test_zipf_cdf_bit_identical_to_oracle(d, s):
    a = G.zipf_cdf(d, s)
    b = O.zipf_cdf(d, s)
    assert a.tobytes() == b.tobytes()


def _cv(ptr=0x1000, n=10, t=G.type_id.INT32):
    return G.column_view(ptr, n, t)


def test_no_key_columns_is_invalid():
    """groupby_hash.cu:32-40 / groupby_sort.cu:208-220: keys are required."""
    with pytest.raises(G.GamaxError) as e:
        G.groupby([], [_cv()], [(0, G.aggregation_id.SUM)])
    assert e.value.status == G.EG_ERR_INVALID


def test_multi_column_key_requests_rejected_before_device_work():
    """Multi-column keys go through key fusion; the sort and partition
    strategy entry points, and a forced non-hash strategy without enable_skf,
    take one key column (groupby_sort.cu:209-211, groupby.cu:287-294).  A key
    column without a type cannot be fused; key columns must have equal
    lengths."""
    two = [_cv(), _cv()]
    for fn in (G.sort_groupby, G.sort_groupby_gfur, G.partition_groupby, G.partition_groupby_gfur):
        with pytest.raises(G.GamaxError) as e:
            fn(two, [_cv()], [(0, G.aggregation_id.SUM)])
        assert e.value.status == G.EG_ERR_UNSUPPORTED
    for algo in (G.groupby_algo.SORT, G.groupby_algo.PARTITION):
        with pytest.raises(G.GamaxError) as e:
            G.groupby(two, [_cv()], [(0, G.aggregation_id.SUM)], G.groupby_hint(groupby_algo=algo))
        assert e.value.status == G.EG_ERR_UNSUPPORTED
    with pytest.raises(G.GamaxError) as e:
        G.groupby([_cv(), _cv(t=G.type_id.EMPTY)], [_cv()], [(0, G.aggregation_id.SUM)])
    assert e.value.status == G.EG_ERR_UNSUPPORTED
    with pytest.raises(G.GamaxError) as e:
        G.groupby([_cv(n=10), _cv(n=11)], [_cv()], [(0, G.aggregation_id.SUM)])
    assert e.value.status == G.EG_ERR_INVALID


def test_dictionary_encode_arguments_rejected_before_device_work():
    """compress/dictionary_encode*.cu: the dictionary-ready form takes
    integral key columns, encodes to INT32 only and needs a dictionary of the
    column's type; the self-built form needs an integral encoded type."""
    def status(fn):
        with pytest.raises(G.GamaxError) as e:
            fn()
        return e.value.status
    f32 = _cv(t=G.type_id.FLOAT32)
    assert status(lambda: G.dictionary_encode_with_dict(f32, f32)) == G.EG_ERR_UNSUPPORTED
    assert status(lambda: G.dictionary_encode_with_dict(_cv(), _cv(), encode_type=G.type_id.INT64)) == \
        G.EG_ERR_UNSUPPORTED
    assert status(lambda: G.dictionary_encode_with_dict(_cv(), _cv(t=G.type_id.INT64))) == G.EG_ERR_INVALID
    assert status(lambda: G.dictionary_encode(_cv(), G.type_id.FLOAT64)) == G.EG_ERR_UNSUPPORTED
    assert status(lambda: G.dictionary_encode_with_dict(_cv(ptr=0), _cv())) == G.EG_ERR_INVALID  # null data


def test_no_requests_is_invalid():
    with pytest.raises(G.GamaxError) as e:
        G.groupby([_cv()], [_cv()], [])
    assert e.value.status == G.EG_ERR_INVALID
    assert "No aggregation" in str(e.value)


def test_float_keys_are_unsupported_by_partition():
    """groupby_partition.cu:556-567: non-integral keys are rejected by the
    partition strategy (forced or through the hint); STRING keys everywhere."""
    for t in (G.type_id.FLOAT32, G.type_id.FLOAT64):
        for fn in (G.partition_groupby, G.partition_groupby_gfur):
            with pytest.raises(G.GamaxError) as e:
                fn([_cv(t=t)], [_cv()], [(0, G.aggregation_id.SUM)])
            assert e.value.status == G.EG_ERR_UNSUPPORTED
        with pytest.raises(G.GamaxError) as e:
            G.groupby([_cv(t=t)], [_cv()], [(0, G.aggregation_id.SUM)],
                      G.groupby_hint(groupby_algo=G.groupby_algo.PARTITION))
        assert e.value.status == G.EG_ERR_UNSUPPORTED
    with pytest.raises(G.GamaxError) as e:
        G.groupby([_cv(t=G.type_id.STRING)], [_cv()], [(0, G.aggregation_id.SUM)])
    assert e.value.status == G.EG_ERR_UNSUPPORTED


def test_bad_column_index_and_size_mismatch():
    with pytest.raises(G.GamaxError) as e:
        G.groupby([_cv()], [_cv()], [(3, G.aggregation_id.SUM)])
    assert e.value.status == G.EG_ERR_INVALID
    with pytest.raises(G.GamaxError) as e:
        G.groupby([_cv(n=10)], [_cv(n=11)], [(0, G.aggregation_id.SUM)])
    assert e.value.status == G.EG_ERR_INVALID


def test_avg_is_rejected_at_the_strategy_level():
    """AVG is rewritten only by the top-level groupby (groupby.cu:121-244)."""
    with pytest.raises(G.GamaxError) as e:
        G.hash_groupby([_cv()], [_cv()], [(0, G.aggregation_id.AVG)])
    assert e.value.status == G.EG_ERR_UNSUPPORTED


def test_unsupported_aggregation_kind():
    with pytest.raises(G.GamaxError) as e:
        G.groupby([_cv()], [_cv()], [(0, 17)])
    assert e.value.status == G.EG_ERR_UNSUPPORTED


def test_string_value_column_is_unsupported():
    with pytest.raises(G.GamaxError):
        G.groupby([_cv()], [_cv(t=G.type_id.STRING)], [(0, G.aggregation_id.SUM)])
