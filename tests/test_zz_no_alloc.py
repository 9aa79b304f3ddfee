"""eg_groupby_hash_noalloc: the hash strategy with its HBM table in the
library's static device scratch, writing into caller-owned columns.  Checked
against eg_groupby_hash on the same inputs (keys, COUNT and integer
SUM/MIN/MAX bit-exact, float32 SUM within 1e-6 relative) on two calls in a
row, the second one smaller, so groups left in the scratch or the outputs by
the first call would show; and against the oracle.  Capacity limits give
EG_ERR_CAPACITY and leave the outputs untouched."""
import numpy as np
import pytest

import oracle as O
from eiger_amd import gamax as G

pytestmark = pytest.mark.gpu
A, T = G.aggregation_id, G.type_id
REQS = [(-1, A.COUNT_ALL), (0, A.SUM), (0, A.MIN), (0, A.MAX), (1, A.SUM)]


def _inputs(n, d, seed, key_dtype):
    rng = np.random.default_rng(seed)
    info = np.iinfo(key_dtype)
    domain = rng.choice(np.arange(-d * 7, d * 7, 7, dtype=np.int64), d, replace=False).astype(key_dtype)
    domain[:2] = (info.min, info.max)  # the LDS table's empty encoding and the top key
    keys = domain[rng.integers(0, d, n)]
    v64 = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    vf = rng.random(n).astype(np.float32)
    return keys, v64, vf


def _same(got, exp):
    assert len(got) == len(exp)
    for j, (g, e) in enumerate(zip(got, exp)):
        assert g.dtype == e.dtype and g.shape == e.shape, (j, g.dtype, e.dtype, g.shape, e.shape)
        if g.dtype.kind == "f":
            np.testing.assert_allclose(g, e, rtol=1e-6, atol=0, err_msg=f"column {j}")
        else:
            assert np.array_equal(g, e), f"column {j}: {int(np.count_nonzero(g != e))} groups differ"


def _out_columns(key_type, capacity):
    types = [key_type] + [G.agg_output_data_type(k, T.INT64 if c == 0 else T.FLOAT32) if k != A.COUNT_ALL
                          else T.INT64 for c, k in REQS]
    return [G.make_column(t, capacity, clear=True) for t in types]


def test_noalloc_matches_hash_groupby_twice():
    capacity = 8192
    # int32 keys: 2^20 rows, 3000 groups; then 4096 rows, 50 groups
    outs = _out_columns(T.INT32, capacity)
    try:
        for n, d, seed in ((1 << 20, 3000, 1), (4096, 50, 2)):
            keys, v64, vf = _inputs(n, d, seed, np.int32)
            kc, vc, fc = (G.make_column_from_host(a) for a in (keys, v64, vf))
            used_before = G.memory_usage()[0]
            g = G.hash_groupby_noalloc([kc], [vc, fc], REQS, outs, groups_upper_bound=d)
            assert G.memory_usage()[0] == used_before  # nothing from the library's allocator
            assert g == d
            got = [G.copy_column_to_host(G.column_view(c.data, g, c.type)) for c in outs]
            ref = G.hash_groupby([kc], [vc, fc], REQS, G.groupby_hint(groups_upper_bound=d)).to_numpy()
            _same(O.sort_by_key(got), O.sort_by_key(ref))
            _same(O.sort_by_key(got), O.groupby(keys, [v64, vf], REQS))
            for c in (kc, vc, fc):
                c.free()
    finally:
        for c in outs:
            c.free()


def test_noalloc_int64_keys_and_offsets():
    """8-byte keys, output views at an element offset, no bound (the row count)."""
    n, d = 50_000, 700
    keys, v64, vf = _inputs(n, d, 3, np.int64)
    kc, vc, fc = (G.make_column_from_host(a) for a in (keys, v64, vf))
    outs = _out_columns(T.INT64, 1000 + 16)
    try:
        views = [G.column_view(c.data, 1000, c.type, 16) for c in outs]
        g = G.hash_groupby_noalloc([kc], [vc, fc], REQS, views)
        got = [G.copy_column_to_host(G.column_view(c.data, g, c.type, 16)) for c in outs]
        ref = G.hash_groupby([kc], [vc, fc], REQS).to_numpy()
        _same(O.sort_by_key(got), O.sort_by_key(ref))
    finally:
        for c in outs + [kc, vc, fc]:
            c.free()


def test_noalloc_capacity_errors_leave_outputs_untouched():
    n, d = 300_000, 5000
    keys, v64, vf = _inputs(n, d, 4, np.int32)
    kc, vc, fc = (G.make_column_from_host(a) for a in (keys, v64, vf))
    outs = _out_columns(T.INT32, 4000)  # zero-filled
    try:
        # a bound above what the 32 MiB table holds for five states (2^17
        # groups; the bound is first clamped to the 300000 rows)
        with pytest.raises(G.GamaxError) as e:
            G.hash_groupby_noalloc([kc], [vc, fc], REQS, outs, groups_upper_bound=(1 << 17) + 1)
        assert e.value.status == G.EG_ERR_CAPACITY
        # more groups than the outputs hold
        with pytest.raises(G.GamaxError) as e:
            G.hash_groupby_noalloc([kc], [vc, fc], REQS, outs, groups_upper_bound=d)
        assert e.value.status == G.EG_ERR_CAPACITY
        for c in outs:
            assert not c.to_numpy().view(np.uint8).any()  # still all zero bytes
        # an undersized bound still succeeds (retry in the scratch's largest table)
        big = _out_columns(T.INT32, 8192)
        try:
            g = G.hash_groupby_noalloc([kc], [vc, fc], REQS, big, groups_upper_bound=10)
            got = [G.copy_column_to_host(G.column_view(c.data, g, c.type)) for c in big]
            _same(O.sort_by_key(got), O.groupby(keys, [v64, vf], REQS))
        finally:
            for c in big:
                c.free()
    finally:
        for c in outs + [kc, vc, fc]:
            c.free()
