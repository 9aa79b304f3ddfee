"""Key columns of 1- and 2-byte integers, BOOL8, FLOAT32 and FLOAT64: the
group-by (every strategy that takes them), distinct and the self-built
dictionary encoding, against the oracle.  Floats group by operator==
(memory/row_utils.cuh:16-21: -0.0 meets +0.0) with every NaN in one group;
partition takes integral keys only (groupby_partition.cu:556-567)."""
import numpy as np
import pytest

import oracle as O
from eiger_amd import gamax as G

pytestmark = pytest.mark.gpu
A, T = G.aggregation_id, G.type_id
REQS = [(-1, A.COUNT_ALL), (0, A.SUM), (0, A.MIN), (0, A.MAX), (1, A.SUM)]
TYPES = [T.INT8, T.INT16, T.UINT8, T.UINT16, T.BOOL8, T.FLOAT32, T.FLOAT64]


def _keys(t, n, rng):
    dt = G.to_numpy_dtype(t)
    if t == T.BOOL8:
        return rng.integers(0, 2, n).astype(np.uint8)
    if dt.kind == "f":
        pool = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1.5, -1.5, 1e-30, -2.0, 3.25,
                         np.finfo(dt).max, np.finfo(dt).min, np.finfo(dt).tiny], dtype=dt)
        pool = np.concatenate([pool, rng.standard_normal(2000).astype(dt)])
        return pool[rng.integers(0, len(pool), n)]
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, n, endpoint=True).astype(dt)


def _same(got, exp):
    assert len(got) == len(exp)
    np.testing.assert_array_equal(got[0], exp[0])  # keys: numeric equality, NaN == NaN
    for j, (g, e) in enumerate(zip(got[1:], exp[1:]), 1):
        assert g.dtype == e.dtype and g.shape == e.shape, (j, g.dtype, e.dtype)
        if g.dtype.kind == "f":
            np.testing.assert_allclose(g, e, rtol=1e-6, atol=0, err_msg=f"column {j}")
        else:
            assert np.array_equal(g, e), f"column {j}"


def _sorted(cols):
    """Rows in the oracle's order: ascending key, -0.0 with +0.0, NaN last."""
    order = np.argsort(cols[0], kind="stable")  # one zero group at most; NaNs sort last
    return [c[order] for c in cols]


STRATS = {"auto": G.groupby, "hash": G.hash_groupby, "sort": G.sort_groupby, "partition": G.partition_groupby}


@pytest.mark.parametrize("t", TYPES, ids=lambda t: t.name)
@pytest.mark.parametrize("name", list(STRATS))
def test_groupby_key_types_match_oracle(t, name):
    rng = np.random.default_rng(int(t) * 7 + len(name))
    n = 200_000
    keys = _keys(t, n, rng)
    v64 = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    vf = rng.random(n).astype(np.float32)
    kc = G.make_column_from_host(keys, type=t)
    vc, fc = G.make_column_from_host(v64), G.make_column_from_host(vf)
    try:
        if name == "partition" and t in (T.FLOAT32, T.FLOAT64):
            with pytest.raises(G.GamaxError) as e:
                STRATS[name]([kc], [vc, fc], REQS)
            assert e.value.status == G.EG_ERR_UNSUPPORTED
            return
        res = STRATS[name]([kc], [vc, fc], REQS)
        assert res[0].type == t
        got = res.to_numpy()
        res.free()
        exp = O.groupby(keys, [v64, vf], REQS)
        if name == "sort":
            _same(got, exp)  # ascending key order, NaN last
        else:
            _same(_sorted(got), exp)
        if G.to_numpy_dtype(t).kind == "f":
            assert np.count_nonzero(got[0] == 0) == 1 and np.count_nonzero(np.isnan(got[0])) == 1
    finally:
        for c in (kc, vc, fc):
            c.free()


@pytest.mark.parametrize("t", TYPES, ids=lambda t: t.name)
def test_distinct_key_types(t):
    rng = np.random.default_rng(int(t))
    keys = _keys(t, 50_000, rng)
    kc = G.make_column_from_host(keys, type=t)
    try:
        exp = O.groupby(keys, [], [(-1, A.COUNT_ALL)])[0]
        s = G.sort_distinct([kc])
        np.testing.assert_array_equal(s[0].to_numpy(), exp)
        s.free()
        h = G.hash_distinct([kc])
        np.testing.assert_array_equal(_sorted([h[0].to_numpy()])[0], exp)
        h.free()
    finally:
        kc.free()


@pytest.mark.parametrize("t", [T.FLOAT32, T.FLOAT64, T.INT8, T.UINT16], ids=lambda t: t.name)
@pytest.mark.parametrize("ordered", [True, False])
def test_dictionary_encode_float_and_narrow_keys(t, ordered):
    """dictionary_encode / _with_order over float columns (the reference's
    FLOAT64_TO_INT32 and FLOAT32_TO_INT16 cases,
    tests/compress/dictionary_encode_test.cu:139-145,276): the dictionary
    holds each distinct key once; decoding the codes gives the column back."""
    rng = np.random.default_rng(31 + int(t))
    keys = _keys(t, 4096, rng)
    kc = G.make_column_from_host(keys, type=t)
    try:
        enc, size, dct = G.dictionary_encode(kc, T.INT32, with_order=ordered, return_dictionary=True)
        codes, d = enc.to_numpy(), dct.to_numpy()
        exp = O.groupby(keys, [], [(-1, A.COUNT_ALL)])[0]
        assert size == len(exp) == len(d)
        if ordered:
            np.testing.assert_array_equal(d, exp)  # ascending, NaN last
        else:
            np.testing.assert_array_equal(_sorted([d])[0], exp)
        assert codes.min() >= 0 and codes.max() < size
        back = d[codes]
        np.testing.assert_array_equal(back, keys)  # -0.0 decodes to +0.0 (==), NaN to NaN
        for c in (enc, dct):
            c.free()
    finally:
        kc.free()
