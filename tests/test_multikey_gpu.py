"""Multi-column group keys (key fusion into one 64-bit key, the reference's
SKF path).  The reference's own multi-key fixtures (tests/groupby/
groupby_multikeys_test.cu) under the same strategy/SKF parameter sets
(:1475-1486), random multi-key inputs against pandas' group-by (an
independent implementation: bit-exact for integer results, float32 SUM and
AVG within 1e-6 relative), the error cases, and empty input."""
import numpy as np
import pandas as pd
import pytest

from eiger_amd import gamax as G

pytestmark = pytest.mark.gpu

A = G.aggregation_id
ALGO = G.groupby_algo

# groupby_multikeys_test.cu:1478-1486
PARAMS = [("Hash_NoSKF", ALGO.HASH, False), ("Hash_SKF", ALGO.HASH, True), ("Sort_SKF", ALGO.SORT, True),
          ("SortOpt_SKF", ALGO.SORT_OPT, True), ("SortGFUR_SKF", ALGO.SORT_GFUR, True),
          ("Partition_SKF", ALGO.PARTITION, True), ("PartitionGFUR_SKF", ALGO.PARTITION_GFUR, True)]

I32 = np.int32
# (name, key columns, value columns, requests, expected {key tuple: values})
FIXTURES = [
    ("TwoIntKeyColumns_Sum",  # :54-108
     [np.array([1, 1, 2, 2, 1], I32), np.array([10, 20, 10, 20, 10], I32)],
     [np.array([100, 200, 300, 400, 500], I32)], [(0, A.SUM)],
     {(1, 10): (600,), (1, 20): (200,), (2, 10): (300,), (2, 20): (400,)}),
    ("ThreeIntKeyColumns_Sum",  # :111-171
     [np.array([1, 1, 1, 1, 2, 2], I32), np.array([1, 1, 2, 2, 1, 1], I32), np.array([1, 2, 1, 2, 1, 2], I32)],
     [np.array([10, 20, 30, 40, 50, 60], I32)], [(0, A.SUM)],
     {(1, 1, 1): (10,), (1, 1, 2): (20,), (1, 2, 1): (30,), (1, 2, 2): (40,), (2, 1, 1): (50,), (2, 1, 2): (60,)}),
    ("TwoIntKeyColumns_Max",  # :174-226
     [np.array([1, 1, 1, 2, 2, 2], I32), np.array([1, 1, 2, 1, 2, 2], I32)],
     [np.array([10, 50, 30, 40, 20, 60], I32)], [(0, A.MAX)],
     {(1, 1): (50,), (1, 2): (30,), (2, 1): (40,), (2, 2): (60,)}),
    ("MixedTypeKeys_Int32AndInt64",  # :229-282
     [np.array([1, 1, 2, 2], I32), np.array([100, 200, 100, 200], np.int64)],
     [np.array([10, 20, 30, 40], I32)], [(0, A.SUM)],
     {(1, 100): (10,), (1, 200): (20,), (2, 100): (30,), (2, 200): (40,)}),
    ("TwoKeys_MultipleAggregations",  # :285-346
     [np.array([1, 1, 2, 2, 1], I32), np.array([1, 2, 1, 2, 1], I32)],
     [np.array([10, 20, 30, 40, 50], I32), np.array([5, 15, 25, 35, 45], I32)], [(0, A.SUM), (1, A.SUM)],
     {(1, 1): (60, 50), (1, 2): (20, 15), (2, 1): (30, 25), (2, 2): (40, 35)}),
    ("SingleGroup_AllSameKeys",  # :349-390
     [np.array([1, 1, 1, 1, 1], I32), np.array([100, 100, 100, 100, 100], I32)],
     [np.array([10, 20, 30, 40, 50], I32)], [(0, A.SUM)],
     {(1, 100): (150,)}),
    ("AllUniqueKeys_EachRowIsGroup",  # :393-442
     [np.array([1, 2, 3, 4, 5], I32), np.array([10, 20, 30, 40, 50], I32)],
     [np.array([100, 200, 300, 400, 500], I32)], [(0, A.SUM)],
     {(1, 10): (100,), (2, 20): (200,), (3, 30): (300,), (4, 40): (400,), (5, 50): (500,)}),
    ("NegativeKeyValues",  # :445-491
     [np.array([-1, -1, 1, 1], I32), np.array([-10, 10, -10, 10], I32)],
     [np.array([100, 200, 300, 400], I32)], [(0, A.SUM)],
     {(-1, -10): (100,), (-1, 10): (200,), (1, -10): (300,), (1, 10): (400,)}),
    ("FirstKeyMatchesSecondDiffers",  # :494-541
     [np.array([1, 1, 1, 1, 1, 1], I32), np.array([1, 2, 3, 1, 2, 3], I32)],
     [np.array([10, 20, 30, 40, 50, 60], I32)], [(0, A.SUM)],
     {(1, 1): (50,), (1, 2): (70,), (1, 3): (90,)}),
]


def _key_col(k):
    # numpy bools travel as BOOL8 columns (one byte, read back as uint8)
    if k.dtype == np.bool_:
        return G.make_column_from_host(k.astype(np.uint8), type=G.type_id.BOOL8)
    return G.make_column_from_host(k)


def _out_dtype(k):
    return np.dtype(np.uint8) if k.dtype == np.bool_ else k.dtype


def run(keys, vals, reqs, hint=None, fn=G.groupby):
    kcs = [_key_col(k) for k in keys]
    vcs = [G.make_column_from_host(v) for v in vals]
    t = fn(kcs, vcs, reqs, hint or G.groupby_hint())
    out = t.to_numpy()
    t.free()
    return out


@pytest.mark.parametrize("pname,algo,skf", PARAMS, ids=[p[0] for p in PARAMS])
@pytest.mark.parametrize("case", FIXTURES, ids=[f[0] for f in FIXTURES])
def test_reference_multikey_fixture(case, pname, algo, skf):
    name, keys, vals, reqs, expected = case
    hint = G.groupby_hint(groupby_algo=algo, enable_skf=skf, groups_upper_bound=10)
    out = run(keys, vals, reqs, hint)
    nk = len(keys)
    assert len(out) == nk + len(reqs)
    for j, k in enumerate(keys):
        assert out[j].dtype == k.dtype
    got = {tuple(int(out[j][i]) for j in range(nk)): tuple(int(out[nk + r][i]) for r in range(len(reqs)))
           for i in range(len(out[0]))}
    assert len(got) == len(out[0]) == len(expected)
    assert got == expected
    if algo in (ALGO.SORT, ALGO.SORT_OPT, ALGO.SORT_GFUR):
        # the fused key orders the tuples lexicographically
        tuples = [tuple(int(out[j][i]) for j in range(nk)) for i in range(len(out[0]))]
        assert tuples == sorted(tuples)


def random_case(n, seed):
    rng = np.random.default_rng(seed)
    k1 = rng.integers(-50, 50, n).astype(np.int32)
    k2 = (rng.integers(0, 40, n) * 1_000_000_007 - (1 << 40)).astype(np.int64)
    k3 = rng.integers(0, 3, n).astype(np.uint8)
    v = rng.integers(-(1 << 31), 1 << 31, n).astype(np.int64)
    f = rng.random(n).astype(np.float32)
    return [k1, k2, k3], [v, f]


REQS = [(0, A.SUM), (-1, A.COUNT_ALL), (0, A.MIN), (0, A.MAX), (1, A.SUM), (1, A.AVG)]


def pandas_expected(keys, vals):
    df = pd.DataFrame({"k1": keys[0], "k2": keys[1], "k3": keys[2], "v": vals[0],
                       "f": vals[1].astype(np.float64)})
    g = df.groupby(["k1", "k2", "k3"], sort=True)
    e = pd.DataFrame({"sum": g["v"].sum(), "count": g["v"].count(), "min": g["v"].min(), "max": g["v"].max(),
                      "fsum": g["f"].sum(), "favg": g["f"].mean()}).reset_index()
    return e


@pytest.mark.parametrize("algo,skf", [(ALGO.AUTO_ALGO, False), (ALGO.HASH, False), (ALGO.SORT, True),
                                      (ALGO.PARTITION, True), (ALGO.PARTITION_GFUR, True)])
def test_random_multikey_against_pandas(algo, skf):
    n = 1 << 18
    keys, vals = random_case(n, seed=int(algo) * 7 + 1)
    out = run(keys, vals, REQS, G.groupby_hint(groupby_algo=algo, enable_skf=skf))
    exp = pandas_expected(keys, vals)
    order = np.lexsort((out[2], out[1], out[0]))
    out = [c[order] for c in out]
    assert len(out[0]) == len(exp)
    for j, name in enumerate(["k1", "k2", "k3"]):
        assert out[j].dtype == keys[j].dtype
        np.testing.assert_array_equal(out[j], exp[name].to_numpy())
    np.testing.assert_array_equal(out[3], exp["sum"].to_numpy())
    np.testing.assert_array_equal(out[4], exp["count"].to_numpy())
    np.testing.assert_array_equal(out[5], exp["min"].to_numpy())
    np.testing.assert_array_equal(out[6], exp["max"].to_numpy())
    np.testing.assert_allclose(out[7], exp["fsum"].to_numpy(), rtol=1e-6, atol=0)
    np.testing.assert_allclose(out[8], exp["favg"].to_numpy(), rtol=1e-6, atol=0)


def test_multikey_errors_and_empty_input():
    keys, vals = random_case(1000, seed=3)
    # the sort and partition strategy entry points take one key column
    for fn in (G.sort_groupby, G.partition_groupby):
        with pytest.raises(G.GamaxError):
            run(keys, vals, [(0, A.SUM)], fn=fn)
    # a forced non-hash strategy without key fusion (groupby.cu:287-294)
    with pytest.raises(G.GamaxError):
        run(keys, vals, [(0, A.SUM)], G.groupby_hint(groupby_algo=ALGO.SORT))
    # two full-range int64 columns need 128 bits: folded through dense ranks
    wide = [np.array([np.iinfo(np.int64).min, np.iinfo(np.int64).max] * 4, np.int64)] * 2
    out = run(wide, [np.arange(8, dtype=np.int64)], [(0, A.SUM)], G.groupby_hint(groupby_algo=ALGO.SORT,
                                                                                   enable_skf=True))
    assert [c.tolist() for c in out] == [[np.iinfo(np.int64).min, np.iinfo(np.int64).max],
                                         [np.iinfo(np.int64).min, np.iinfo(np.int64).max], [12, 16]]
    empty = run([k[:0] for k in keys], [v[:0] for v in vals], [(0, A.SUM), (-1, A.COUNT_ALL)])
    assert len(empty) == 5 and all(len(c) == 0 for c in empty)
    assert [c.dtype for c in empty[:3]] == [np.int32, np.int64, np.uint8]


def wide_case(kind, n, seed):
    """Key sets wider than 64 bits after range reduction, drawn from a pool of
    tuples so that groups repeat."""
    rng = np.random.default_rng(seed)
    if kind == "two_int64":      # 128 bits: the prefix and the second column both ranked
        pool = [rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, 700, dtype=np.int64,
                             endpoint=True) for _ in range(2)]
        pool[1][:50] = pool[1][0]  # tuples sharing a second column
    elif kind == "mixed":        # 64 + 32 + 16 + 8 bits of four types
        pool = [rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, 900, dtype=np.int64, endpoint=True),
                rng.integers(0, 1 << 32, 900, dtype=np.uint64).astype(np.uint32),
                rng.integers(-(1 << 15), 1 << 15, 900).astype(np.int16),
                rng.integers(0, 256, 900).astype(np.uint8)]
        pool[0][:300] = pool[0][7]  # tuples sharing the first column
    elif kind == "twelve_narrow":  # more fields than one fusion pass holds, three of them constant
        pool = []
        for c in range(12):
            t = [np.uint8, np.int16, np.bool_, np.int32][c % 4]
            if c in (2, 7, 10):
                col = np.full(600, 1 if t is np.bool_ else -3 if t in (np.int16, np.int32) else 5, t)
            elif t is np.bool_:
                col = rng.integers(0, 2, 600).astype(np.bool_)
            else:
                col = rng.integers(-2 if t in (np.int16, np.int32) else 0, 3, 600).astype(t)
            pool.append(col)
    else:                        # eight full-range uint64 columns: a chain of ranked stages
        pool = [rng.integers(0, np.iinfo(np.uint64).max, 400, dtype=np.uint64, endpoint=True) for _ in range(8)]
        for c in range(1, 8):
            pool[c][:100] = pool[c][0]
    pick = rng.integers(0, len(pool[0]), n)
    keys = [p[pick] for p in pool]
    vals = [rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64), rng.random(n).astype(np.float32)]
    return keys, vals


@pytest.mark.parametrize("algo,skf", [(ALGO.AUTO_ALGO, False), (ALGO.HASH, False), (ALGO.SORT, True),
                                      (ALGO.SORT_GFUR, True), (ALGO.PARTITION, True)])
@pytest.mark.parametrize("kind", ["two_int64", "mixed", "eight_uint64", "twelve_narrow"])
def test_wide_multikey_against_pandas(kind, algo, skf):
    """Keys needing more than 64 bits (the reference's multi-column hash table,
    groupby_hash.cu:212-270): per-group results against pandas, key columns
    decoded in their own types, sort output in ascending tuple order."""
    n = 1 << 17
    keys, vals = wide_case(kind, n, seed=len(kind) * 31 + int(algo))
    reqs = [(0, A.SUM), (-1, A.COUNT_ALL), (0, A.MIN), (1, A.SUM)]
    out = run(keys, vals, reqs, G.groupby_hint(groupby_algo=algo, enable_skf=skf))
    nk = len(keys)
    for j in range(nk):
        assert out[j].dtype == _out_dtype(keys[j])
    got = list(zip(*[c.tolist() for c in out[:nk]]))
    if algo in (ALGO.SORT, ALGO.SORT_GFUR):
        assert got == sorted(got)
    names = [f"k{j}" for j in range(nk)]
    df = pd.DataFrame({**{names[j]: keys[j] for j in range(nk)}, "v": vals[0], "f": vals[1].astype(np.float64)})
    g = df.groupby(names, sort=True)
    exp = pd.DataFrame({"sum": g["v"].sum(), "count": g["v"].count(), "min": g["v"].min(),
                        "fsum": g["f"].sum()}).reset_index()
    order = np.lexsort(tuple(out[j] for j in reversed(range(nk))))
    out = [c[order] for c in out]
    assert len(out[0]) == len(exp)
    for j in range(nk):
        np.testing.assert_array_equal(out[j], exp[names[j]].to_numpy().astype(_out_dtype(keys[j])))
    np.testing.assert_array_equal(out[nk], exp["sum"].to_numpy())
    np.testing.assert_array_equal(out[nk + 1], exp["count"].to_numpy())
    np.testing.assert_array_equal(out[nk + 2], exp["min"].to_numpy())
    np.testing.assert_allclose(out[nk + 3], exp["fsum"].to_numpy(), rtol=1e-6, atol=0)


@pytest.mark.parametrize("algo,skf", [(ALGO.HASH, False), (ALGO.SORT, True), (ALGO.PARTITION, True)])
def test_wide_multikey_many_groups(algo, skf):
    """Two full-range int64 columns, 2^20 rows in about 2^19 groups: the dense
    ranks run over 256 tiles with many groups in each, so every tile's rank
    offset comes from the scan across tiles."""
    rng = np.random.default_rng(77 + int(algo))
    n = 1 << 20
    pool = [rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, 1 << 19, dtype=np.int64, endpoint=True)
            for _ in range(2)]
    pool[0][: 1 << 16] = pool[0][0]  # a first column shared by 2^16 tuples
    pick = rng.integers(0, 1 << 19, n)
    keys = [p[pick] for p in pool]
    vals = [rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)]
    out = run(keys, vals, [(0, A.SUM), (-1, A.COUNT_ALL)], G.groupby_hint(groupby_algo=algo, enable_skf=skf))
    order = np.lexsort((out[1], out[0]))
    if algo == ALGO.SORT:
        np.testing.assert_array_equal(order, np.arange(len(order)))
    out = [c[order] for c in out]
    df = pd.DataFrame({"a": keys[0], "b": keys[1], "v": vals[0]})
    g = df.groupby(["a", "b"], sort=True)["v"]
    exp = pd.DataFrame({"sum": g.sum(), "count": g.count()}).reset_index()
    assert len(out[0]) == len(exp)
    np.testing.assert_array_equal(out[0], exp["a"].to_numpy())
    np.testing.assert_array_equal(out[1], exp["b"].to_numpy())
    np.testing.assert_array_equal(out[2], exp["sum"].to_numpy())
    np.testing.assert_array_equal(out[3], exp["count"].to_numpy())


@pytest.mark.parametrize("algo,skf", [(ALGO.AUTO_ALGO, False), (ALGO.HASH, False), (ALGO.SORT, True),
                                      (ALGO.PARTITION, True)])
def test_float_key_columns_against_pandas(algo, skf):
    """Float key columns among integer ones: -0.0 groups with +0.0 and every
    NaN (any payload) falls in one group, as pandas' group-by with
    dropna=False does; sort output in ascending order, NaN last."""
    rng = np.random.default_rng(int(algo) + 40)
    n = 1 << 16
    other_nan = np.array([0x7fc00123], np.uint32).view(np.float32)[0]
    f32 = np.array([0.0, -0.0, np.nan, other_nan, 1.5, -1.5, np.inf, -np.inf, 3e-39, 7.0], np.float32)
    f64 = np.array([0.0, -0.0, np.nan, 2.5, -2.5, 1e300, -1e-300, 4.0], np.float64)
    keys = [f32[rng.integers(0, len(f32), n)], rng.integers(-3, 3, n).astype(np.int32),
            f64[rng.integers(0, len(f64), n)]]
    vals = [rng.integers(-1000, 1000, n).astype(np.int64)]
    out = run(keys, vals, [(0, A.SUM), (-1, A.COUNT_ALL)], G.groupby_hint(groupby_algo=algo, enable_skf=skf))
    assert [c.dtype for c in out[:3]] == [np.float32, np.int32, np.float64]
    order = np.lexsort((out[2], out[1], out[0]))
    if algo == ALGO.SORT:
        np.testing.assert_array_equal(order, np.arange(len(order)))
    out = [c[order] for c in out]
    df = pd.DataFrame({"a": keys[0], "b": keys[1], "c": keys[2], "v": vals[0]})
    g = df.groupby(["a", "b", "c"], sort=True, dropna=False)["v"]
    exp = pd.DataFrame({"sum": g.sum(), "count": g.count()}).reset_index()
    assert len(out[0]) == len(exp)
    np.testing.assert_array_equal(out[0], exp["a"].to_numpy())
    np.testing.assert_array_equal(out[1], exp["b"].to_numpy())
    np.testing.assert_array_equal(out[2], exp["c"].to_numpy())
    np.testing.assert_array_equal(out[3], exp["sum"].to_numpy())
    np.testing.assert_array_equal(out[4], exp["count"].to_numpy())
    assert not np.any(np.signbit(out[0]) & (out[0] == 0))  # the zero group's key is +0.0


@pytest.mark.parametrize("algo,skf", [(ALGO.HASH, False), (ALGO.SORT, True), (ALGO.PARTITION, True)])
def test_ten_constant_key_columns(algo, skf):
    """Ten constant columns: a full fusion pass of zero-width fields folds
    into a zero-width rank; one group holding every row."""
    n = 5000
    types = [np.int8, np.uint16, np.int32, np.int64, np.uint8, np.bool_, np.uint32, np.int16, np.uint64, np.int64]
    keys = [np.full(n, 1 if t is np.bool_ else 7 * (i + 1), t) for i, t in enumerate(types)]
    vals = [np.arange(n, dtype=np.int64)]
    out = run(keys, vals, [(0, A.SUM), (-1, A.COUNT_ALL)], G.groupby_hint(groupby_algo=algo, enable_skf=skf))
    assert len(out) == 12 and all(len(c) == 1 for c in out)
    for j, k in enumerate(keys):
        assert out[j].dtype == _out_dtype(k) and out[j][0] == k[0]
    assert out[10][0] == n * (n - 1) // 2 and out[11][0] == n


@pytest.mark.parametrize("fn", [G.hash_distinct, G.sort_distinct])
def test_wide_multikey_distinct(fn):
    keys, _ = wide_case("mixed", 1 << 16, seed=5)
    out = fn([_key_col(k) for k in keys], G.groupby_hint(enable_skf=True)).to_numpy()
    assert [c.dtype for c in out] == [_out_dtype(k) for k in keys]
    exp = sorted(set(zip(*[k.tolist() for k in keys])))
    if fn is G.sort_distinct:
        assert list(zip(*[c.tolist() for c in out])) == exp
    else:
        assert _rows(out) == exp


def _rows(out):
    return sorted(zip(*[c.tolist() for c in out]))


@pytest.mark.parametrize("fn,skf", [(G.hash_distinct, False), (G.sort_distinct, False), (G.sort_distinct, True)])
def test_distinct_single_and_multi_column(fn, skf):
    """groupby_hash.hpp:24-32 / groupby_sort.hpp:31-48: one row per unique key
    (tuple), exactly the key columns; sort order is ascending.  Checked
    against numpy's unique."""
    keys, _ = random_case(1 << 17, seed=21)
    single = [keys[1]]
    out = fn([G.make_column_from_host(k) for k in single], G.groupby_hint(enable_skf=skf)).to_numpy()
    assert len(out) == 1 and out[0].dtype == np.int64
    np.testing.assert_array_equal(np.sort(out[0]), np.unique(keys[1]))
    if fn is G.sort_distinct:
        assert np.all(np.diff(out[0]) > 0)
    multi = [G.make_column_from_host(k) for k in keys]
    if fn is G.sort_distinct and not skf:
        with pytest.raises(G.GamaxError):
            fn(multi, G.groupby_hint())
        return
    out = fn(multi, G.groupby_hint(enable_skf=skf)).to_numpy()
    assert [c.dtype for c in out] == [np.int32, np.int64, np.uint8]
    exp = sorted(set(zip(keys[0].tolist(), keys[1].tolist(), keys[2].tolist())))
    assert _rows(out) == exp
    if fn is G.sort_distinct:
        assert list(zip(*[c.tolist() for c in out])) == exp  # ascending tuples


def test_distinct_empty_input():
    for fn in (G.hash_distinct, G.sort_distinct):
        out = fn([G.make_column_from_host(np.zeros(0, np.int32))]).to_numpy()
        assert len(out) == 1 and len(out[0]) == 0 and out[0].dtype == np.int32
