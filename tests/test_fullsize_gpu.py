"""Full-size parity: the benchmark workloads at their real sizes, compared per
group against the oracle on the same inputs.

* The strategy sweep at 2^28 rows (int32 keys, D in {2^4, 2^16, 2^26} x
  {uniform, Zipf 1.0, Zipf 1.5}; COUNT, SUM/MIN/MAX(int64), SUM(float32)),
  through the top-level dispatcher and each forced strategy (hash, sort,
  partition).  Keys, COUNT and integer SUM/MIN/MAX bit-exact; the float32
  SUM within 1e-6 relative of the oracle's fp64 sum.  The oracle runs once
  per case (OpenMP group-by on the host) on the very arrays the device
  holds; slices of those arrays are checked against the host generator.
* The 4096-query batched mix in one persistent launch: per-query parity on a
  seeded sample of queries, every query's status clean, and the counts of
  every query summing to its row count.
* Column views with a non-zero element offset (memory/column.hpp:13-30)
  through eg_groupby and the sort and partition sub-operators.
"""
import numpy as np
import pytest

import oracle as O
from eiger_amd import gamax as G
from eiger_amd import workloads as W

pytestmark = pytest.mark.gpu
A = G.aggregation_id
T = G.type_id
N = W.SWEEP_ROWS
REQS = W.sweep_requests(A)
CASES = [(d, s) for d in (1 << 4, 1 << 16, 1 << 26) for s in (0.0, 1.0, 1.5)]
STRATS = {"auto": G.groupby, "hash": G.hash_groupby, "sort": G.sort_groupby, "partition": G.partition_groupby}
PARAMS = [(d, s, name) for (d, s) in CASES for name in STRATS]
CHUNK = 1 << 16


def _assert_same(got, exp):
    assert len(got) == len(exp)
    for j, (g, e) in enumerate(zip(got, exp)):
        assert g.dtype == e.dtype and g.shape == e.shape, (j, g.dtype, e.dtype, g.shape, e.shape)
        if g.dtype.kind == "f":
            np.testing.assert_allclose(g, e, rtol=1e-6, atol=0, err_msg=f"column {j}")
        else:
            assert np.array_equal(g, e), f"column {j}: {int(np.count_nonzero(g != e))} groups differ"


class _Sweep:
    """Device inputs of one case at a time (the values are shared)."""

    def __init__(self):
        self.v64, self.vf = W.sweep_values(G)
        self.hv64, self.hvf = self.v64.to_numpy(), self.vf.to_numpy()
        for b in (0, N // 2 + 12345, N - CHUNK):
            np.testing.assert_array_equal(self.hv64[b:b + CHUNK],
                                          O.gen_values(CHUNK, O.VAL_UNIFORM_I64, np.int64, seed=W.SEED, stream_id=2,
                                                       begin=b))
            np.testing.assert_array_equal(self.hvf[b:b + CHUNK],
                                          O.gen_values(CHUNK, O.VAL_UNIFORM_F32, np.float32, seed=W.SEED, stream_id=3,
                                                       begin=b))
        self.case, self.key, self.expected = None, None, None

    def load(self, d, s):
        if self.case == (d, s):
            return
        if self.key is not None:
            self.key.free()
        self.case, self.expected = None, None
        self.key = W.sweep_keys(G, d, s)
        hk = self.key.to_numpy()
        cdf = O.zipf_cdf(d, s) if s else None
        for b in (0, N // 3 + 7, N - CHUNK):
            exp = (O.gen_keys_uniform(N, d, seed=W.SEED, begin=b, count=CHUNK) if s == 0.0 else
                   O.gen_keys_zipf(CHUNK, cdf, seed=W.SEED, begin=b))
            np.testing.assert_array_equal(hk[b:b + CHUNK], exp)
        self.expected = O.sort_by_key_fast(O.groupby(hk, [self.hv64, self.hvf], REQS, parallel=True))
        self.case = (d, s)

    def free(self):
        for c in (self.key, self.v64, self.vf):
            if c is not None:
                c.free()


@pytest.fixture(scope="module")
def sweep():
    st = _Sweep()
    yield st
    st.free()


@pytest.mark.parametrize("d,s,name", PARAMS, ids=[f"D{d.bit_length() - 1}-s{s}-{n}" for d, s, n in PARAMS])
def test_sweep_full_size_matches_oracle(sweep, d, s, name):
    sweep.load(d, s)
    exp = sweep.expected
    t = STRATS[name]([sweep.key], [sweep.v64, sweep.vf], REQS, G.groupby_hint(groups_upper_bound=d))
    got = t.to_numpy()
    t.free()
    assert len(got[0]) == len(exp[0])
    if name == "sort":
        assert np.all(np.diff(got[0].astype(np.int64)) > 0)  # ascending, one row per key
    _assert_same(O.sort_by_key_fast(got), exp)
    assert int(got[1].sum()) == N


# ------------------------------------------------------------------ batched mix
def test_batched_mix_full_size():
    """The 4096-query mix (rows log-uniform in [1e3, 4e6], about 2e9 rows in
    all) in one persistent launch, replayed from its graph."""
    specs = W.mixed_batch_specs()
    mb = W.MixedBatch(G, specs)
    reqs = [(0, A.SUM), (-1, A.COUNT_ALL)]
    try:
        plan = G.batch_plan(mb.queries, T.INT32, T.INT64, reqs)
        plan.graph_launch()
        ng, st = plan.results()
        assert np.all(st == 0)
        rng = np.random.default_rng(7)
        sample = set(rng.choice(len(specs), 40, replace=False).tolist())
        sample |= {int(np.argmax([r for r, _, _ in specs])), int(np.argmax([d for _, d, _ in specs]))}
        for q, (rows, d, s) in enumerate(specs):
            g = int(ng[q])
            assert 1 <= g <= min(rows, d), q
            qq = mb.queries[q]
            cnt = G.copy_column_to_host(G.column_view(qq.out_aggs[1].data, g, T.INT64))
            assert int(cnt.sum()) == rows, q
            if q not in sample:
                continue
            got = [G.copy_column_to_host(G.column_view(c.data, g, c.type)) for c in [qq.out_keys] + qq.out_aggs]
            k = mb.keys[q].to_numpy()
            v = mb.value_view(q).to_numpy()
            exp = O.sort_by_key_fast(O.groupby(k, [v], reqs, parallel=True))
            _assert_same(O.sort_by_key_fast(got), exp)
        plan.free()
    finally:
        mb.free()


# ------------------------------------------------------------------ column offsets
def test_column_view_offsets_through_groupby_and_subops():
    """gamax::column_view's offset counts elements (memory/column.hpp:13-30):
    a view at offset o of size m covers rows [o, o + m) of its buffer."""
    rng = np.random.default_rng(11)
    n, off = 300_007, 12_345
    keys = rng.integers(-500, 500, n).astype(np.int32)
    v64 = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    vf = rng.random(n).astype(np.float32)
    kc, vc, fc = G.make_column_from_host(keys), G.make_column_from_host(v64), G.make_column_from_host(vf)
    kv, vv, fv = kc.view(off), vc.view(off), fc.view(off)
    assert (kv.size, kv.offset) == (n - off, off)
    hk, h64, hf = keys[off:], v64[off:], vf[off:]
    for fn in STRATS.values():
        got = fn([kv], [vv, fv], REQS, G.groupby_hint()).to_numpy()
        _assert_same(O.sort_by_key(got), O.groupby(hk, [h64, hf], REQS))
    # sort and partition sub-operators on offset views
    sk, sv = G.sort_column_pair(kv, vv)
    ek, ev = O.sort_pairs(hk, h64)
    np.testing.assert_array_equal(sk.to_numpy(), ek)
    np.testing.assert_array_equal(sv.to_numpy(), ev)
    pk, pv = G.partition_column_pair(kv, vv, 0, 6)
    ek, ev = O.sort_pairs(hk, h64, 0, 6)
    np.testing.assert_array_equal(pk.to_numpy(), ek)
    np.testing.assert_array_equal(pv.to_numpy(), ev)
    # a view whose offset leaves rows on both sides
    mid = G.column_view(kc.data, 1000, T.INT32, 777)
    got = G.groupby([mid], [G.column_view(vc.data, 1000, T.INT64, 777)], [(0, A.SUM)], G.groupby_hint()).to_numpy()
    _assert_same(O.sort_by_key(got), O.groupby(keys[777:1777], [v64[777:1777]], [(0, A.SUM)]))
    for c in (kc, vc, fc):
        c.free()
