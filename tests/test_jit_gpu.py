"""GPU parity of the run-time specialised aggregation kernels: the hash and
partition strategies with the request's ops compiled in (hipRTC) must give the
same results as the oracle, bit-exact for integer work and within 1e-6
relative for float32 SUM, over fixed and typed-load value layouts, 4- and
8-byte keys, skewed keys (wave combining) and LDS-table spills."""
import numpy as np
import pytest

import oracle as O
from eiger_amd import gamax as G

pytestmark = pytest.mark.gpu

A = G.aggregation_id


@pytest.fixture
def jit_sync():
    before = G.jit_stats()
    G.jit_configure(G.jit_mode.SYNC, 0)
    yield
    G.jit_configure(before["mode"], before["min_rows"])


def run(fn, keys, vals, reqs):
    kc = G.make_column_from_host(keys)
    vcs = [G.make_column_from_host(v) for v in vals]
    t = fn([kc], vcs, reqs, G.groupby_hint())
    out = t.to_numpy()
    t.free()
    order = np.argsort(out[0], kind="stable")
    return [c[order] for c in out]


def assert_same(got, exp):
    assert len(got) == len(exp)
    for j, (g, e) in enumerate(zip(got, exp)):
        assert g.dtype == e.dtype and g.shape == e.shape, (j, g.dtype, e.dtype, g.shape, e.shape)
        if g.dtype.kind == "f":
            np.testing.assert_allclose(g, e, rtol=1e-6, atol=0, err_msg=f"column {j}")
        else:
            np.testing.assert_array_equal(g, e, err_msg=f"column {j}")


def launches():
    return G.jit_stats()["launches"]


BENCH_REQS = [(-1, A.COUNT_ALL), (0, A.SUM), (0, A.MIN), (0, A.MAX), (1, A.SUM)]


@pytest.mark.parametrize("name,fn,distinct", [
    ("hash", G.hash_groupby, 16),
    ("hash", G.hash_groupby, 1000),
    ("hash", G.hash_groupby, 300_000),      # beyond the LDS table: HBM-table rows
    ("partition", G.partition_groupby, 1 << 16),
    ("partition_gfur", G.partition_groupby_gfur, 5000),
])
def test_specialised_int64_f32_requests(jit_sync, name, fn, distinct):
    n = 1 << 20
    keys = O.gen_keys_uniform(n, distinct, seed=5)
    v64 = O.gen_values(n, O.VAL_UNIFORM_I64, np.int64, seed=5, stream_id=2)
    f32 = O.gen_values(n, O.VAL_UNIFORM_F32, np.float32, seed=5, stream_id=3)
    l0 = launches()
    got = run(fn, keys, [v64, f32], BENCH_REQS)
    assert launches() > l0, f"{name}: the specialised kernel did not run"
    assert G.jit_stats()["failed"] == 0
    assert_same(got, O.groupby(keys, [v64, f32], BENCH_REQS))


@pytest.mark.parametrize("fn", [G.hash_groupby, G.partition_groupby])
@pytest.mark.parametrize("s", [1.0, 1.5])
def test_specialised_skewed_keys(jit_sync, fn, s):
    """Zipf keys: many lanes of a wave hit one group (wave-combined updates)."""
    n = 1 << 20
    keys = O.gen_keys_zipf(n, O.zipf_cdf(20_000, s), seed=8)
    vals = O.gen_values(n, O.VAL_UNIFORM_I32_1000, np.int32, seed=8, stream_id=4)
    reqs = [(0, A.SUM), (-1, A.COUNT_ALL), (0, A.MAX)]
    l0 = launches()
    got = run(fn, keys, [vals], reqs)
    assert launches() > l0
    assert_same(got, O.groupby(keys, [vals], reqs))


@pytest.mark.parametrize("algo", [G.groupby_algo.HASH, G.groupby_algo.PARTITION])
def test_specialised_typed_load_layout_int64_keys(jit_sync, algo):
    """Three value columns (typed-load layout), 8-byte signed keys including
    INT64_MIN (the side slot), float MIN/MAX, unsigned SUM, AVG."""
    n = 1 << 19
    rng = np.random.default_rng(3)
    keys = rng.integers(-(1 << 40), 1 << 40, size=3000, dtype=np.int64)[rng.integers(0, 3000, size=n)]
    keys[::97] = np.iinfo(np.int64).min
    f64 = rng.standard_normal(n)
    u32 = rng.integers(0, 1 << 32, size=n, dtype=np.uint32)
    f32 = rng.standard_normal(n).astype(np.float32)
    reqs = [(0, A.MAX), (1, A.SUM), (2, A.MIN), (0, A.AVG), (-1, A.COUNT_ALL)]
    kc = G.make_column_from_host(keys)
    vcs = [G.make_column_from_host(v) for v in (f64, u32, f32)]
    l0 = launches()
    t = G.groupby([kc], vcs, reqs, G.groupby_hint(groupby_algo=algo))
    out = t.to_numpy()
    t.free()
    assert launches() > l0
    order = np.argsort(out[0], kind="stable")
    assert_same([c[order] for c in out], O.groupby(keys, [f64, u32, f32], reqs))


def test_async_mode_switches_to_specialised_kernel():
    """ASYNC: the first call runs the generic kernel while the specialised one
    compiles; after jit_wait the same request runs specialised.  Both agree
    with the oracle."""
    before = G.jit_stats()
    try:
        G.jit_configure(G.jit_mode.ASYNC, 0)
        n = 1 << 18
        keys = O.gen_keys_uniform(n, 700, seed=13)
        vals = O.gen_values(n, O.VAL_UNIFORM_I64, np.int64, seed=13, stream_id=6)
        reqs = [(0, A.MIN), (-1, A.COUNT_ALL), (0, A.SUM), (0, A.MAX)]  # an op order no other test uses
        exp = O.groupby(keys, [vals], reqs)
        assert_same(run(G.hash_groupby, keys, [vals], reqs), exp)
        G.jit_wait()
        assert G.jit_stats()["pending"] == 0
        l0 = launches()
        assert_same(run(G.hash_groupby, keys, [vals], reqs), exp)
        assert launches() > l0
    finally:
        G.jit_configure(before["mode"], before["min_rows"])


def test_off_mode_uses_generic_kernels():
    before = G.jit_stats()
    try:
        G.jit_configure(G.jit_mode.OFF, 0)
        n = 1 << 18
        keys = O.gen_keys_uniform(n, 100, seed=1)
        vals = O.gen_values(n, O.VAL_UNIFORM_I32_1000, np.int32, seed=1, stream_id=4)
        reqs = [(-1, A.COUNT_ALL), (0, A.SUM)]
        l0 = launches()
        assert_same(run(G.hash_groupby, keys, [vals], reqs), O.groupby(keys, [vals], reqs))
        assert launches() == l0
    finally:
        G.jit_configure(before["mode"], before["min_rows"])
