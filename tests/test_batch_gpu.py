"""Batched group-by (eg_batch_*): every query of one persistent launch equals
the oracle's group-by of that query, replays of the captured hipGraph equal
the eager run, and per-query overflow is reported instead of a wrong result."""
import numpy as np
import pytest

import oracle as O
from eiger_amd import gamax as G

pytestmark = pytest.mark.gpu
A = G.aggregation_id
T = G.type_id


def _make_batch(specs, key_dtype=np.int32, reqs=((0, A.SUM), (-1, A.COUNT_ALL)), seed=0):
    """specs: list of (rows, distinct, zipf_s).  Returns host data, queries and
    the device columns that keep them alive."""
    rng = np.random.default_rng(seed)
    host, queries, keep = [], [], []
    out_types = [G.agg_output_data_type(k, T.INT64) if k != A.COUNT_ALL else T.INT64 for _, k in reqs]
    for i, (rows, d, s) in enumerate(specs):
        if s == 0.0:
            k = O.gen_keys_uniform(rows, d, seed=100 + i, dtype=key_dtype) if rows else np.empty(0, key_dtype)
        else:
            k = O.gen_keys_zipf(rows, O.zipf_cdf(d, s), seed=100 + i, dtype=key_dtype)
        v = rng.integers(-(1 << 40), 1 << 40, rows).astype(np.int64)
        kc, vc = G.make_column_from_host(k), G.make_column_from_host(v)
        cap = max(rows, 1)
        ok = G.make_column(G.to_type_id(key_dtype), cap)
        oa = [G.make_column(t, cap) for t in out_types]
        queries.append(G.batch_query(kc, vc, d, ok, oa))
        keep += [kc, vc, ok] + oa
        host.append((k, v))
    return host, queries, keep, list(reqs)


def _check_query(q, ng, k, v, reqs):
    got = [G.copy_column_to_host(G.column_view(c.data, ng, c.type)) for c in [q.out_keys] + q.out_aggs]
    order = np.argsort(got[0], kind="stable")
    got = [g[order] for g in got]
    exp = O.groupby(k, [v], reqs)
    assert len(got[0]) == len(exp[0])
    for g, e in zip(got, exp):
        np.testing.assert_array_equal(g, e)


SPECS = [(1000, 1, 0.0), (1000, 500, 0.0), (50_000, 17, 0.0), (50_000, 3000, 0.0), (200_000, 100_000, 0.0),
         (300_000, 64, 1.2), (120_000, 20_000, 0.8), (70_001, 2048, 0.0), (1, 1, 0.0), (0, 1, 0.0),
         (400_000, 150_000, 0.5), (65_536, 4096, 1.5)]


@pytest.mark.parametrize("radix_bits", [0, 4, 12])
@pytest.mark.parametrize("key_dtype", [np.int32, np.int64])
def test_batch_matches_oracle_per_query(key_dtype, radix_bits):
    """radix_bits 0: queries above the LDS bound use per-row HBM-table
    atomics; 4 and 12: they are partitioned first (4 bits: partitions far
    above half an LDS table, so keys spill to the query's HBM table)."""
    host, queries, keep, reqs = _make_batch(SPECS, key_dtype)
    plan = G.batch_plan(queries, G.to_type_id(key_dtype), T.INT64, reqs, tuning=G.batch_tuning(radix_bits=radix_bits))
    plan.run()
    ng, st = plan.results()
    assert np.all(st == 0)
    for q, n, (k, v) in zip(queries, ng, host):
        _check_query(q, int(n), k, v, reqs)
    plan.free()


def test_batch_rerun_and_graph_replays_equal_eager():
    host, queries, keep, reqs = _make_batch(SPECS[:8], np.int32, reqs=((0, A.SUM), (0, A.MIN), (0, A.MAX),
                                                                         (-1, A.COUNT_ALL)))
    plan = G.batch_plan(queries, T.INT32, T.INT64, reqs, tuning=G.batch_tuning(tile_rows_log2=14))
    plan.run()
    ng0, _ = plan.results()
    for rep in range(3):
        plan.graph_launch()
        ng, st = plan.results()
        assert np.all(st == 0) and np.array_equal(ng, ng0), rep
        for q, n, (k, v) in zip(queries, ng, host):
            _check_query(q, int(n), k, v, reqs)
        plan.run()  # eager work between replays
        ng2, _ = plan.results()
        assert np.array_equal(ng2, ng0)
    plan.free()


def test_batch_reports_table_full_and_capacity():
    rows = 20_000
    k = O.gen_keys_uniform(rows, 5000, seed=1)
    v = np.ones(rows, np.int64)
    kc, vc = G.make_column_from_host(k), G.make_column_from_host(v)
    reqs = [(0, A.SUM)]
    ok_small, oa_small = G.make_column(T.INT32, 10), [G.make_column(T.INT64, 10)]
    ok, oa = G.make_column(T.INT32, rows), [G.make_column(T.INT64, rows)]
    qs = [G.batch_query(kc, vc, 4, ok, oa),             # bound 4: 16-slot table fills
          G.batch_query(kc, vc, 5000, ok_small, oa_small),  # 10 output rows for 5000 groups
          G.batch_query(kc, vc, 5000, ok, oa)]             # fine
    plan = G.batch_plan(qs[2:], T.INT32, T.INT64, reqs)
    plan.run()
    ng, st = plan.results()
    assert st[0] == 0 and ng[0] == 5000
    plan.free()
    plan = G.batch_plan(qs[:2], T.INT32, T.INT64, reqs)
    plan.run()
    ng, st = plan.results()
    assert st[0] & 1 and st[1] == 2
    plan.free()


def test_batch_autotune_picks_a_candidate():
    host, queries, keep, reqs = _make_batch(SPECS[:6], np.int32)
    cands = [G.batch_tuning(tile_rows_log2=t, lds_slots_log2=c) for t in (14, 16) for c in (10, 11)]
    best, times = G.batch_autotune(queries, T.INT32, T.INT64, reqs, cands, reps=1)
    assert np.all(times > 0)
    i = int(np.argmin(times))
    assert (best.tile_rows_log2, best.lds_slots_log2) == (cands[i].tile_rows_log2, cands[i].lds_slots_log2)
