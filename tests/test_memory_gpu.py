"""The library's device memory manager (gamax::MemoryManager,
memory/mem_manager.hpp:12-158): the default caching mode reuses freed blocks,
the slab mode serves allocations from one slab, and group-by results are the
same under every mode."""
import numpy as np
import pytest

import oracle as O
from eiger_amd import gamax as G

pytestmark = pytest.mark.gpu
A = G.aggregation_id
EG_MM_ASYNC, EG_MM_POOL, EG_MM_CACHE = 0, 1, 2


def _groupby_once():
    keys = O.gen_keys_uniform(1 << 18, 5000, seed=3)
    vals = O.gen_values(1 << 18, O.VAL_UNIFORM_I64, np.int64, seed=3, stream_id=2)
    kc, vc = G.make_column_from_host(keys), G.make_column_from_host(vals)
    reqs = [(-1, A.COUNT_ALL), (0, A.SUM)]
    out = []
    for fn in (G.hash_groupby, G.sort_groupby, G.partition_groupby):
        t = fn([kc], [vc], reqs, G.groupby_hint())
        got = t.to_numpy()
        t.free()
        order = np.argsort(got[0], kind="stable")
        out.append([g[order] for g in got])
    kc.free()
    vc.free()
    exp = O.groupby(keys, [vals], reqs)
    for got in out:
        for g, e in zip(got, exp):
            np.testing.assert_array_equal(g, e)


def test_cache_mode_reuses_freed_blocks():
    a = G.make_column(G.type_id.INT64, 100_000)
    p = a.data
    a.free()
    b = G.make_column(G.type_id.INT64, 99_000)  # same size class
    assert b.data == p
    cur, peak = G.memory_usage()
    assert cur >= 99_000 * 8 and peak >= cur
    b.free()
    _groupby_once()


@pytest.mark.parametrize("mode", [EG_MM_POOL, EG_MM_ASYNC])
def test_other_modes_give_the_same_results(mode):
    G.set_memory_mode(mode, 1 << 30)
    try:
        _groupby_once()
    finally:
        G.set_memory_mode(EG_MM_CACHE, 0)
