"""Synthetic workloads shared by the benchmark and the tests: the single-query
strategy sweep (2^28 rows, int32 keys) and the batched uneven-query mix.

Shapes follow the reference benchmark's axes (benchmarks/groupby/
groupby_nvbench.cu:114-122: uniform "i % G then shuffle" keys or Zipf over G
distinct values; data_gen/generator.cuh:345-525), generated on the device by
the counter-based generator (gen.hip), so every array can be regenerated
bit-identically on the host by the oracle's generator.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

SEED = 42
SWEEP_ROWS = 1 << 28
SWEEP_DS = [1 << 4, 1 << 10, 1 << 16, 1 << 22, 1 << 26]
SWEEP_DISTS = [("uniform", 0.0), ("zipf", 0.5), ("zipf", 1.0), ("zipf", 1.5)]


def sweep_requests(A):
    """COUNT, SUM/MIN/MAX(int64 column 0), SUM(float32 column 1)."""
    return [(-1, A.COUNT_ALL), (0, A.SUM), (0, A.MIN), (0, A.MAX), (1, A.SUM)]


def sweep_cases() -> List[Tuple[int, str, float]]:
    return [(d, name, s) for d in SWEEP_DS for (name, s) in SWEEP_DISTS]


def case_name(c) -> str:
    d, name, s = c
    return f"D=2^{d.bit_length() - 1} {name}" + ("" if name == "uniform" else f" s={s}")


def sweep_values(G, n=SWEEP_ROWS, seed=SEED):
    """The int64 value column (uniform in [-2^31, 2^31), stream 2) and the
    float32 value column (uniform in [0, 1), stream 3) on the device."""
    v64 = G.gen_values(n, G.val_dist.UNIFORM_I64, G.type_id.INT64, seed=seed, stream_id=2)
    vf = G.gen_values(n, G.val_dist.UNIFORM_F32, G.type_id.FLOAT32, seed=seed, stream_id=3)
    return v64, vf


def sweep_keys(G, d: int, s: float, n=SWEEP_ROWS, seed=SEED, t=None):
    t = G.type_id.INT32 if t is None else t
    if s == 0.0:
        return G.gen_keys_uniform(n, d, seed=seed, t=t)
    return G.gen_keys_zipf(n, d, s, seed=seed, t=t)


def mixed_batch_specs(nq=4096, seed=SEED) -> List[Tuple[int, int, float]]:
    """(rows, distinct, zipf_s) per query: rows log-uniform in [1e3, 4e6],
    distinct log-uniform in [1, rows/2], half uniform and half Zipf with
    s ~ U[0.5, 1.5]."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nq):
        rows = int(math.exp(rng.uniform(math.log(1e3), math.log(4e6))))
        d = max(1, int(math.exp(rng.uniform(0.0, math.log(max(rows / 2, 1.0))))))
        s = 0.0 if rng.random() < 0.5 else float(rng.uniform(0.5, 1.5))
        specs.append((rows, d, s))
    return specs


class MixedBatch:
    """Device data of a batched mix: one int64 value column holding every
    query's rows back to back (uniform in [-2^31, 2^31), stream 2), each
    query's int32 keys (seed + q), and output columns sized min(rows, d)."""

    def __init__(self, G, specs: Sequence[Tuple[int, int, float]], seed=SEED):
        T = G.type_id
        self.G, self.specs, self.seed = G, list(specs), seed
        self.total = sum(r for r, _, _ in specs)
        self.vals = G.gen_values(self.total, G.val_dist.UNIFORM_I64, T.INT64, seed=seed, stream_id=2)
        self.keys, self.outs, self.queries, self.offsets = [], [], [], []
        off = 0
        for q, (rows, d, s) in enumerate(specs):
            k = G.gen_keys_uniform(rows, d, seed=seed + q, t=T.INT32) if s == 0.0 else \
                G.gen_keys_zipf(rows, d, s, seed=seed + q, t=T.INT32)
            self.keys.append(k)
            cap = min(rows, d)
            ok, o1, o2 = G.make_column(T.INT32, cap), G.make_column(T.INT64, cap), G.make_column(T.INT64, cap)
            self.outs += [ok, o1, o2]
            self.queries.append(G.batch_query(k, G.column_view(self.vals.data + off * 8, rows, T.INT64), d, ok,
                                              [o1, o2]))
            self.offsets.append(off)
            off += rows

    def value_view(self, q):
        rows = self.specs[q][0]
        return self.G.column_view(self.vals.data + self.offsets[q] * 8, rows, self.G.type_id.INT64)

    def free(self):
        for c in self.keys + self.outs:
            c.free()
        self.vals.free()
        self.keys, self.outs = [], []
