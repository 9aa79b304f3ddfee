"""Trace generation, training and evaluation of the learned strategy dispatcher.

    python -m eiger_amd.dispatch_train --train 50000 --heldout 10000 --out DIR

(or in parts: --trace-part train:0:25000 ... then --from-traces)

1. Queries are drawn like the two benchmark workloads: (a) batched uneven
   queries -- rows log-uniform in [1e3, 4e6], distinct count log-uniform in
   [1, rows/2], uniform or Zipf s ~ U[0.5, 1.5], SUM(int64) + COUNT; (b)
   strategy-sweep queries -- rows log-uniform in [2^20, 2^23] (one in ten in
   [2^24, 2^28]), distinct count log-uniform in [2^4, 2^26] (capped by rows),
   uniform or Zipf s in {0.5, 1, 1.5}, COUNT + SUM/MIN/MAX(int64) +
   SUM(float32).  Keys are int32, generated on the device.  The aggregation
   kernels specialised for each request list are compiled before timing, as
   a long-running caller would have them.
2. Each query's 16 sample features come from the HIP feature kernel; each
   strategy (hash, sort, partition) is timed on it with the group bound the
   learned dispatcher would size it with (from the true group count), so the
   label is what each strategy costs when the dispatcher picks it.
3. The MLP is trained on the device by eg_mlp_train (HIP forward/backward +
   Adam) against soft targets p_i ~ exp(-(t_i / t_best - 1) / tau) and the
   log2 group count.
4. Held-out evaluation: total time of the strategies each dispatcher picks,
   plus each dispatcher's own cost (the rule's HyperLogLog pass over the whole
   key column; the MLP's sample features + inference), against the
   exhaustive best per query and against fixed single strategies.

Writes DIR/dispatch_mlp.f32 (5508 little-endian float32), DIR/dispatch_eval.json
and DIR/dispatch_trace.npz.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import gamax as G

A = G.aggregation_id
T = G.type_id
STRATS = ("hash", "sort", "partition")
FNS = {"hash": G.hash_groupby, "sort": G.sort_groupby, "partition": G.partition_groupby}
ALGO_INDEX = {G.groupby_algo.HASH: 0, G.groupby_algo.SORT: 1, G.groupby_algo.PARTITION: 2}
REQS_BATCH = [(0, A.SUM), (-1, A.COUNT_ALL)]
REQS_SWEEP = [(-1, A.COUNT_ALL), (0, A.SUM), (0, A.MIN), (0, A.MAX), (1, A.SUM)]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def draw_queries(n: int, seed: int):
    """Query specs: (kind, rows, distinct, zipf_s)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        if rng.random() < 0.5:
            rows = int(math.exp(rng.uniform(math.log(1e3), math.log(4e6))))
            d = int(math.exp(rng.uniform(0.0, math.log(max(rows / 2, 1.0)))))
            s = 0.0 if rng.random() < 0.5 else float(rng.uniform(0.5, 1.5))
            out.append(("batch", rows, max(d, 1), s))
        else:
            rows = int(2 ** (rng.uniform(24, 28) if rng.random() < 0.1 else rng.uniform(20, 23)))
            d = int(2 ** rng.uniform(4, 26))
            d = max(1, min(d, rows))
            s = float(rng.choice([0.0, 0.5, 1.0, 1.5]))
            out.append(("sweep", rows, d, s))
    return out


class Workload:
    """Device-resident value columns shared by every query (prefix views)."""

    def __init__(self, max_rows: int):
        self.v64 = G.gen_values(max_rows, G.val_dist.UNIFORM_I64, T.INT64, seed=42, stream_id=2)
        self.vf = G.gen_values(max_rows, G.val_dist.UNIFORM_F32, T.FLOAT32, seed=42, stream_id=3)

    def query(self, spec, seed: int):
        kind, rows, d, s = spec
        if s == 0.0:
            k = G.gen_keys_uniform(rows, d, seed=seed, t=T.INT32)
        else:
            k = G.gen_keys_zipf(rows, d, s, seed=seed, t=T.INT32)
        if kind == "batch":
            vals = [G.column_view(self.v64.data, rows, T.INT64)]
            reqs, nops, vb = REQS_BATCH, 2, 8
        else:
            vals = [G.column_view(self.v64.data, rows, T.INT64), G.column_view(self.vf.data, rows, T.FLOAT32)]
            reqs, nops, vb = REQS_SWEEP, 5, 12
        return k, vals, reqs, nops, vb


def timed(fn, *args):
    a = time.perf_counter()
    r = fn(*args)
    return r, (time.perf_counter() - a) * 1e3


def trace(specs, seed0: int, wl: Workload, progress_every: int = 500):
    """Features, per-strategy times (ms), group counts, rule choice and the
    two dispatchers' own costs for every query."""
    n = len(specs)
    feats = np.zeros((n, 16), np.float32)
    times = np.zeros((n, 3), np.float64)
    groups = np.zeros(n, np.int64)
    rule = np.zeros(n, np.int32)
    rule_ms = np.zeros(n, np.float64)
    feat_ms = np.zeros(n, np.float64)
    t0 = time.time()
    for i, spec in enumerate(specs):
        k, vals, reqs, nops, vb = wl.query(spec, seed0 + i)
        feats[i] = G.mlp_features([k], None, nops, vb)[0]
        _, feat_ms[i] = timed(G.mlp_select, k, 0, nops, vb)  # the dispatcher's own cost
        h = G.groupby_hint()
        _, rule_ms[i] = timed(G.auto_select_groupby_algo, [k], h)
        rule[i] = ALGO_INDEX.get(h.groupby_algo, 0)
        # the true group count (untimed), then every strategy sized the way the
        # learned dispatcher sizes the one it picks (api.hip: 2.5x the
        # estimate, or the HyperLogLog estimate + 25% when the sample is mostly
        # singletons), with the true count standing in for the estimates
        t = FNS["hash"]([k], vals, reqs, G.groupby_hint(groups_upper_bound=spec[2]))
        true = int(t.num_rows())
        t.free()
        rows = spec[1]
        caller = spec[2] if 0 < spec[2] < rows else rows
        bound = min(caller, 2 * true + true // 2 + 1024)
        if bound < caller and feats[i][7] > 0.25:
            bound = min(caller, true + true // 4 + 1024)
        groups[i] = true
        for j, name in enumerate(STRATS):
            hint = G.groupby_hint(groups_upper_bound=bound)
            t, times[i, j] = timed(FNS[name], [k], vals, reqs, hint)
            t.free()
        k.free()
        if (i + 1) % progress_every == 0:
            log(f"traced {i + 1}/{n} queries in {time.time() - t0:.0f}s")
    return dict(features=feats, times=times, groups=groups, rule=rule, rule_ms=rule_ms, feat_ms=feat_ms)


def soft_targets(times: np.ndarray, tau: float) -> np.ndarray:
    rel = times / times.min(axis=1, keepdims=True) - 1.0
    w = np.exp(-rel / tau)
    return (w / w.sum(axis=1, keepdims=True)).astype(np.float32)


def picks(y: np.ndarray, features: np.ndarray) -> np.ndarray:
    """Strategy index (hash, sort, partition) the library picks from model
    outputs y: the best class score, except that inputs of >= 2^24 rows whose
    estimated group count, or group bound, is <= 2048 go to hash (mlp.hip
    algo_of, bound_picks_hash)."""
    pick = np.argmax(y[:, :3], axis=1)
    rows = np.exp2(features[:, 0].astype(np.float64) * 32) - 1
    card = np.exp2(np.clip(y[:, 3].astype(np.float64) * 32, 0, 62))
    bound = np.exp2(features[:, 11].astype(np.float64) * 32) - 1
    return np.where((rows >= 2.0 ** 24) & ((card <= 2048) | (bound <= 2048.5)), 0, pick)


def choice_ms(params: np.ndarray, tr: dict) -> float:
    y = G.mlp_forward(params, tr["features"])
    pick = picks(y, tr["features"])
    return float(tr["times"][np.arange(len(pick)), pick].sum())


def evaluate(params: np.ndarray, tr: dict, mlp_overhead_ms: float, rule_overhead_ms: float) -> dict:
    """Totals of the strategies each dispatcher picks plus its own cost per
    query (mlp/rule_overhead_ms: per-query dispatch cost)."""
    x, t = tr["features"], tr["times"]
    n = len(x)
    y = G.mlp_forward(params, x)
    pick = picks(y, x)
    best = t.min(axis=1)
    mlp_choice = float(t[np.arange(n), pick].sum())
    rule_choice = float(t[np.arange(n), tr["rule"]].sum())
    mlp_total = mlp_choice + mlp_overhead_ms * n
    rule_total = rule_choice + rule_overhead_ms * n
    return {
        "queries": n,
        "best_ms": float(best.sum()),
        "mlp_ms": mlp_total,
        "rule_ms": rule_total,
        "mlp_choice_only_ms": mlp_choice,
        "rule_choice_only_ms": rule_choice,
        "mlp_overhead_ms_per_query": mlp_overhead_ms,
        "rule_overhead_ms_per_query": rule_overhead_ms,
        "fixed_ms": {s: float(t[:, j].sum()) for j, s in enumerate(STRATS)},
        "mlp_over_best": mlp_total / float(best.sum()),
        "mlp_choice_over_best": mlp_choice / float(best.sum()),
        "rule_over_mlp": rule_total / mlp_total,
        "rule_choice_over_mlp_choice": rule_choice / mlp_choice,
        "mlp_picks_best": float((pick == t.argmin(axis=1)).mean()),
        "rule_picks_best": float((tr["rule"] == t.argmin(axis=1)).mean()),
        "mlp_within_10pct": float((t[np.arange(n), pick] <= 1.1 * best).mean()),
        "card_log2_mae": float(np.abs(y[:, 3] * 32 - np.log2(np.maximum(tr["groups"], 1))).mean()),
    }


def batched_overheads(specs, seed0: int, wl: Workload, chunk: int = 512):
    """Per-query dispatch cost when a batch of queries is dispatched at once
    (the batched workload): the MLP's one feature + forward launch for the
    chunk against the rule's per-column HyperLogLog with one synchronisation.
    Returns (mlp_ms_per_query, rule_ms_per_query)."""
    mlp_ms = rule_ms = 0.0
    done = 0
    for c0 in range(0, len(specs), chunk):
        part = specs[c0:c0 + chunk]
        qs = [wl.query(sp, seed0 + c0 + i) for i, sp in enumerate(part)]
        keys = [q[0] for q in qs]
        nops = [q[3] for q in qs]
        vb = [q[4] for q in qs]
        G.mlp_select_batch(keys[:4], None, nops[:4], vb[:4])  # warm
        G.auto_select_groupby_algo_batch(keys[:4])
        _, a = timed(G.mlp_select_batch, keys, None, nops, vb)
        _, b = timed(G.auto_select_groupby_algo_batch, keys)
        mlp_ms += a
        rule_ms += b
        done += len(part)
        for k in keys:
            k.free()
    return mlp_ms / done, rule_ms / done


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--train", type=int, default=50000)
    p.add_argument("--heldout", type=int, default=10000)
    p.add_argument("--epochs", type=int, default=80)
    p.add_argument("--out", default="dispatch_out")
    p.add_argument("--trace-part", default=None,
                   help="SET:A:B (SET train or heldout): trace queries [A, B) of SET into "
                        "OUT/trace_SET_A_B.npz and stop, so a long trace can span several runs")
    p.add_argument("--from-traces", action="store_true",
                   help="train and evaluate from the OUT/trace_*.npz parts instead of tracing")
    args = p.parse_args()
    os.makedirs(args.out, exist_ok=True)

    specs_tr = draw_queries(args.train, seed=1)
    specs_ho = draw_queries(args.heldout, seed=2)
    seeds = {"train": 1_000_000, "heldout": 2_000_000}
    wl = Workload(max(s[1] for s in specs_tr + specs_ho))
    G.jit_configure(G.jit_mode.SYNC, None)  # specialised kernels compile on first use, before timing
    if not args.from_traces:
        # warm every strategy once, on both request lists at sizes that use the
        # specialised kernels
        warm = [("batch", 1 << 21, 1000, 0.0), ("sweep", 1 << 21, 1000, 0.0), ("sweep", 1 << 21, 1 << 20, 0.0)]
        for spec in warm + specs_tr[:8]:
            k, vals, reqs, _, _ = wl.query(spec, 0)
            for name in STRATS:
                FNS[name]([k], vals, reqs, G.groupby_hint(groups_upper_bound=spec[2])).free()
            k.free()
    G.mlp_load(G.mlp_init(0))  # any weights: the trace times the dispatcher's select path
    if args.trace_part:
        name, a, b = args.trace_part.split(":")
        a, b = int(a), int(b)
        specs = specs_tr if name == "train" else specs_ho
        log(f"tracing {name} queries [{a}, {b})")
        part = trace(specs[a:b], seeds[name] + a, wl)
        np.savez_compressed(os.path.join(args.out, f"trace_{name}_{a}_{b}.npz"), **part)
        return
    if args.from_traces:
        def load(name, n):
            parts = []
            for f in os.listdir(args.out):
                if f.startswith(f"trace_{name}_") and f.endswith(".npz"):
                    a, b = (int(x) for x in f[len(f"trace_{name}_"):-4].split("_"))
                    parts.append((a, b, f))
            parts.sort()
            if not parts or parts[0][0] != 0 or parts[-1][1] != n or any(
                    parts[i][1] != parts[i + 1][0] for i in range(len(parts) - 1)):
                raise SystemExit(f"trace parts of {name} do not cover [0, {n}): {[(a, b) for a, b, _ in parts]}")
            data = [np.load(os.path.join(args.out, f)) for _, _, f in parts]
            return {k: np.concatenate([d[k] for d in data]) for k in data[0].files}
        tr, ho = load("train", len(specs_tr)), load("heldout", len(specs_ho))
    else:
        log("tracing training queries")
        tr = trace(specs_tr, seeds["train"], wl)
        log("tracing held-out queries")
        ho = trace(specs_ho, seeds["heldout"], wl)

    card = (np.log2(np.maximum(tr["groups"], 1)) / 32).astype(np.float32)
    # model selection on a validation split of the training trace: the soft
    # target temperature and the epoch count that pick the fastest strategies
    nfit = int(0.8 * len(card))
    fit = {k: v[:nfit] for k, v in tr.items()}
    val = {k: v[nfit:] for k, v in tr.items()}
    batch = max(64, min(1024, (nfit // 64) * 64))
    grid = []
    for tau in (0.01, 0.03, 0.1):
        for epochs in (args.epochs, 3 * args.epochs):
            p_, _ = G.mlp_train(G.mlp_init(0), fit["features"], soft_targets(fit["times"], tau), card[:nfit],
                                epochs=epochs, batch=batch, lr=2e-3, lam=4.0, weight_decay=1e-5, seed=3)
            grid.append((choice_ms(p_, val), tau, epochs))
            log(f"tau={tau} epochs={epochs}: validation choice time {grid[-1][0]:.1f} ms")
    _, tau, epochs = min(grid)
    target = soft_targets(tr["times"], tau)
    batch = max(64, min(1024, (len(target) // 64) * 64))
    a = time.perf_counter()
    params, losses = G.mlp_train(G.mlp_init(0), tr["features"], target, card, epochs=epochs, batch=batch, lr=2e-3,
                                 lam=4.0, weight_decay=1e-5, seed=3)
    train_s = time.perf_counter() - a
    log(f"trained tau={tau} {epochs} epochs in {train_s:.1f}s: loss {losses[0]:.4f} -> {losses[-1]:.4f}")

    G.mlp_load(params)
    bm, br = batched_overheads(specs_ho[:2048], 2_000_000, wl)
    G.mlp_load(None)
    sm = float(np.median(ho["feat_ms"]))  # single-query select path
    sr = float(np.median(ho["rule_ms"]))  # single-query rule (HyperLogLog + sync)
    ev = {"heldout_batched": evaluate(params, ho, bm, br),
          "heldout_single_query": evaluate(params, ho, sm, sr),
          "train_batched": evaluate(params, tr, bm, br),
          "selection": [{"validation_choice_ms": g[0], "tau": g[1], "epochs": g[2]} for g in grid],
          "epochs": epochs, "tau": tau, "train_seconds": train_s, "train_queries": len(card),
          "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1])}
    log(json.dumps(ev["heldout_batched"], indent=1))
    log(json.dumps(ev["heldout_single_query"], indent=1))
    params.astype("<f4").tofile(os.path.join(args.out, "dispatch_mlp.f32"))
    with open(os.path.join(args.out, "dispatch_eval.json"), "w") as f:
        json.dump(ev, f, indent=1)
    np.savez_compressed(os.path.join(args.out, "dispatch_trace.npz"),
                        **{f"train_{k}": v for k, v in tr.items()}, **{f"heldout_{k}": v for k, v in ho.items()})


if __name__ == "__main__":
    main()
