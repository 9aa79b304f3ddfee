"""Benchmark: single-key group-by throughput on one MI355X.

Headline workload (the strategy sweep the metric is quoted on): 2^28 rows,
int32 keys with distinct count D in {2^4, 2^10, 2^16, 2^22, 2^26}, drawn
uniform (exact counts, shuffled) or Zipf s in {0.5, 1.0, 1.5}; an int64 value
uniform in [-2^31, 2^31) and a float32 value uniform in [0, 1); aggregates
COUNT, SUM/MIN/MAX(int64), SUM(float32).  20 cases.  Inputs are generated on
the device and stay resident in HBM; each case passes
hint.groups_upper_bound = D as the reference's benchmark does
(benchmarks/groupby/groupby_nvbench.cu:57-59).

One step = one group-by of every case through the top-level dispatcher.
value = rows / second over the whole sweep (Grows/s = N / t / 1e9, the
reference's Elem/s).  Algorithmic bytes per case = 16 N + 40 G (each input
column read once, each output written once).

Also measured in the same run (each its own key of the result line):
  sweep[*].ms_by_strategy  every case under forced hash, sort and partition
  partition_heavy          2^30 int64 rows, 2^28 groups, SUM + COUNT
  batched                  4096 uneven queries in one persistent launch
  dispatch                 rule vs learned dispatch: on the sweep, and on a
                           held-out mixed set of queries drawn like the sweep
                           and the batched mix, against the exhaustive best
  cpu_baseline             the oracle's OpenMP group-by on the host's cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--strategy auto]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "group-by Grows/s @256M rows, %HBM peak; learned-vs-fixed dispatch speedup"
STRATS = {"auto": "groupby", "hash": "hash_groupby", "sort": "sort_groupby", "sort_gfur": "sort_groupby_gfur",
          "partition": "partition_groupby", "partition_gfur": "partition_groupby_gfur"}
# Kernel timers of the library (eg_kernel_timing_*) and the kernel functions
# rocprofv3 reports for them.
KERNELS = ("hash_agg", "hash_compact", "hll", "key_range", "radix_hist", "radix_scan", "radix_scatter", "seg_heads",
           "segreduce", "partition_bounds", "partition_agg", "partition_compact", "fill_zero", "mlp_select")
# Retries counted by the library (a table or output sized too small for the
# groups found: that strategy's work is redone inside the call).
RETRIES = ("hash_retry", "partition_retry_table", "partition_retry_capacity")
TIMER_KERNEL = {"partition_agg": ("part_agg_kernel", "eg_jit_part_agg"), "segreduce": ("seg_reduce_kernel",),
                "hash_agg": ("hash_agg_kernel", "eg_jit_hash_agg"),
                "partition_bounds": ("part_bounds_kernel",), "partition_compact": ("part_compact_kernel",)}
# Algorithmic bytes per input row of one launch: the workload's input columns
# read once (int32 key + int64 + float32 values = 16 B), whatever the kernel
# itself moves (a radix pass reads and writes them; that intermediate is not
# counted).  Kernels that read only the key: 4 B.
KERNEL_BYTES_PER_ROW = {"hash_agg": 16, "partition_agg": 16, "segreduce": 16, "radix_scatter": 16,
                        "radix_hist": 4, "seg_heads": 4, "key_range": 4, "hll": 4, "partition_bounds": 4}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--rows", type=int, default=1 << 28)
    p.add_argument("--strategy", default="auto", choices=list(STRATS))
    p.add_argument("--rule-only", action="store_true", help="auto dispatch without the learned model")
    p.add_argument("--cases", default="all", help="'all' or comma list like 16:uniform,1024:zipf1.0")
    p.add_argument("--no-strategies", action="store_true", help="skip the forced-strategy times per case")
    p.add_argument("--no-heavy", action="store_true", help="skip the 2^30-row partition-heavy run")
    p.add_argument("--no-batch", action="store_true", help="skip the batched uneven-query workload")
    p.add_argument("--no-dispatch-eval", action="store_true", help="skip the held-out dispatch evaluation")
    p.add_argument("--dispatch-queries", type=int, default=1000)
    p.add_argument("--no-cpu-baseline", action="store_true")
    return p.parse_args()


def case_list(W, spec):
    cases = W.sweep_cases()
    if spec == "all":
        return cases
    want = []
    for tok in spec.split(","):
        d, dist = tok.split(":")
        s = 0.0 if dist == "uniform" else float(dist.replace("zipf", ""))
        want.append((int(d), "uniform" if s == 0.0 else "zipf", s))
    return want


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the operator runs on one GPU)", file=sys.stderr)
        sys.exit(2)

    from eiger_amd import gamax as G
    from eiger_amd import workloads as W

    A = G.aggregation_id
    n = args.rows
    cases = case_list(W, args.cases)
    fn = getattr(G, STRATS[args.strategy])
    reqs = W.sweep_requests(A)

    # ---- inputs resident in HBM -------------------------------------------------
    t0 = time.time()
    v64, vf = W.sweep_values(G, n)
    keys = [W.sweep_keys(G, d, s, n) for d, _, s in cases]
    log(f"generated {len(cases)} cases x {n} rows in {time.time() - t0:.1f}s")

    def run_case(i, f=fn):
        h = G.groupby_hint(groups_upper_bound=cases[i][0])
        t = f([keys[i]], [v64, vf], reqs, h)
        g = t.num_rows()
        t.free()
        return g

    # ---- copy roofline probe (also the PMC calibration kernel) -------------------
    copy_ms = G.probe_copy(1 << 32, iters=5)
    copy_gbs = 2 * (1 << 32) / (copy_ms * 1e-3) / 1e9
    log(f"copy probe: {copy_gbs:.0f} GB/s (read+write of 4 GiB)")

    # ---- specialised kernels: compiled before anything is timed -------------------
    T = G.type_id
    if G.jit_stats()["mode"] != G.jit_mode.OFF:
        G.jit_prefetch(T.INT32, [T.INT64, T.FLOAT32], reqs)
        G.jit_prefetch(T.INT32, [T.INT64], [(0, A.SUM), (-1, A.COUNT_ALL)])
        G.jit_prefetch(T.INT64, [T.INT64], [(0, A.SUM), (-1, A.COUNT_ALL)])
        G.jit_wait()
    jit0 = G.jit_stats()
    log(f"specialised kernels: {jit0['compiled']} compiled, {jit0['failed']} failed")

    # ---- warmup ----------------------------------------------------------------
    groups = [0] * len(cases)
    for _ in range(max(args.warmup, 1)):
        for i in range(len(cases)):
            groups[i] = run_case(i)

    # ---- dispatch on the sweep: the reference rule, then the learned model -------
    dispatch = None
    learned = args.strategy == "auto" and not args.rule_only and G.load_dispatch_model()
    if args.strategy == "auto":
        G.mlp_load(None)  # the rule of auto_select_groupby_algo
        for i in range(len(cases)):
            run_case(i)
        a = time.perf_counter()
        for _ in range(args.steps):
            for i in range(len(cases)):
                run_case(i)
        rule_value = n * len(cases) * args.steps / (time.perf_counter() - a) / 1e9
        log(f"rule dispatch: {rule_value:.3f} Grows/s")
        if learned:
            G.load_dispatch_model()
            for i in range(len(cases)):
                run_case(i)
        dispatch = {"sweep": {"rule_grows_per_s": round(rule_value, 3), "learned": bool(learned)}}

    # ---- timed steps (the headline) ------------------------------------------------
    G.kernel_timing_reset()
    G.kernel_timing(True)
    jit_timed0 = G.jit_stats()["launches"]
    per_case = [[] for _ in cases]
    t_start = time.perf_counter()
    for k in range(args.steps):
        for i in range(len(cases)):
            a = time.perf_counter()
            groups[i] = run_case(i)
            per_case[i].append(time.perf_counter() - a)
        log(f"step {k + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t_start
    G.kernel_timing(False)
    jit_end = G.jit_stats()
    value = n * len(cases) * args.steps / elapsed / 1e9
    if dispatch is not None:
        sw = dispatch["sweep"]
        sw["learned_grows_per_s" if learned else "timed_grows_per_s"] = round(value, 3)
        if learned:
            sw["learned_vs_rule"] = round(value / sw["rule_grows_per_s"], 3)
    alg_bytes = sum(16 * n + 40 * g for g in groups) * args.steps

    kernels = {}
    for name in KERNELS:
        cnt, ms = G.kernel_time(name)
        if cnt:
            kernels[name] = (cnt, ms)
    roofline = roofline_of(kernels, n, copy_gbs)
    retries = {name: G.kernel_time(name)[0] for name in RETRIES}

    # ---- every case under each forced strategy (not part of the timed steps) --------
    by_strategy = {}
    if not args.no_strategies:
        for name in ("hash", "sort", "partition"):
            f = getattr(G, STRATS[name])
            for i in range(len(cases)):
                run_case(i, f)
                ts = []
                for _ in range(2):
                    a = time.perf_counter()
                    run_case(i, f)
                    ts.append(time.perf_counter() - a)
                by_strategy.setdefault(i, {})[name] = round(min(ts) * 1e3, 3)
            log(f"forced {name}: done")

    sweep = []
    for i, c in enumerate(cases):
        t = statistics.median(per_case[i])
        b = 16 * n + 40 * groups[i]
        rec = {"case": W.case_name(c), "groups": groups[i], "ms": round(t * 1e3, 3),
               "grows_per_s": round(n / t / 1e9, 2), "hbm_frac": round(b / t / PEAK_HBM, 4)}
        if i in by_strategy:
            rec["ms_by_strategy"] = by_strategy[i]
            best = min(by_strategy[i].values())
            rec["auto_over_best_forced"] = round(t * 1e3 / best, 3)
        sweep.append(rec)
        log(f"  {W.case_name(c):24s} G={groups[i]:>9d}  {t * 1e3:8.3f} ms  {n / t / 1e9:7.2f} Grows/s  "
            f"{b / t / PEAK_HBM * 100:5.1f}% HBM  {by_strategy.get(i, '')}")
    for k in (keys):
        k.free()
    v64.free()
    vf.free()

    heavy = None if args.no_heavy else partition_heavy(G)
    batched = None if args.no_batch else batched_workload(G, W)
    if batched:
        log(f"batched: {batched['grows_per_s']} Grows/s, {batched['ms']} ms for {batched['queries']} queries "
            f"({batched['speedup_vs_sequential']}x the sequential calls)")
    if dispatch is not None and learned and not args.no_dispatch_eval:
        dispatch["heldout_mixed"] = dispatch_eval(G, args.dispatch_queries)

    cpu = None if args.no_cpu_baseline else cpu_baseline(W, n)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Grows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (counter-based Philox generator, seed 42, resident in HBM)",
        "config": {"workload": "single-query strategy sweep: 2^28-row int32-key group-by, D in "
                               "{2^4,2^10,2^16,2^22,2^26} x {uniform, zipf 0.5/1.0/1.5}; COUNT, "
                               "SUM/MIN/MAX(int64), SUM(float32)",
                   "rows": n, "cases": len(cases),
                   "strategy": ("auto: learned dispatch" if learned else
                                "auto: rule dispatch" if args.strategy == "auto" else args.strategy),
                   "hbm_frac_algorithmic": round(alg_bytes / elapsed / PEAK_HBM, 4),
                   "specialised_kernels": {"mode": ["off", "async", "sync"][jit_end["mode"]],
                                           "compiled": jit_end["compiled"], "failed": jit_end["failed"],
                                           "launches_timed": jit_end["launches"] - jit_timed0}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernel_launches": {k: {"launches": c, "total_ms": round(m, 3), "avg_ms": round(m / c, 4)}
                            for k, (c, m) in kernels.items()},
        "retries_timed": retries,
        "dispatch": dispatch,
        "partition_heavy": heavy,
        "batched": batched,
        "sweep": sweep,
    }
    G.jit_wait()
    print(json.dumps(out), flush=True)


def roofline_of(kernels, n, copy_gbs):
    """The dominant kernel of the timed steps: algorithmic bytes per launch
    (the rows' 16 input bytes, or 4 for key-only kernels) / mean launch time,
    and the HBM traffic per launch that rocprofv3 counters measured for it."""
    if not kernels:
        return None
    dom = max(kernels, key=lambda k: kernels[k][1])
    cnt, ms = kernels[dom]
    avg_s = ms / cnt * 1e-3
    bpr = KERNEL_BYTES_PER_ROW.get(dom, 16)
    per_launch = bpr * n  # every launch covers the n rows of one case
    achieved = per_launch / avg_s / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tf):
        # measured HBM bytes per launch of this kernel on this workload
        # (rocprofv3 FETCH_SIZE and WRITE_SIZE, scaled by the copy kernel)
        for kname, rec in json.load(open(tf)).get("kernels", {}).items():
            if kname.startswith(TIMER_KERNEL.get(dom, (dom + "_kernel",))):
                traffic = int(rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"])
                break
    log(f"dominant kernel {dom}: {cnt} launches, avg {ms / cnt:.3f} ms, {achieved:.0f} GB/s algorithmic")
    for k, (c, m) in kernels.items():
        log(f"  kernel {k:16s} launches={c:4d} total={m:9.3f} ms")
    return {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM / 1e9,
            "unit": "GB/s", "frac": round(achieved * 1e9 / PEAK_HBM, 4), "traffic": traffic,
            "traffic_over_algorithmic": round(traffic / per_launch, 3) if traffic else None,
            "launches": cnt, "avg_ms": round(ms / cnt, 4), "bytes_per_row": bpr,
            "bytes_per_launch": per_launch, "copy_probe_gbs": round(copy_gbs, 1)}


def partition_heavy(G, n=1 << 30, d=1 << 28, reps=3):
    """2^30 rows of int64 keys, 2^28 distinct values (each exactly four
    times), int64 values; SUM + COUNT with hint.groups_upper_bound = 2^28.
    Algorithmic bytes 16 N + 24 G."""
    A, T = G.aggregation_id, G.type_id
    k = G.gen_keys_uniform(n, d, seed=42, t=T.INT64)
    v = G.gen_values(n, G.val_dist.UNIFORM_I64, T.INT64, seed=42, stream_id=2)
    reqs = [(0, A.SUM), (-1, A.COUNT_ALL)]
    times, groups, algo = [], 0, None
    for r in range(reps + 1):
        h = G.groupby_hint(groups_upper_bound=d)
        a = time.perf_counter()
        t = G.groupby([k], [v], reqs, h)
        dt = time.perf_counter() - a
        groups = t.num_rows()
        t.free()
        algo = int(h.groupby_algo)
        if r:
            times.append(dt)
    k.free()
    v.free()
    t = statistics.median(times)
    log(f"partition-heavy: {t * 1e3:.1f} ms, {groups} groups")
    return {"rows": n, "groups": groups, "ms": round(t * 1e3, 2), "grows_per_s": round(n / t / 1e9, 2),
            "hbm_frac": round((16 * n + 24 * groups) / t / PEAK_HBM, 4), "strategy": algo,
            "runs_ms": [round(x * 1e3, 2) for x in times]}


def batched_workload(G, W, nq=4096):
    """Batched uneven queries: nq independent group-bys (rows log-uniform in
    [1e3, 4e6], distinct count log-uniform in [1, rows/2], int32 keys uniform
    or Zipf s ~ U[0.5, 1.5], int64 values; SUM + COUNT) in one persistent
    launch replayed from a hipGraph, tuned by eg_batch_autotune.  Compared
    with the same queries run one by one through the top-level groupby."""
    import numpy as np

    A, T = G.aggregation_id, G.type_id
    specs = W.mixed_batch_specs(nq)
    t0 = time.time()
    mb = W.MixedBatch(G, specs)
    total = mb.total
    log(f"batched workload: {nq} queries, {total} rows generated in {time.time() - t0:.1f}s")
    reqs = [(0, A.SUM), (-1, A.COUNT_ALL)]
    cands = [G.batch_tuning(tile_rows_log2=17, lds_slots_log2=c, lds_max_bound_log2=b, blocks_per_cu=w, radix_bits=r)
             for c in (10, 11) for b in (10, 12) for w in (1, 2) for r in (0, 10, 12)]
    best, times = G.batch_autotune(mb.queries, T.INT32, T.INT64, reqs, cands, reps=1)
    plan = G.batch_plan(mb.queries, T.INT32, T.INT64, reqs, tuning=best)
    plan.graph_launch()
    ng, st = plan.results()
    reps = 5
    a = time.perf_counter()
    for _ in range(reps):
        plan.graph_launch()
    ng, st = plan.results()
    batch_s = (time.perf_counter() - a) / reps
    # the same queries one by one (top-level groupby, bound hint)
    a = time.perf_counter()
    for q, (rows, d, s) in enumerate(specs):
        t = G.groupby([mb.keys[q]], [mb.queries[q].vals], reqs, G.groupby_hint(groups_upper_bound=d))
        t.free()
    seq_s = time.perf_counter() - a
    bad = int(np.count_nonzero(st))
    mb.free()
    plan.free()
    bt = best
    return {"queries": nq, "rows": total, "grows_per_s": round(total / batch_s / 1e9, 3),
            "ms": round(batch_s * 1e3, 3), "sequential_ms": round(seq_s * 1e3, 1),
            "speedup_vs_sequential": round(seq_s / batch_s, 2), "queries_with_status": bad,
            "groups": int(ng.sum()),
            "algorithmic_hbm_frac": round((12 * total + 20 * int(ng.sum())) / batch_s / PEAK_HBM, 4),
            "tuning": {"tile_rows_log2": bt.tile_rows_log2, "lds_slots_log2": bt.lds_slots_log2,
                       "lds_max_bound_log2": bt.lds_max_bound_log2, "blocks_per_cu": bt.blocks_per_cu,
                       "radix_bits": bt.radix_bits},
            "autotune_candidates": len(cands), "autotune_ms": [round(float(x), 3) for x in times]}


def dispatch_eval(G, nq, seed=3, key_seed0=3_000_000):
    """Held-out mixed queries drawn like the sweep and the batched mix (query
    seed 3; the shipped model was trained on seeds 1 and 2), each run through
    the top-level group-by with the learned model, with the reference rule,
    and under each forced strategy, timed in this run (best of two calls
    each).  Time = the whole synchronous call, dispatch cost included."""
    import numpy as np

    from eiger_amd import dispatch_train as D

    specs = D.draw_queries(nq, seed=seed)
    wl = D.Workload(max(s[1] for s in specs))
    params = np.fromfile(G.DEFAULT_DISPATCH_MODEL, dtype="<f4")
    names = ("hash", "sort", "partition")
    fns = [G.hash_groupby, G.sort_groupby, G.partition_groupby]
    t = np.zeros((nq, 5))  # hash, sort, partition, mlp, rule
    picks = np.zeros((nq, 2), np.int64)
    t0 = time.time()

    def best_of(f, k, vals, reqs, d):
        ts, algo = [], 0
        for _ in range(2):
            h = G.groupby_hint(groups_upper_bound=d)
            a = time.perf_counter()
            f([k], vals, reqs, h).free()
            ts.append(time.perf_counter() - a)
            algo = int(h.groupby_algo)
        return min(ts) * 1e3, algo

    idx = {G.groupby_algo.HASH: 0, G.groupby_algo.SORT: 1, G.groupby_algo.PARTITION: 2}
    for i, spec in enumerate(specs):
        k, vals, reqs, _, _ = wl.query(spec, key_seed0 + i)
        d = spec[2]
        for j, f in enumerate(fns):
            t[i, j], _ = best_of(f, k, vals, reqs, d)
        G.mlp_load(params)
        t[i, 3], a = best_of(G.groupby, k, vals, reqs, d)
        picks[i, 0] = idx.get(a, -1)
        G.mlp_load(None)
        t[i, 4], a = best_of(G.groupby, k, vals, reqs, d)
        picks[i, 1] = idx.get(a, -1)
        k.free()
        if (i + 1) % 250 == 0:
            log(f"dispatch evaluation: {i + 1}/{nq} queries in {time.time() - t0:.0f}s")
    G.mlp_load(params)
    wl.v64.free()
    wl.vf.free()
    best = t[:, :3].min(axis=1)
    arg = t[:, :3].argmin(axis=1)
    tot = t.sum(axis=0)
    rows = np.array([sp[1] for sp in specs], dtype=np.int64)
    by_rows = {}
    for lo, hi, name in ((0, 1 << 17, "rows<2^17"), (1 << 17, 1 << 22, "2^17<=rows<2^22"), (1 << 22, 1 << 62, "rows>=2^22")):
        sel = (rows >= lo) & (rows < hi)
        if sel.any():
            by_rows[name] = {"queries": int(sel.sum()), "best_ms": round(float(best[sel].sum()), 2),
                             "mlp_ms": round(float(t[sel, 3].sum()), 2), "rule_ms": round(float(t[sel, 4].sum()), 2),
                             "fixed_ms": {nm: round(float(t[sel, j].sum()), 2) for j, nm in enumerate(names)},
                             "mlp_picks_best": round(float((picks[sel, 0] == arg[sel]).mean()), 4)}
    return {"queries": nq, "query_seed": seed, "best_ms": round(float(best.sum()), 1), "by_rows": by_rows,
            "mlp_ms": round(float(tot[3]), 1), "rule_ms": round(float(tot[4]), 1),
            "fixed_ms": {nm: round(float(tot[j]), 1) for j, nm in enumerate(names)},
            "mlp_over_best": round(float(tot[3] / best.sum()), 4),
            "rule_over_mlp": round(float(tot[4] / tot[3]), 4),
            "rule_over_best": round(float(tot[4] / best.sum()), 4),
            "mlp_picks_best": round(float((picks[:, 0] == arg).mean()), 4),
            "rule_picks_best": round(float((picks[:, 1] == arg).mean()), 4),
            "mlp_within_10pct_per_query": round(float((t[:, 3] <= 1.1 * best).mean()), 4),
            "seconds": round(time.time() - t0, 1)}


def host_cpu():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        mask = len(os.sched_getaffinity(0))
    except AttributeError:
        mask = os.cpu_count() or 1
    return model, os.cpu_count(), mask


def cpu_baseline(W, n, ds=(1 << 16, 1 << 26), reps=5):
    """The oracle's OpenMP group-by (hash partition + per-partition tables)
    on the full 2^28-row uniform cases at D = 2^16 and 2^26, same inputs and
    aggregates as the sweep, every thread of this process's CPU set
    (OMP_PROC_BIND=close); one warm-up run, then the median of `reps`."""
    model, nproc, threads = host_cpu()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    import numpy as np

    import oracle as O

    v64 = O.gen_values(n, O.VAL_UNIFORM_I64, np.int64, seed=W.SEED, stream_id=2)
    vf = O.gen_values(n, O.VAL_UNIFORM_F32, np.float32, seed=W.SEED, stream_id=3)
    reqs = [(-1, O.COUNT_ALL), (0, O.SUM), (0, O.MIN), (0, O.MAX), (1, O.SUM)]
    per = {}
    for d in ds:
        k = O.gen_keys_uniform(n, d, seed=W.SEED)
        ts = []
        for r in range(reps + 1):
            a = time.perf_counter()
            O.groupby(k, [v64, vf], reqs, parallel=True, nthreads=threads)
            if r:
                ts.append(time.perf_counter() - a)
        per[f"D=2^{d.bit_length() - 1} uniform"] = round(statistics.median(ts) * 1e3, 1)
        log(f"cpu baseline D={d}: median {statistics.median(ts):.2f}s over {reps} runs on {threads} threads")
        del k
    total_ms = sum(per.values())
    return {"value": round(n * len(ds) / (total_ms * 1e-3) / 1e9, 4), "unit": "Grows/s", "cores": threads,
            "kind": "port", "ms_per_case": per, "cpu_model": model, "nproc": nproc,
            "sample": f"full 2^28-row uniform cases at D = 2^16 and 2^26 (same generator, same 5 aggregates); "
                      f"oracle OpenMP hash-partition group-by, {threads} threads, 1 warm-up + median of {reps}"}


if __name__ == "__main__":
    main()
