"""Benchmark: single-key group-by throughput on one MI355X.

Workload (the strategy sweep the headline metric is quoted on): 2^28 rows,
int32 keys with distinct count D in {2^4, 2^10, 2^16, 2^22, 2^26}, drawn
uniform (exact counts, shuffled) or Zipf s in {0.5, 1.0, 1.5}; an int64 value
uniform in [-2^31, 2^31) and a float32 value uniform in [0, 1); aggregates
COUNT, SUM/MIN/MAX(int64), SUM(float32).  20 cases.  Inputs are generated on
the device and stay resident in HBM; each case passes
hint.groups_upper_bound = D as the reference's benchmark does
(benchmarks/groupby/groupby_nvbench.cu:57-59).

One step = one group-by of every case.  value = rows / second over the whole
sweep (Grows/s = N / t / 1e9, the reference's Elem/s).  Algorithmic bytes per
case = 16 N + 40 G (each input column read once, each output written once).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--strategy auto]
    --strategy auto   the library's dispatcher (HyperLogLog estimate + rule)
    --strategy hash|sort|sort_gfur|partition|partition_gfur   one strategy
    --explore         also time every strategy once per case (stderr + sweep)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "group-by Grows/s @256M rows, %HBM peak; learned-vs-fixed dispatch speedup"
DS = [1 << 4, 1 << 10, 1 << 16, 1 << 22, 1 << 26]
DISTS = [("uniform", 0.0), ("zipf", 0.5), ("zipf", 1.0), ("zipf", 1.5)]
SEED = 42
STRATS = {"auto": "groupby", "hash": "hash_groupby", "sort": "sort_groupby", "sort_gfur": "sort_groupby_gfur",
          "partition": "partition_groupby", "partition_gfur": "partition_groupby_gfur"}
# Algorithmic HBM bytes per row of one launch of each timed kernel (int32 key,
# int64 + float32 values): inputs read once; a radix scatter reads and writes
# the key and both values once (2 x 16 B).
KERNELS = ("hash_agg", "hash_compact", "hll", "key_range", "radix_hist", "radix_scan", "radix_scatter", "seg_heads",
           "segreduce", "partition_bounds", "partition_agg", "partition_compact", "fill_zero", "mlp_select")
# timer name -> kernel function name (rocprofv3 reports the function)
TIMER_KERNEL = {"partition_agg": ("part_agg_kernel", "eg_jit_part_agg"), "segreduce": ("seg_reduce_kernel",),
                "hash_agg": ("hash_agg_kernel", "eg_jit_hash_agg"),
                "partition_bounds": ("part_bounds_kernel",), "partition_compact": ("part_compact_kernel",)}
KERNEL_BYTES_PER_ROW = {"hash_agg": 16, "partition_agg": 16, "segreduce": 16, "radix_scatter": 32,
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
    p.add_argument("--explore", action="store_true")
    p.add_argument("--rule-only", action="store_true", help="auto dispatch without the learned model")
    p.add_argument("--no-batch", action="store_true", help="skip the batched uneven-query workload")
    p.add_argument("--cases", default="all", help="'all' or comma list like 16:uniform,1024:zipf1.0")
    p.add_argument("--cpu-sample-rows", type=int, default=1 << 22)
    p.add_argument("--no-cpu-baseline", action="store_true")
    return p.parse_args()


def case_list(spec):
    cases = [(d, name, s) for d in DS for (name, s) in DISTS]
    if spec == "all":
        return cases
    want = []
    for tok in spec.split(","):
        d, dist = tok.split(":")
        s = 0.0 if dist == "uniform" else float(dist.replace("zipf", ""))
        want.append((int(d), "uniform" if s == 0.0 else "zipf", s))
    return want


def case_name(c):
    d, name, s = c
    return f"D=2^{d.bit_length() - 1} {name}" + ("" if name == "uniform" else f" s={s}")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the operator runs on one GPU)", file=sys.stderr)
        sys.exit(2)

    from eiger_amd import gamax as G

    A = G.aggregation_id
    n = args.rows
    cases = case_list(args.cases)
    fn = getattr(G, STRATS[args.strategy])

    # ---- inputs resident in HBM -------------------------------------------------
    t0 = time.time()
    v64 = G.gen_values(n, G.val_dist.UNIFORM_I64, G.type_id.INT64, seed=SEED, stream_id=2)
    vf = G.gen_values(n, G.val_dist.UNIFORM_F32, G.type_id.FLOAT32, seed=SEED, stream_id=3)
    keys = []
    for c in cases:
        d, name, s = c
        if name == "uniform":
            keys.append(G.gen_keys_uniform(n, d, seed=SEED, t=G.type_id.INT32))
        else:
            keys.append(G.gen_keys_zipf(n, d, s, seed=SEED, t=G.type_id.INT32))
    log(f"generated {len(cases)} cases x {n} rows in {time.time() - t0:.1f}s")
    reqs = [(-1, A.COUNT_ALL), (0, A.SUM), (0, A.MIN), (0, A.MAX), (1, A.SUM)]

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

    # ---- warmup + timed steps ----------------------------------------------------
    groups = [0] * len(cases)
    for w in range(args.warmup):
        for i in range(len(cases)):
            groups[i] = run_case(i)
    # The aggregation kernels specialised for this request list compile in the
    # background on first use; let both (hash and partition) finish and load
    # before anything is timed.
    jit0 = G.jit_stats()
    if jit0["mode"] != G.jit_mode.OFF:
        for f in (G.hash_groupby, G.partition_groupby):
            run_case(0, f)
        G.jit_wait()
        for i in range(len(cases)):
            run_case(i)
        for f in (G.hash_groupby, G.partition_groupby):
            run_case(0, f)
    jit_warm = G.jit_stats()
    log(f"specialised kernels: {jit_warm['compiled']} compiled, {jit_warm['failed']} failed")
    explore = {}
    if args.explore:
        for name in ("hash", "sort", "sort_gfur", "partition", "partition_gfur"):
            f = getattr(G, STRATS[name])
            for i in range(len(cases)):
                run_case(i, f)  # the operator returns after its stream is synchronised
                a = time.perf_counter()
                run_case(i, f)
                explore.setdefault(i, {})[name] = round((time.perf_counter() - a) * 1e3, 3)
            log(f"explored {name}")
    # ---- dispatch: the learned model (when shipped) against the reference rule ------
    dispatch = None
    learned = args.strategy == "auto" and not args.rule_only and G.load_dispatch_model()
    if args.strategy == "auto":
        G.mlp_load(None)  # the rule of auto_select_groupby_algo
        for i in range(len(cases)):
            run_case(i)
        rule_t = time.perf_counter()
        for k in range(args.steps):
            for i in range(len(cases)):
                run_case(i)
        rule_elapsed = time.perf_counter() - rule_t
        rule_value = n * len(cases) * args.steps / rule_elapsed / 1e9
        log(f"rule dispatch: {rule_value:.3f} Grows/s")
        if learned:
            G.load_dispatch_model()
            for i in range(len(cases)):
                run_case(i)
        dispatch = {"rule_grows_per_s": round(rule_value, 3), "learned": bool(learned)}

    G.kernel_timing_reset()
    G.kernel_timing(True)
    jit_timed0 = G.jit_stats()["launches"]
    per_case = [0.0] * len(cases)
    t_start = time.perf_counter()
    for k in range(args.steps):
        for i in range(len(cases)):
            a = time.perf_counter()
            groups[i] = run_case(i)
            per_case[i] += time.perf_counter() - a
        log(f"step {k + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t_start
    G.kernel_timing(False)
    jit_end = G.jit_stats()
    if dispatch is not None:
        v = n * len(cases) * args.steps / elapsed / 1e9
        dispatch["learned_grows_per_s" if learned else "timed_grows_per_s"] = round(v, 3)
        if learned:
            dispatch["learned_vs_rule"] = round(v / dispatch["rule_grows_per_s"], 3)
            ev = os.path.join(ROOT, "eiger_amd", "data", "dispatch_eval.json")
            if os.path.exists(ev):
                e = json.load(open(ev)).get("heldout_batched", {})
                dispatch["heldout_mixed_queries"] = {k: e.get(k) for k in (
                    "queries", "mlp_over_best", "rule_over_mlp", "mlp_picks_best", "rule_picks_best")}

    rows = n * len(cases) * args.steps
    value = rows / elapsed / 1e9
    alg_bytes = sum(16 * n + 40 * g for g in groups) * args.steps
    sweep = []
    for i, c in enumerate(cases):
        t = per_case[i] / args.steps
        b = 16 * n + 40 * groups[i]
        sweep.append({"case": case_name(c), "groups": groups[i], "ms": round(t * 1e3, 3),
                      "grows_per_s": round(n / t / 1e9, 2), "hbm_frac": round(b / t / PEAK_HBM, 4)})
        if i in explore:
            sweep[-1]["ms_by_strategy"] = explore[i]
        log(f"  {case_name(c):24s} G={groups[i]:>9d}  {t * 1e3:8.3f} ms  {n / t / 1e9:7.2f} Grows/s  "
            f"{b / t / PEAK_HBM * 100:5.1f}% HBM" + (f"  {explore[i]}" if i in explore else ""))

    # ---- dominant kernel roofline --------------------------------------------------
    kernels = {}
    for name in KERNELS:
        cnt, ms = G.kernel_time(name)
        if cnt:
            kernels[name] = (cnt, ms)
    roofline = None
    if kernels:
        dom = max(kernels, key=lambda k: kernels[k][1])
        cnt, ms = kernels[dom]
        avg_s = ms / cnt * 1e-3
        per_launch = KERNEL_BYTES_PER_ROW.get(dom, 16) * n  # every launch covers the n rows of one case
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
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM / 1e9,
                    "unit": "GB/s", "frac": round(achieved * 1e9 / PEAK_HBM, 4), "traffic": traffic,
                    "launches": cnt, "avg_ms": round(ms / cnt, 4), "bytes_per_row": KERNEL_BYTES_PER_ROW.get(dom, 16),
                    "bytes_per_launch": per_launch, "copy_probe_gbs": round(copy_gbs, 1)}
        log(f"dominant kernel {dom}: {cnt} launches, avg {ms / cnt:.3f} ms, {achieved:.0f} GB/s algorithmic")
        for k, (c, m) in kernels.items():
            log(f"  kernel {k:16s} launches={c:4d} total={m:9.3f} ms")

    batched = None
    if not args.no_batch:
        batched = batched_workload(G)
        log(f"batched: {batched['grows_per_s']} Grows/s, {batched['ms']} ms for {batched['queries']} queries "
            f"({batched['speedup_vs_sequential']}x the sequential calls)")

    # ---- CPU baseline (oracle, OpenMP) on a bounded sample -------------------------
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cases, n, args.cpu_sample_rows)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Grows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (counter-based Philox generator, seed 42, resident in HBM)",
        "config": {"workload": "single-query strategy sweep: 2^28-row int32-key group-by, D in "
                               "{2^4,2^10,2^16,2^22,2^26} x {uniform, zipf 0.5/1.0/1.5}; COUNT, "
                               "SUM/MIN/MAX(int64), SUM(float32)",
                   "rows": n, "cases": len(cases),
                   "strategy": ("auto: learned dispatch" if dispatch and dispatch.get("learned") else
                                "auto: rule dispatch" if args.strategy == "auto" else args.strategy),
                   "hbm_frac_algorithmic": round(alg_bytes / elapsed / PEAK_HBM, 4),
                   "specialised_kernels": {"mode": ["off", "async", "sync"][jit_end["mode"]],
                                           "compiled": jit_end["compiled"], "failed": jit_end["failed"],
                                           "launches_timed": jit_end["launches"] - jit_timed0}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "dispatch": dispatch,
        "batched": batched,
        "sweep": sweep,
    }
    G.jit_wait()  # compiles queued by the batched comparison finish before exit
    print(json.dumps(out), flush=True)


def batched_workload(G, nq=4096, seed=42):
    """Batched uneven queries: nq independent group-bys (rows log-uniform in
    [1e3, 4e6], distinct count log-uniform in [1, rows/2], int32 keys uniform
    or Zipf s ~ U[0.5, 1.5], int64 values; SUM + COUNT) in one persistent
    launch replayed from a hipGraph, tuned by eg_batch_autotune.  Compared
    with the same queries run one by one through the top-level groupby."""
    import math

    import numpy as np

    A = G.aggregation_id
    T = G.type_id
    rng = np.random.default_rng(seed)
    specs = []
    for q in range(nq):
        rows = int(math.exp(rng.uniform(math.log(1e3), math.log(4e6))))
        d = max(1, int(math.exp(rng.uniform(0.0, math.log(max(rows / 2, 1.0))))))
        s = 0.0 if rng.random() < 0.5 else float(rng.uniform(0.5, 1.5))
        specs.append((rows, d, s))
    total = sum(r for r, _, _ in specs)
    t0 = time.time()
    vals = G.gen_values(total, G.val_dist.UNIFORM_I64, T.INT64, seed=seed, stream_id=2)
    keys, outs, queries = [], [], []
    off = 0
    for q, (rows, d, s) in enumerate(specs):
        k = G.gen_keys_uniform(rows, d, seed=seed + q, t=T.INT32) if s == 0.0 else \
            G.gen_keys_zipf(rows, d, s, seed=seed + q, t=T.INT32)
        keys.append(k)
        cap = min(rows, d)
        ok, o1, o2 = G.make_column(T.INT32, cap), G.make_column(T.INT64, cap), G.make_column(T.INT64, cap)
        outs += [ok, o1, o2]
        queries.append(G.batch_query(k, G.column_view(vals.data + off * 8, rows, T.INT64), d, ok, [o1, o2]))
        off += rows
    log(f"batched workload: {nq} queries, {total} rows generated in {time.time() - t0:.1f}s")
    reqs = [(0, A.SUM), (-1, A.COUNT_ALL)]
    cands = [G.batch_tuning(tile_rows_log2=t, lds_slots_log2=c, lds_max_bound_log2=b, blocks_per_cu=w)
             for t in (15, 17) for c in (10, 11) for b in (10, 12) for w in (1, 2)]
    best, times = G.batch_autotune(queries, T.INT32, T.INT64, reqs, cands, reps=1)
    plan = G.batch_plan(queries, T.INT32, T.INT64, reqs, tuning=best)
    plan.graph_launch()
    ng, st = plan.results()
    reps = 5
    a = time.perf_counter()
    for _ in range(reps):
        plan.graph_launch()
    plan.results()
    batch_s = (time.perf_counter() - a) / reps
    # the same queries one by one (top-level groupby, rule dispatch, bound hint)
    G.mlp_load(None)
    a = time.perf_counter()
    for q, (rows, d, s) in enumerate(specs):
        t = G.groupby([keys[q]], [queries[q].vals], reqs, G.groupby_hint(groups_upper_bound=d))
        t.free()
    seq_s = time.perf_counter() - a
    bad = int(np.count_nonzero(st))
    for c in keys + outs:
        c.free()
    vals.free()
    plan.free()
    bt = best
    return {"queries": nq, "rows": total, "grows_per_s": round(total / batch_s / 1e9, 3),
            "ms": round(batch_s * 1e3, 3), "sequential_ms": round(seq_s * 1e3, 1),
            "speedup_vs_sequential": round(seq_s / batch_s, 2), "queries_with_status": bad,
            "algorithmic_hbm_frac": round((12 * total + 20 * int(ng.sum())) / batch_s / PEAK_HBM, 4),
            "tuning": {"tile_rows_log2": bt.tile_rows_log2, "lds_slots_log2": bt.lds_slots_log2,
                       "lds_max_bound_log2": bt.lds_max_bound_log2, "blocks_per_cu": bt.blocks_per_cu},
            "autotune_candidates": len(cands), "autotune_ms": [round(float(x), 3) for x in times]}


def cpu_baseline(cases, n, sample):
    """The oracle's OpenMP group-by (hash partition + per-partition tables) on
    the first `sample` rows of every case, same aggregates."""
    import numpy as np

    import oracle as O

    try:
        threads = min(16, len(os.sched_getaffinity(0)))
    except AttributeError:
        threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sample = min(sample, n)
    v64 = O.gen_values(sample, O.VAL_UNIFORM_I64, np.int64, seed=SEED, stream_id=2)
    vf = O.gen_values(sample, O.VAL_UNIFORM_F32, np.float32, seed=SEED, stream_id=3)
    reqs = [(-1, O.COUNT_ALL), (0, O.SUM), (0, O.MIN), (0, O.MAX), (1, O.SUM)]
    total = 0.0
    for d, name, s in cases:
        if name == "uniform":
            k = O.gen_keys_uniform(n, d, seed=SEED, begin=0, count=sample)
        else:
            k = O.gen_keys_zipf(sample, O.zipf_cdf(d, s), seed=SEED)
        a = time.perf_counter()
        O.groupby(k, [v64, vf], reqs, parallel=True, nthreads=threads)
        total += time.perf_counter() - a
    rows = sample * len(cases)
    log(f"cpu baseline: {rows} rows in {total:.2f}s on {threads} threads")
    return {"value": round(rows / total / 1e9, 4), "unit": "Grows/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} rows of each of the {len(cases)} cases (same generator, same 5 "
                      f"aggregates), oracle OpenMP hash-partition group-by"}


if __name__ == "__main__":
    main()
