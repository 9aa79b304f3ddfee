"""Timing of the dictionary lookup (gamax::compress::dictionary_encode with a
given dictionary) on the reference's own test sizes
(dictionary_encode_dict_ready_test.cu:93-96): 268435356 keys (0..n-1) mod D
shuffled against the dictionary 0..D-1, plus int64 keys, a 30000-key
dictionary (the largest direct table) and a 2^20-key one.  Prints one line per case: wall ms of the whole call (range
reduction, table build, lookup) and the lookup's algorithmic rate
(key + id bytes per row).

    python3 -m eiger_amd.bench_dict [--reps 4]
"""
import argparse
import time

from eiger_amd import gamax as G


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=4)
    a = ap.parse_args()
    n = 268435356
    for d, t in [(1024, G.type_id.INT32), (4096, G.type_id.INT32), (4096, G.type_id.INT64),
                 (30000, G.type_id.INT32), (1 << 20, G.type_id.INT32)]:
        keys = G.gen_keys_uniform(n, d, seed=2, t=t)
        dict_col = G.gen_keys_uniform(d, d, seed=3, t=t)  # a permutation of 0..d-1
        best = float("inf")
        for _ in range(a.reps):
            t0 = time.perf_counter()
            e = G.dictionary_encode_with_dict(keys, dict_col)
            best = min(best, time.perf_counter() - t0)
            e.free()
        kb = 4 if t == G.type_id.INT32 else 8
        print(f"rows={n} D={d} key_bytes={kb} wall_ms={1e3 * best:.3f} "
              f"wall_GBps={(kb + 4) * n / best / 1e9:.0f}", flush=True)
        keys.free()
        dict_col.free()


if __name__ == "__main__":
    main()
