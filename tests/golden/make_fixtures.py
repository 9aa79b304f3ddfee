"""Writes tests/golden/reference_cases.json: the reference's own group-by test
vectors (inputs and expected outputs) as data.

Sources (src/cuda/ of the reference repository):
  tests/types/floating_point_test.cu:19-72      MAX over float/double, 4 keys
  tests/phys/phys_test.cu:533-548                GroupBySum, 3 groups
  tests/groupby/groupby_multikeys_test.cu:54-1473 expected maps of the
      multi-key suite.  This build is single-key, so each composite key is fused
      into one int64 key (k1 << 32 | k2 as uint32; three small keys as
      k1 << 40 | k2 << 20 | k3) and the expected map is re-keyed the same way;
      the expected aggregate values are the reference's.
  tests/aggregate/aggregate_test.cu:11-111      scalar SUM/MAX closed forms

Run: python tests/golden/make_fixtures.py
"""
import json
import os

SUM, MAX, MIN, AVG, ANY, COUNT_ALL = 0, 1, 2, 3, 4, 5
I32, I64, F32, F64 = "int32", "int64", "float32", "float64"


def fuse2(a, b):
    v = (a << 32) | (b & 0xFFFFFFFF)
    return v - (1 << 64) if v >= (1 << 63) else v


def fuse3(a, b, c):
    return (a << 40) | (b << 20) | c


def mk(name, src, k1, k2, vals, vtypes, reqs, expected, agg_count=False, hint=10, out_types=None, tol=None):
    keys = [fuse2(a, b) for a, b in zip(k1, k2)]
    exp_keys = sorted(fuse2(a, b) for (a, b) in expected)
    rows = {fuse2(a, b): v for (a, b), v in expected.items()}
    cols = []
    ncol = len(next(iter(expected.values())))
    for j in range(ncol):
        cols.append([rows[k][j] for k in exp_keys])
    return dict(name=name, source=src, key_type=I64, keys=keys, vals=vals, val_types=vtypes,
                requests=reqs, agg_count=agg_count, groups_upper_bound=hint,
                expected_keys=exp_keys, expected_cols=cols, expected_types=out_types, tol=tol)


cases = []

# floating_point_test.cu:19-37 (keys single int32 column; MAX)
for vt in (F32, F64):
    cases.append(dict(
        name=f"fp_max_{vt}", source="tests/types/floating_point_test.cu:19-72", key_type=I32,
        keys=[1, 2, 2, 3, 3, 3, 4, 4, 4, 4],
        vals=[[1.1, 2.1, 2.2, 3.1, 3.2, 3.3, 4.1, 4.2, 4.3, 4.4]], val_types=[vt],
        requests=[[0, MAX]], agg_count=False, groups_upper_bound=4,
        expected_keys=[1, 2, 3, 4], expected_cols=[[1.1, 2.2, 3.3, 4.4]], expected_types=[vt], tol=0.0))

# phys_test.cu:533-548
cases.append(dict(
    name="phys_groupby_sum", source="tests/phys/phys_test.cu:533-548", key_type=I32,
    keys=[1, 1, 2, 2, 3], vals=[[10, 20, 30, 40, 50]], val_types=[I32], requests=[[0, SUM]],
    agg_count=False, groups_upper_bound=(1 << 63) - 1,
    expected_keys=[1, 2, 3], expected_cols=[[30, 70, 50]], expected_types=[I64], tol=None))

S = "tests/groupby/groupby_multikeys_test.cu"
cases.append(mk("mk_two_keys_sum", S + ":54-106", [1, 1, 2, 2, 1], [10, 20, 10, 20, 10],
                [[100, 200, 300, 400, 500]], [I32], [[0, SUM]],
                {(1, 10): [600], (1, 20): [200], (2, 10): [300], (2, 20): [400]}, out_types=[I64]))
# three keys
k1, k2, k3 = [1, 1, 1, 1, 2, 2], [1, 1, 2, 2, 1, 1], [1, 2, 1, 2, 1, 2]
exp3 = {(1, 1, 1): 10, (1, 1, 2): 20, (1, 2, 1): 30, (1, 2, 2): 40, (2, 1, 1): 50, (2, 1, 2): 60}
ek = sorted(fuse3(*t) for t in exp3)
cases.append(dict(name="mk_three_keys_sum", source=S + ":111-169", key_type=I64,
                  keys=[fuse3(a, b, c) for a, b, c in zip(k1, k2, k3)], vals=[[10, 20, 30, 40, 50, 60]],
                  val_types=[I32], requests=[[0, SUM]], agg_count=False, groups_upper_bound=10,
                  expected_keys=ek, expected_cols=[[{fuse3(*t): v for t, v in exp3.items()}[k] for k in ek]],
                  expected_types=[I64], tol=None))
cases.append(mk("mk_two_keys_max", S + ":174-224", [1, 1, 1, 2, 2, 2], [1, 1, 2, 1, 2, 2],
                [[10, 50, 30, 40, 20, 60]], [I32], [[0, MAX]],
                {(1, 1): [50], (1, 2): [30], (2, 1): [40], (2, 2): [60]}, out_types=[I32]))
cases.append(mk("mk_mixed_int32_int64", S + ":229-280", [1, 1, 2, 2], [100, 200, 100, 200],
                [[10, 20, 30, 40]], [I32], [[0, SUM]],
                {(1, 100): [10], (1, 200): [20], (2, 100): [30], (2, 200): [40]}, out_types=[I64]))
cases.append(mk("mk_multiple_aggs", S + ":285-344", [1, 1, 2, 2, 1], [1, 2, 1, 2, 1],
                [[10, 20, 30, 40, 50], [5, 15, 25, 35, 45]], [I32, I32], [[0, SUM], [1, SUM]],
                {(1, 1): [60, 50], (1, 2): [20, 15], (2, 1): [30, 25], (2, 2): [40, 35]},
                out_types=[I64, I64]))
cases.append(mk("mk_single_group", S + ":349-388", [1] * 5, [100] * 5, [[10, 20, 30, 40, 50]], [I32],
                [[0, SUM]], {(1, 100): [150]}, out_types=[I64]))
cases.append(mk("mk_all_unique", S + ":393-440", [1, 2, 3, 4, 5], [10, 20, 30, 40, 50],
                [[100, 200, 300, 400, 500]], [I32], [[0, SUM]],
                {(1, 10): [100], (2, 20): [200], (3, 30): [300], (4, 40): [400], (5, 50): [500]}, out_types=[I64]))
cases.append(mk("mk_negative_keys", S + ":445-489", [-1, -1, 1, 1], [-10, 10, -10, 10],
                [[100, 200, 300, 400]], [I32], [[0, SUM]],
                {(-1, -10): [100], (-1, 10): [200], (1, -10): [300], (1, 10): [400]}, out_types=[I64]))
cases.append(mk("mk_first_key_matches", S + ":494-539", [1] * 6, [1, 2, 3, 1, 2, 3],
                [[10, 20, 30, 40, 50, 60]], [I32], [[0, SUM]],
                {(1, 1): [50], (1, 2): [70], (1, 3): [90]}, out_types=[I64]))
cases.append(mk("mk_large_number_of_groups", S + ":544-599", [i // 10 for i in range(100)],
                [i % 10 for i in range(100)], [[1] * 100], [I32], [[0, SUM]],
                {(i, j): [1] for i in range(10) for j in range(10)}, hint=200, out_types=[I64]))
cases.append(mk("mk_sum_and_max_same_col", S + ":604-658", [1, 1, 1, 2, 2], [1, 1, 2, 1, 1],
                [[10, 30, 20, 50, 40]], [I32], [[0, SUM], [0, MAX]],
                {(1, 1): [40, 30], (1, 2): [20, 20], (2, 1): [90, 50]}, out_types=[I64, I32]))
cases.append(mk("mk_float_values_sum", S + ":663-711", [1, 1, 2, 2], [1, 2, 1, 2], [[1.5, 2.5, 3.5, 4.5]],
                [F32], [[0, SUM]], {(1, 1): [1.5], (1, 2): [2.5], (2, 1): [3.5], (2, 2): [4.5]},
                out_types=[F32], tol=0.001))
cases.append(mk("mk_double_values_max", S + ":716-762", [1, 1, 1, 2, 2, 2], [1, 1, 2, 1, 2, 2],
                [[1.1, 3.3, 2.2, 5.5, 4.4, 6.6]], [F64], [[0, MAX]],
                {(1, 1): [3.3], (1, 2): [2.2], (2, 1): [5.5], (2, 2): [6.6]}, out_types=[F64], tol=0.001))
cases.append(mk("mk_basic_avg", S + ":771-827", [1, 1, 2, 2, 1], [10, 10, 20, 20, 10], [[10, 20, 30, 40, 30]],
                [I32], [[0, AVG]], {(1, 10): [20.0], (2, 20): [35.0]}, out_types=[F64], tol=0.001))
cases.append(mk("mk_avg_and_sum_same_col", S + ":832-893", [1, 1, 2, 2], [1, 1, 2, 2], [[10, 30, 20, 40]],
                [I32], [[0, SUM], [0, AVG]], {(1, 1): [40, 20.0], (2, 2): [60, 30.0]},
                out_types=[I64, F64], tol=0.001))
cases.append(mk("mk_avg_before_sum", S + ":898-954", [1, 1, 2, 2], [1, 1, 2, 2], [[10, 30, 20, 40]],
                [I32], [[0, AVG], [0, SUM]], {(1, 1): [20.0, 40], (2, 2): [30.0, 60]},
                out_types=[F64, I64], tol=0.001))
cases.append(mk("mk_multiple_avg_diff_cols", S + ":959-1024", [1, 1, 2, 2], [1, 1, 2, 2],
                [[10, 20, 30, 40], [100, 200, 300, 400]], [I32, I32], [[0, AVG], [1, AVG]],
                {(1, 1): [15.0, 150.0], (2, 2): [35.0, 350.0]}, out_types=[F64, F64], tol=0.001))
cases.append(mk("mk_multiple_avg_same_col", S + ":1029-1082", [1, 1, 2, 2], [1, 1, 2, 2], [[10, 30, 20, 40]],
                [I32], [[0, AVG], [0, AVG]], {(1, 1): [20.0, 20.0], (2, 2): [30.0, 30.0]},
                out_types=[F64, F64], tol=0.001))
cases.append(mk("mk_avg_and_max_same_col", S + ":1087-1144", [1, 1, 1, 2, 2], [1, 1, 1, 2, 2],
                [[10, 20, 30, 40, 50]], [I32], [[0, AVG], [0, MAX]], {(1, 1): [20.0, 30], (2, 2): [45.0, 50]},
                out_types=[F64, I32], tol=0.001))
cases.append(mk("mk_avg_sum_max_same_col", S + ":1149-1207", [1, 1, 2, 2], [1, 1, 2, 2], [[10, 30, 20, 60]],
                [I32], [[0, SUM], [0, AVG], [0, MAX]], {(1, 1): [40, 20.0, 30], (2, 2): [80, 40.0, 60]},
                out_types=[I64, F64, I32], tol=0.001))
cases.append(mk("mk_avg_float_values", S + ":1212-1259", [1, 1, 2, 2], [1, 1, 2, 2], [[1.5, 2.5, 3.5, 4.5]],
                [F32], [[0, AVG]], {(1, 1): [2.0], (2, 2): [4.0]}, out_types=[F64], tol=0.001))
cases.append(mk("mk_single_group_avg", S + ":1264-1299", [1] * 5, [10] * 5, [[10, 20, 30, 40, 50]], [I32],
                [[0, AVG]], {(1, 10): [30.0]}, out_types=[F64], tol=0.001))
cases.append(mk("mk_all_unique_avg", S + ":1304-1348", [1, 2, 3, 4, 5], [10, 20, 30, 40, 50],
                [[100, 200, 300, 400, 500]], [I32], [[0, AVG]],
                {(1, 10): [100.0], (2, 20): [200.0], (3, 30): [300.0], (4, 40): [400.0], (5, 50): [500.0]},
                out_types=[F64], tol=0.001))
cases.append(mk("mk_avg_with_count_hint", S + ":1353-1402", [1, 1, 2, 2, 2], [1, 1, 2, 2, 2],
                [[10, 20, 30, 40, 50]], [I32], [[0, AVG]], {(1, 1): [15.0, 2], (2, 2): [40.0, 3]},
                agg_count=True, out_types=[F64, I64], tol=0.001))
cases.append(mk("mk_complex_mixed", S + ":1407-1473", [1, 1, 2, 2], [1, 1, 2, 2],
                [[10, 20, 30, 40], [100, 200, 300, 400]], [I32, I32], [[0, SUM], [1, AVG], [0, AVG]],
                {(1, 1): [30, 150.0, 15.0], (2, 2): [70, 350.0, 35.0]}, out_types=[I64, F64, F64], tol=0.001))

# aggregate_test.cu:11-111: scalar SUM = n(n-1)/2 over 0..n-1, MAX = n-1
scalar = []
for t, n, kind in ((I32, 1 << 14, SUM), (I32, 1 << 25, MAX), (I64, 1 << 14, SUM), (I64, 1 << 25, MAX)):
    scalar.append(dict(name=f"aggregate_{'sum' if kind == SUM else 'max'}_{t}",
                       source="tests/aggregate/aggregate_test.cu:44-58", type=t, rows=n, kind=kind,
                       expected=(n * (n - 1) // 2) if kind == SUM else n - 1))

# stats_compute_test.cu:197-244: HLL estimate bounds
src = "tests/stats/stats_compute_test.cu"
ndv = [dict(name="ndv_all_distinct", source=src + ":197-207", data=list(range(1, 11)), lo=8, hi=12),
       dict(name="ndv_all_same", source=src + ":209-218", data=[42] * 10, lo=1, hi=1),
       dict(name="ndv_some_duplicates", source=src + ":220-231", data=[1, 2, 3, 4, 5] * 2, lo=4, hi=6),
       dict(name="ndv_larger", source=src + ":233-244", data=[i % 100 for i in range(1000)], lo=80, hi=120)]

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_cases.json")
with open(out, "w") as f:
    json.dump(dict(groupby=cases, scalar=scalar, ndv=ndv), f, indent=1)
print(f"wrote {len(cases)} group-by cases, {len(scalar)} scalar cases -> {out}")
