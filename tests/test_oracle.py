"""CPU checks of the oracle itself: the reference's golden vectors, published
known-answer tests, and independent textbook restatements (numpy), chosen so
that a dropped term, wrong sign/index or transposed operand fails one."""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "reference_cases.json")))


def _np(vals, t):
    return np.array(vals, dtype=np.dtype(t))


# ------------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("case", FIX["groupby"], ids=lambda c: c["name"])
def test_oracle_reference_groupby_cases(case):
    keys = _np(case["keys"], case["key_type"])
    vals = [_np(v, t) for v, t in zip(case["vals"], case["val_types"])]
    out = O.groupby(keys, vals, [tuple(r) for r in case["requests"]], agg_count=case["agg_count"])
    assert out[0].tolist() == case["expected_keys"]
    assert len(out) == 1 + len(case["expected_cols"])
    for got, exp, et in zip(out[1:], case["expected_cols"], case["expected_types"]):
        assert got.dtype == np.dtype(et)
        if case["tol"] is None:
            assert got.tolist() == exp
        else:
            np.testing.assert_allclose(got.astype(np.float64), np.array(exp, dtype=np.float64),
                                       rtol=0, atol=max(case["tol"], 1e-6))


@pytest.mark.parametrize("case", FIX["scalar"], ids=lambda c: c["name"])
def test_oracle_reference_scalar_cases(case):
    n = case["rows"]
    col = np.arange(n, dtype=np.dtype(case["type"]))
    out = O.aggregate([col, col], [(0, case["kind"]), (1, case["kind"])])
    assert len(out) == 2
    assert int(out[0][0]) == case["expected"] and int(out[1][0]) == case["expected"]


# ------------------------------------------------------------------ known-answer tests
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors published with Random123 (kat_vectors)."""
    import ctypes as C
    L = O.lib()
    L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32 * 4), C.POINTER(C.c_uint32 * 2), C.POINTER(C.c_uint32 * 4)]
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, exp in kat:
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        L.orc_philox4x32_10(C.byref(c), C.byref(k), C.byref(o))
        assert tuple(o) == exp


def _murmur3_x86_32(data: bytes, seed: int) -> int:
    """MurmurHash3_x86_32 body with the oracle's fmix32 as its finalizer."""
    c1, c2 = 0xcc9e2d51, 0x1b873593
    h = seed & 0xFFFFFFFF
    rotl = lambda x, r: ((x << r) | (x >> (32 - r))) & 0xFFFFFFFF
    nblocks = len(data) // 4
    for i in range(nblocks):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (k * c1) & 0xFFFFFFFF
        k = rotl(k, 15)
        k = (k * c2) & 0xFFFFFFFF
        h ^= k
        h = rotl(h, 13)
        h = (h * 5 + 0xe6546b64) & 0xFFFFFFFF
    tail = data[4 * nblocks:]
    k = 0
    if len(tail) >= 3:
        k ^= tail[2] << 16
    if len(tail) >= 2:
        k ^= tail[1] << 8
    if len(tail) >= 1:
        k ^= tail[0]
        k = (k * c1) & 0xFFFFFFFF
        k = rotl(k, 15)
        k = (k * c2) & 0xFFFFFFFF
        h ^= k
    h ^= len(data)
    return O.fmix32(h)


def test_fmix32_against_murmurhash3_vectors():
    """Published MurmurHash3_x86_32 test vectors pin the fmix32 finalizer the
    reference uses for partition ids (groupby_partition.cu:64-90)."""
    vectors = [
        (b"", 0, 0x00000000), (b"", 1, 0x514E28B7), (b"", 0xffffffff, 0x81F16F39),
        (b"\x00\x00\x00\x00", 0, 0x2362F9DE), (b"aaaa", 0x9747b28c, 0x5A97808A),
        (b"Hello, world!", 0x9747b28c, 0x24884CBA),
        (b"The quick brown fox jumps over the lazy dog", 0x9747b28c, 0x2FA826CD),
    ]
    for data, seed, exp in vectors:
        assert _murmur3_x86_32(data, seed) == exp, (data, seed)


def _murmur3_x64_128(data: bytes, seed: int = 0) -> bytes:
    """MurmurHash3_x64_128 (public domain, A. Appleby) with the oracle's fmix64."""
    M = (1 << 64) - 1
    c1, c2 = 0x87c37b91114253d5, 0x4cf5ad432745937f

    def rotl(x, r):
        return ((x << r) | (x >> (64 - r))) & M

    h1 = h2 = seed
    n = len(data)
    for i in range(n // 16):
        k1 = int.from_bytes(data[16 * i:16 * i + 8], "little")
        k2 = int.from_bytes(data[16 * i + 8:16 * i + 16], "little")
        h1 ^= (rotl((k1 * c1) & M, 31) * c2) & M
        h1 = (rotl(h1, 27) + h2) & M
        h1 = (h1 * 5 + 0x52dce729) & M
        h2 ^= (rotl((k2 * c2) & M, 33) * c1) & M
        h2 = (rotl(h2, 31) + h1) & M
        h2 = (h2 * 5 + 0x38495ab5) & M
    tail = data[16 * (n // 16):]
    if len(tail) > 8:
        h2 ^= (rotl((int.from_bytes(tail[8:], "little") * c2) & M, 33) * c1) & M
    if len(tail) > 0:
        h1 ^= (rotl((int.from_bytes(tail[:8], "little") * c1) & M, 31) * c2) & M
    h1 ^= n
    h2 ^= n
    h1 = (h1 + h2) & M
    h2 = (h2 + h1) & M
    h1, h2 = O.fmix64(h1), O.fmix64(h2)
    h1 = (h1 + h2) & M
    h2 = (h2 + h1) & M
    return h1.to_bytes(8, "little") + h2.to_bytes(8, "little")


def test_fmix64_against_murmurhash3_x64_128_vectors():
    """Published MurmurHash3_x64_128 digests (seed 0) pin fmix64, the 8-byte
    key finalizer (groupby_partition.cu:64-90)."""
    assert _murmur3_x64_128(b"The quick brown fox jumps over the lazy dog").hex() == \
        "6c1b07bc7bbc4be347939ac4a93c437a"
    assert _murmur3_x64_128(b"").hex() == "00" * 16


def test_fmix64_is_the_invertible_murmur_finalizer():
    c1, c2 = 0xff51afd7ed558ccd, 0xc4ceb9fe1a85ec53
    i1, i2 = pow(c1, -1, 1 << 64), pow(c2, -1, 1 << 64)
    M = (1 << 64) - 1

    def unshift(y, s):
        x = y
        for _ in range(64 // s + 1):
            x = y ^ (x >> s)
        return x & M

    rng = np.random.default_rng(7)
    for x in [0, 1, M, 1 << 63] + [int(v) for v in rng.integers(0, 1 << 63, 50, dtype=np.int64)]:
        h = O.fmix64(x)
        y = unshift(h, 33)
        y = (y * i2) & M
        y = unshift(y, 33)
        y = (y * i1) & M
        y = unshift(y, 33)
        assert y == x
    assert O.fmix64(0) == 0


# ------------------------------------------------------------------ independent restatement
def _numpy_groupby(keys, vals, reqs, agg_count=False):
    """Textbook group-by with numpy: np.unique + reduceat in sorted order."""
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    uniq, starts, counts = np.unique(sk, return_index=True, return_counts=True)
    out = [uniq]
    for c, k in reqs:
        if k == O.COUNT_ALL:
            out.append(counts.astype(np.int64))
            continue
        v = vals[c][order]
        if k in (O.SUM, O.AVG):
            if v.dtype.kind == "f":
                s = np.add.reduceat(v.astype(np.float64), starts)
            elif v.dtype.kind == "u":
                s = np.add.reduceat(v.astype(np.uint64), starts)
            else:
                s = np.add.reduceat(v.astype(np.int64), starts)
            if k == O.AVG:
                out.append(s.astype(np.float64) / counts)
            else:
                out.append(s.astype(np.float32) if v.dtype == np.float32 else s)
        elif k == O.MAX:
            out.append(np.maximum.reduceat(v, starts))
        elif k == O.MIN:
            out.append(np.minimum.reduceat(v, starts))
    if agg_count:
        out.append(counts.astype(np.int64))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("kt", [np.int32, np.int64, np.uint32, np.uint64])
def test_oracle_matches_numpy_restatement(seed, kt):
    rng = np.random.default_rng(seed)
    n = 5000
    info = np.iinfo(kt)
    pool = rng.integers(info.min // 2 if info.min < 0 else 0, info.max // 2, 300, dtype=np.int64).astype(kt)
    pool[:4] = [info.min, info.max, 0, info.min + 1]
    keys = pool[rng.integers(0, len(pool), n)]
    v64 = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    v32 = rng.integers(-1000, 1000, n).astype(np.int32)
    vu = rng.integers(0, 1 << 31, n).astype(np.uint32)
    vf = rng.random(n).astype(np.float32)
    vd = rng.standard_normal(n)
    vals = [v64, v32, vu, vf, vd]
    reqs = [(0, O.SUM), (0, O.MIN), (0, O.MAX), (1, O.SUM), (1, O.MAX), (2, O.SUM), (2, O.MIN),
            (3, O.SUM), (3, O.MAX), (4, O.MIN), (4, O.AVG), (1, O.AVG), (-1, O.COUNT_ALL)]
    got = O.groupby(keys, vals, reqs, agg_count=True)
    exp = _numpy_groupby(keys, vals, reqs, agg_count=True)
    assert len(got) == len(exp)
    for j, (g, e) in enumerate(zip(got, exp)):
        if g.dtype.kind == "f":
            np.testing.assert_allclose(g, e, rtol=1e-6, err_msg=f"col {j}")
        else:
            np.testing.assert_array_equal(g, e, err_msg=f"col {j}")


@pytest.mark.parametrize("kt", [np.int32, np.int64])
def test_parallel_baseline_matches_bruteforce(kt):
    rng = np.random.default_rng(11)
    n = 200_000
    keys = rng.integers(-50_000, 50_000, n).astype(kt)
    v = rng.integers(-(1 << 31), 1 << 31, n).astype(np.int64)
    f = rng.random(n).astype(np.float32)
    reqs = [(-1, O.COUNT_ALL), (0, O.SUM), (0, O.MIN), (0, O.MAX), (1, O.SUM)]
    a = O.groupby(keys, [v, f], reqs)
    b = O.sort_by_key(O.groupby(keys, [v, f], reqs, parallel=True, nthreads=4))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_sum_int64_wraps_like_the_reference():
    """SUM over int64 wraps mod 2^64 (atomics/atomics.cuh:42-43)."""
    keys = np.zeros(3, dtype=np.int32)
    v = np.array([np.iinfo(np.int64).max, 1, 1], dtype=np.int64)
    out = O.groupby(keys, [v], [(0, O.SUM)])
    assert int(out[1][0]) == np.iinfo(np.int64).min + 1


def test_output_types_follow_agg_output_data_type():
    """aggregate.cu:20-53 with sum_output_type of agg_func.cuh:18-36."""
    T = O
    assert O.agg_output_type(O.SUM, T.INT32) == T.INT64
    assert O.agg_output_type(O.SUM, T.INT8) == T.INT64
    assert O.agg_output_type(O.SUM, T.UINT32) == T.UINT64
    assert O.agg_output_type(O.SUM, T.FLOAT32) == T.FLOAT32
    assert O.agg_output_type(O.MAX, T.INT32) == T.INT32
    assert O.agg_output_type(O.AVG, T.INT32) == T.FLOAT64
    assert O.agg_output_type(O.COUNT_ALL, T.INT32) == T.INT64


def test_empty_input_gives_zero_rows_with_schema():
    """phys/phys.cu:589-611: empty input -> 0-row table with [key, aggs]."""
    out = O.groupby(np.empty(0, np.int32), [np.empty(0, np.float64)], [(0, O.SUM)])
    assert len(out) == 2 and len(out[0]) == 0 and out[1].dtype == np.float64


# ------------------------------------------------------------------ generators
@pytest.mark.parametrize("n,d", [(1 << 12, 16), (1000, 7), (4096, 4096), (3000, 1)])
def test_uniform_keys_have_exact_counts(n, d):
    k = O.gen_keys_uniform(n, d, seed=42)
    c = np.bincount(k, minlength=d)
    assert c.min() >= n // d and c.max() <= -(-n // d) and c.sum() == n


@pytest.mark.parametrize("n", [1, 2, 5, 1000, 1024, 4097])
def test_row_permutation_is_a_bijection(n):
    p = [O.perm(i, n, 42) for i in range(n)]
    assert sorted(p) == list(range(n))


def test_uniform_keys_are_shuffled():
    k = O.gen_keys_uniform(1 << 16, 1 << 16, seed=42).astype(np.int64)
    # a shuffle: adjacent rows are not consecutive keys
    assert np.mean(np.abs(np.diff(k)) == 1) < 0.01


def test_prefix_generation_matches_full():
    full = O.gen_keys_uniform(10000, 97, seed=5)
    part = O.gen_keys_uniform(10000, 97, seed=5, begin=1234, count=500)
    np.testing.assert_array_equal(full[1234:1734], part)


def test_zipf_cdf_and_frequencies():
    d, s = 64, 1.0
    cdf = O.zipf_cdf(d, s)
    w = 1.0 / np.arange(1, d + 1) ** s
    np.testing.assert_allclose(cdf, np.cumsum(w) / w.sum(), rtol=1e-12)
    assert cdf[-1] == 1.0 and np.all(np.diff(cdf) > 0)
    n = 1 << 18
    z = O.gen_keys_zipf(n, cdf, seed=42)
    freq = np.bincount(z, minlength=d) / n
    p = w / w.sum()
    assert np.max(np.abs(freq - p) / np.sqrt(p * (1 - p) / n)) < 6.0


def test_value_distributions():
    v = O.gen_values(1 << 16, O.VAL_UNIFORM_I64, np.int64, seed=42, stream_id=2)
    assert v.min() >= -(1 << 31) and v.max() < (1 << 31) and abs(v.mean()) < 2e7
    f = O.gen_values(1 << 16, O.VAL_UNIFORM_F32, np.float32, seed=42, stream_id=3)
    assert f.min() >= 0 and f.max() < 1 and np.all(f * (1 << 24) == np.floor(f * (1 << 24)))
    i = O.gen_values(1 << 16, O.VAL_UNIFORM_I32_1000, np.int32, seed=42, stream_id=4)
    assert i.min() >= 0 and i.max() < 1000 and set(np.unique(i)) == set(range(1000))


# ------------------------------------------------------------------ sort / partition / dispatch
def _radix_bits(keys, begin, end):
    k = keys.astype(np.int64) if keys.dtype.kind == "i" else keys.astype(np.uint64)
    w = keys.dtype.itemsize * 8
    u = k.view(np.uint64) if k.dtype == np.int64 else k
    if keys.dtype.kind == "i":
        u = (u & np.uint64((1 << w) - 1 if w < 64 else (1 << 64) - 1)) ^ np.uint64(1 << (w - 1))
    return (u >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1 if end - begin < 64 else (1 << 64) - 1)


@pytest.mark.parametrize("kt", [np.int32, np.int64, np.uint32])
@pytest.mark.parametrize("bits", [(0, -1), (0, 8), (4, 20)])
@pytest.mark.parametrize("desc", [False, True])
def test_sort_pairs_is_a_stable_sort_on_the_bit_range(kt, bits, desc):
    rng = np.random.default_rng(3)
    n = 3000
    info = np.iinfo(kt)
    keys = rng.integers(info.min, info.max, n, dtype=np.int64 if kt != np.uint64 else np.uint64).astype(kt)
    keys[:50] = keys[50:100]
    vals = np.arange(n, dtype=np.int64)
    b, e = bits
    e2 = keys.dtype.itemsize * 8 if e <= 0 else e
    sk, sv = O.sort_pairs(keys, vals, b, e, desc)
    d = _radix_bits(keys, b, e2)
    order = np.argsort(-d.astype(np.float64) if desc else d, kind="stable") if e2 - b <= 52 else None
    if order is None:
        import functools
        order = sorted(range(n), key=lambda i: (-int(d[i]) if desc else int(d[i]), i))
    np.testing.assert_array_equal(sv, vals[order])
    np.testing.assert_array_equal(sk, keys[order])


@pytest.mark.parametrize("kt", [np.float32, np.float64])
@pytest.mark.parametrize("desc", [False, True])
def test_sort_pairs_orders_floats_like_numpy_with_equal_zeros(kt, desc):
    """Full-width float sort equals numpy's stable sort by value: negative
    numbers before positive, infinities at the ends, and -0.0 / +0.0 equal
    (cub::DeviceRadixSort's documented treatment of the two zeros)."""
    rng = np.random.default_rng(4)
    keys = (rng.standard_normal(5000) * 100).astype(kt)
    keys[:12] = [0.0, -0.0, np.inf, -np.inf, 0.0, -0.0, 1e-30, -1e-30, 7.0, 7.0, -0.0, 0.0]
    rng.shuffle(keys)
    vals = np.arange(len(keys), dtype=np.int64)
    sk, sv = O.sort_pairs(keys, vals, 0, -1, desc)
    order = np.argsort(-keys if desc else keys, kind="stable")
    np.testing.assert_array_equal(sv, order)
    np.testing.assert_array_equal(sk.view(np.uint8), keys[order].view(np.uint8))


def test_partition_offsets_are_inclusive_prefix_of_counts():
    keys = np.array([3, 1, 0, 3, 3, 2, 1], dtype=np.int32)
    off = O.partition_offsets(keys, 0, 2)
    assert off.tolist() == [1, 3, 4, 7, 7]


def test_hash_partition_groups_rows_by_fmix_bits():
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 20, 5000).astype(np.int32)
    perm, off = O.hash_partition(keys, 8, 0)
    pid = O.partition_ids(keys, 0xFF)
    assert sorted(perm.tolist()) == list(range(len(keys)))
    for b in range(256):
        rows = perm[off[b]:off[b + 1]]
        assert np.all(pid[rows] == b) and np.all(np.diff(rows) > 0)


def test_rule_dispatcher_thresholds():
    """groupby.cu:104-118."""
    assert O.rule_select(4, 1 << 14) == O.HASH
    assert O.rule_select(4, (1 << 14) + 1) == O.PARTITION
    assert O.rule_select(8, 1 << 18) == O.HASH
    assert O.rule_select(8, (1 << 18) + 1) == O.PARTITION


# ------------------------------------------------------------------ learned dispatch oracle
def _np_mlp(P, x):
    """Textbook numpy forward pass of the 16-64-64-4 ReLU network."""
    W1 = P[:1024].reshape(64, 16); b1 = P[1024:1088]
    W2 = P[1088:5184].reshape(64, 64); b2 = P[5184:5248]
    W3 = P[5248:5504].reshape(4, 64); b3 = P[5504:5508]
    h1 = np.maximum(x.astype(np.float64) @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


def test_mlp_forward_matches_numpy():
    rng = np.random.default_rng(0)
    P = rng.standard_normal(O.MLP_PARAMS) * 0.3
    x = rng.random((33, 16)).astype(np.float32)
    np.testing.assert_allclose(O.mlp_forward(P, x), _np_mlp(P, x), rtol=1e-12, atol=1e-12)


def test_mlp_gradient_matches_finite_differences():
    """Central differences of the oracle's own loss pin its analytic gradient
    (every layer's weights and biases are sampled)."""
    rng = np.random.default_rng(1)
    P = rng.standard_normal(O.MLP_PARAMS) * 0.2
    x = rng.random((7, 16)).astype(np.float32)
    t = rng.dirichlet([1, 1, 1], 7).astype(np.float32)
    c = rng.random(7).astype(np.float32)
    L, g = O.mlp_loss_grad(P, x, t, c, 0.7)
    # the loss at P equals a numpy restatement of the formula
    y = _np_mlp(P, x)
    z = y[:, :3]
    lse = np.log(np.exp(z - z.max(1, keepdims=True)).sum(1)) + z.max(1)
    ref = np.mean(-(t * (z - lse[:, None])).sum(1) + 0.7 * (y[:, 3] - c) ** 2)
    assert abs(L - ref) < 1e-12
    eps = 1e-6
    for i in list(rng.choice(O.MLP_PARAMS, 40, replace=False)) + [0, 1023, 1024, 1087, 1088, 5183, 5184, 5247,
                                                                  5248, 5503, 5504, 5507]:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (O.mlp_loss_grad(Pp, x, t, c, 0.7)[0] - O.mlp_loss_grad(Pm, x, t, c, 0.7)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


@pytest.mark.parametrize("kt", [np.int32, np.int64])
@pytest.mark.parametrize("rows,d", [(100, 7), (4096, 4096), (50_000, 300), (1_000_000, 1 << 20)])
def test_sample_features_match_numpy(kt, rows, d):
    rng = np.random.default_rng(rows)
    keys = rng.integers(-d, d, rows).astype(kt)
    f = O.sample_features(keys, group_bound=0, nops=5, value_bytes=12)
    ns = min(rows, 4096)
    stride = 1 if rows <= 4096 else rows // 4096
    s = keys[np.arange(ns) * stride].astype(np.int64)
    u, cnt = np.unique(s, return_counts=True)
    p = cnt / ns
    f1, f2 = (cnt == 1).sum(), (cnt == 2).sum()
    chao = len(u) + (f1 * f1 / (2 * f2) if f2 else f1 * (f1 - 1) / 2)
    gee = np.sqrt(rows / ns) * f1 + len(u) - f1
    exp = [np.log2(rows + 1) / 32, np.dtype(kt).itemsize / 8, np.log2(len(u) + 1) / 12,
           np.log2(min(chao, rows) + 1) / 32, np.log2(min(gee, rows) + 1) / 32, np.sort(cnt)[::-1][:8].sum() / ns,
           -(p * np.log2(p)).sum() / np.log2(ns), f1 / ns, f2 / ns, cnt.max() / ns,
           np.log2(float(s.max() - s.min() + 1) + 1) / 64, np.log2(rows + 1) / 32, 5 / 16, 12 / 32,
           (np.diff(s) >= 0).sum() / (ns - 1), ns / 4096]
    np.testing.assert_allclose(f, np.array(exp, np.float32), rtol=2e-6, atol=1e-7)


# ---------------------------------------------------------------- dictionary encoding
@pytest.mark.parametrize("d", [1, 16, 4096])
def test_dict_encode_reference_case_is_identity(d):
    """dictionary_encode_dict_ready_test.cu:59-92: dict = 0..D-1, keys =
    (0..n-1) mod D shuffled, so each key's position is the key itself."""
    rng = np.random.default_rng(d)
    keys = (np.arange(4096, dtype=np.int32) % d)
    rng.shuffle(keys)
    np.testing.assert_array_equal(O.dict_encode(keys, np.arange(d, dtype=np.int32)), keys.astype(np.int64))


@pytest.mark.parametrize("dt", [np.int8, np.uint8, np.int16, np.int32, np.uint32, np.int64, np.uint64])
def test_dict_encode_matches_bruteforce(dt):
    """First position in the dictionary, INT32_MIN when absent; extremes of the
    type (the reference's empty-key sentinel among them) are ordinary keys."""
    rng = np.random.default_rng(7)
    info = np.iinfo(dt)
    pool = np.array([info.min, info.max, 0, 1, 2, 3, 5, 8, 13], dtype=dt)
    dict_keys = rng.choice(pool, 7)  # repeats allowed
    keys = rng.choice(pool, 200)
    dl = dict_keys.tolist()
    exp = [dl.index(k) if k in dl else O.INT32_MIN for k in keys.tolist()]
    np.testing.assert_array_equal(O.dict_encode(keys, dict_keys), exp)
    assert (O.dict_encode(keys, dict_keys[:0]) == O.INT32_MIN).all()


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint16])
def test_dict_encode_ordered_matches_numpy_unique(dt):
    rng = np.random.default_rng(11)
    info = np.iinfo(dt)
    keys = rng.integers(info.min, info.max, 5000, dtype=dt, endpoint=True) % dt(97) if dt == np.uint16 else \
        rng.choice(rng.integers(info.min, info.max, 300, dtype=dt, endpoint=True), 5000)
    uniq, inv = np.unique(keys, return_inverse=True)
    ids, d = O.dict_encode_ordered(keys)
    assert d == len(uniq)
    np.testing.assert_array_equal(ids, inv.reshape(-1))
    # the ordered ids are the positions in the sorted distinct keys
    np.testing.assert_array_equal(O.dict_encode(keys, uniq), ids)
    assert O.dict_encode_ordered(keys[:0])[1] == 0


# ---- checker helpers for full-size results ------------------------------------------
@pytest.mark.parametrize("dt", [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64])
def test_argsort_keys_is_numpy_stable_argsort(dt):
    rng = np.random.default_rng(21)
    info = np.iinfo(dt)
    k = rng.integers(info.min, info.max, 50_001, dtype=dt, endpoint=True)
    k[:500] = k[500:1000]  # repeated keys: stability matters
    k[1000:1004] = [info.min, info.max, 0, info.min]
    for nt in (1, 3):
        np.testing.assert_array_equal(O.argsort_keys(k, nthreads=nt), np.argsort(k, kind="stable"))
    assert len(O.argsort_keys(k[:0])) == 0


def test_take_and_sort_by_key_fast():
    rng = np.random.default_rng(22)
    cols = [rng.permutation(100_000).astype(np.int64) - 50_000, rng.random(100_000).astype(np.float32),
            rng.integers(0, 9, 100_000).astype(np.int16)]
    idx = rng.integers(0, 100_000, 70_000)
    for c in cols:
        np.testing.assert_array_equal(O.take(c, idx), c[idx])
    fast, slow = O.sort_by_key_fast(cols), O.sort_by_key(cols)
    for a, b in zip(fast, slow):
        np.testing.assert_array_equal(a, b)


# ---- float and narrow integer keys --------------------------------------------
def _bruteforce_groups(keys, vals):
    """Group-by in pure Python on a dict: float keys compare with == (so -0.0
    and +0.0 meet: equal floats hash equally), every NaN is one group.
    COUNT, SUM, MIN, MAX of an int64 column, groups in ascending key order
    with the NaN group last."""
    g = {}
    for k, v in zip(keys.tolist(), vals.tolist()):
        kk = "nan" if (isinstance(k, float) and k != k) else k
        c, s, lo, hi = g.get(kk, (0, 0, v, v))
        g[kk] = (c + 1, s + v, min(lo, v), max(hi, v))
    order = sorted((k for k in g if k != "nan")) + (["nan"] if "nan" in g else [])
    return order, [g[k] for k in order]


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.int8, np.int16, np.uint8, np.uint16])
def test_oracle_float_and_narrow_keys_match_python_bruteforce(dt):
    rng = np.random.default_rng(5)
    if np.dtype(dt).kind == "f":
        pool = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1.5, -1.5, 1e-30, -2.0, 3.25,
                         np.finfo(dt).max, np.finfo(dt).min, np.finfo(dt).tiny], dtype=dt)
        pool = np.concatenate([pool, rng.standard_normal(40).astype(dt)])
    else:
        info = np.iinfo(dt)
        pool = np.concatenate([[info.min, info.max, 0], rng.integers(info.min, info.max, 30, endpoint=True)]).astype(dt)
    keys = pool[rng.integers(0, len(pool), 3000)]
    vals = rng.integers(-(1 << 40), 1 << 40, len(keys)).astype(np.int64)
    reqs = [(-1, O.COUNT_ALL), (0, O.SUM), (0, O.MIN), (0, O.MAX)]
    got = O.groupby(keys, [vals], reqs)
    order, states = _bruteforce_groups(keys, vals)
    assert len(got[0]) == len(order)
    for i, (k, (c, s, lo, hi)) in enumerate(zip(order, states)):
        if k == "nan":
            assert np.isnan(got[0][i])
        else:
            assert got[0][i] == k  # numeric equality: -0.0 == 0.0
        assert (got[1][i], got[2][i], got[3][i], got[4][i]) == (c, s, lo, hi)
