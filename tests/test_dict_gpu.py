"""Dictionary encoding (gamax::compress::dictionary_encode) on the GPU against
the oracle: the reference's own cases (dictionary_encode_dict_ready_test.cu:
85-96, dictionary 0..D-1 over keys (0..n-1) mod D shuffled, so the encoding
is the key itself), random dictionaries with repeats and absent keys, every
integer key width, the LDS and HBM table paths, and the self-built (ordered
and unordered) forms.  All bit-exact."""
import numpy as np
import pytest

import oracle as O
from eiger_amd import gamax as G

pytestmark = pytest.mark.gpu

T = G.type_id


def encode(keys, dict_keys):
    k, d = G.make_column_from_host(keys), G.make_column_from_host(dict_keys)
    e = G.dictionary_encode_with_dict(k, d)
    out = e.to_numpy()
    for c in (k, d, e):
        c.free()
    return out


@pytest.mark.parametrize("n,d", [(4096, 1), (4096, 16), (4096, 4096)])
def test_reference_small_cases(n, d):
    keys = G.gen_keys_uniform(n, d, seed=1)
    dict_col = G.make_column_from_host(np.arange(d, dtype=np.int32))
    e = G.dictionary_encode_with_dict(keys, dict_col)
    assert e.type == T.INT32 and e.size == n
    host = keys.to_numpy()
    np.testing.assert_array_equal(e.to_numpy(), host)
    np.testing.assert_array_equal(e.to_numpy(), O.dict_encode(host, np.arange(d, dtype=np.int32)))
    for c in (keys, dict_col, e):
        c.free()


@pytest.mark.parametrize("n,d", [(268435356, 1024), (268435356, 2048), (268435356, 4096), (536870712, 1024)])
def test_reference_full_size_cases(n, d):
    """The reference's 2^28 and 2^29-row cases: the encoding equals the key,
    every id in [0, D) occurs (each n/D or n/D + 1 times)."""
    keys = G.gen_keys_uniform(n, d, seed=2)
    dict_col = G.make_column_from_host(np.arange(d, dtype=np.int32))
    e = G.dictionary_encode_with_dict(keys, dict_col)
    got, host = e.to_numpy(), keys.to_numpy()
    for c in (keys, dict_col, e):
        c.free()
    assert np.array_equal(got, host)
    counts = np.bincount(got, minlength=d)
    assert len(counts) == d and counts.min() >= n // d and counts.max() <= n // d + 1


@pytest.mark.parametrize("dt", [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64])
# (dictionary size, key span / size): a span below 32768 takes the direct-indexed
# table, a wider one the hashed table (in LDS up to 8192 slots, else in HBM);
# the 1- and 2-byte types are always direct
@pytest.mark.parametrize("m,spread", [(0, 3), (5, 3), (3000, 3), (10000, 3), (3000, 4096), (20000, 3)])
def test_random_dictionaries_against_oracle(dt, m, spread):
    rng = np.random.default_rng(m * 31 + spread + np.dtype(dt).itemsize)
    info = np.iinfo(dt)
    span = min(int(info.max) - int(info.min), spread * max(m, 4))
    base = rng.integers(info.min, int(info.min) + span, 60000, dtype=dt, endpoint=True)
    base[:2] = [info.min, info.max]  # the extremes are ordinary keys
    dict_keys = rng.choice(base, m) if m else base[:0]  # repeats allowed: first position wins
    out = encode(base, dict_keys)
    assert out.dtype == np.int32
    np.testing.assert_array_equal(out, O.dict_encode(base, dict_keys))


@pytest.mark.parametrize("dt", [np.int32, np.int64, np.uint8])
@pytest.mark.parametrize("with_order", [True, False])
def test_self_built_dictionary(dt, with_order):
    rng = np.random.default_rng(5)
    info = np.iinfo(dt)
    pool = rng.integers(info.min, info.max, 200 if dt != np.uint8 else 50, dtype=dt, endpoint=True)
    keys = rng.choice(pool, 1 << 18)
    k = G.make_column_from_host(keys)
    e, size, dict_col = G.dictionary_encode(k, T.INT64, with_order=with_order, return_dictionary=True)
    ids, dk = e.to_numpy(), dict_col.to_numpy()
    for c in (k, e, dict_col):
        c.free()
    exp_ids, exp_size = O.dict_encode_ordered(keys)
    assert size == exp_size == len(dk) and ids.dtype == np.int64
    # the returned dictionary decodes every id, and every id is used
    np.testing.assert_array_equal(dk[ids], keys)
    assert len(np.unique(ids)) == size and ids.min() == 0 and ids.max() == size - 1
    if with_order:
        np.testing.assert_array_equal(ids, exp_ids)
        np.testing.assert_array_equal(dk, np.unique(keys))  # ascending distinct keys


def test_self_built_types_and_errors():
    keys = np.arange(1000, dtype=np.int32) % 300
    k = G.make_column_from_host(keys)
    e, size = G.dictionary_encode(k, T.INT16, with_order=True)
    assert size == 300 and e.type == T.INT16
    np.testing.assert_array_equal(e.to_numpy(), keys)
    e.free()
    with pytest.raises(G.GamaxError):
        G.dictionary_encode(k, T.INT8)  # 300 ids do not fit
    with pytest.raises(G.GamaxError):
        G.dictionary_encode(k, T.FLOAT32)
    with pytest.raises(G.GamaxError):
        G.dictionary_encode_with_dict(k, k, encode_type=T.INT64)  # the reference encodes to INT32 only
    f = G.make_column_from_host(np.zeros(4, np.float32))
    with pytest.raises(G.GamaxError):
        G.dictionary_encode_with_dict(f, f)
    d64 = G.make_column_from_host(np.arange(4, dtype=np.int64))
    with pytest.raises(G.GamaxError):
        G.dictionary_encode_with_dict(k, d64)
    empty = G.make_column_from_host(np.zeros(0, np.int32))
    e, size = G.dictionary_encode(empty)
    assert size == 0 and e.size == 0
    e2 = G.dictionary_encode_with_dict(empty, k)
    assert e2.size == 0
    for c in (k, f, d64, empty, e, e2):
        c.free()
