"""Learned dispatch on the GPU: the HIP feature, forward and gradient kernels
against the fp64 oracle, training that lowers the loss, and the guarantee
that the dispatcher's choice never changes a group-by result."""
import numpy as np
import pytest

import oracle as O
from eiger_amd import gamax as G

pytestmark = pytest.mark.gpu
A = G.aggregation_id


@pytest.mark.parametrize("kt", [np.int32, np.int64])
def test_features_match_oracle(kt):
    rng = np.random.default_rng(3)
    cases = [(1, 1), (100, 7), (4096, 4096), (5000, 3), (200_000, 50_000), (1_000_003, 1 << 24)]
    cols, exp = [], []
    for i, (rows, d) in enumerate(cases):
        if i % 2:
            keys = rng.integers(-d, d, rows).astype(kt)
        else:
            keys = np.sort(rng.integers(0, d, rows)).astype(kt)
        cols.append(G.make_column_from_host(keys))
        exp.append(O.sample_features(keys, group_bound=d if i == 2 else 0, nops=5, value_bytes=12))
    got = G.mlp_features(cols, group_bounds=[0, 0, 4096, 0, 0, 0], nops=5, value_bytes=12)
    # counts are exact; entropy and logs differ only by fp32 rounding order
    np.testing.assert_allclose(got, np.array(exp), rtol=1e-5, atol=1e-6)


def test_forward_matches_fp64_oracle():
    rng = np.random.default_rng(0)
    P = G.mlp_init(7)
    P[1024:1088] = rng.standard_normal(64) * 0.1
    x = rng.random((1000, 16)).astype(np.float32)
    got = G.mlp_forward(P, x)
    exp = O.mlp_forward(P.astype(np.float64), x)
    np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-5)


def test_gradient_matches_fp64_oracle():
    """eg_mlp_loss_grad (fp32 HIP backward, 64 samples per workgroup, fixed-order
    reduction) against the analytic fp64 gradient of the oracle, itself pinned
    by finite differences in the CPU tests."""
    rng = np.random.default_rng(1)
    P = G.mlp_init(3)
    P[1024:1088] = rng.standard_normal(64) * 0.1
    n = 777
    x = rng.random((n, 16)).astype(np.float32)
    t = rng.dirichlet([1, 1, 1], n).astype(np.float32)
    c = rng.random(n).astype(np.float32)
    loss, g = G.mlp_loss_grad(P, x, t, c, 0.5)
    eloss, eg = O.mlp_loss_grad(P.astype(np.float64), x, t, c, 0.5)
    assert abs(loss - eloss) <= 1e-5 * abs(eloss)
    scale = np.abs(eg).max()
    np.testing.assert_allclose(g, eg, rtol=1e-3, atol=1e-5 * scale)
    assert np.linalg.norm(g - eg) <= 1e-4 * np.linalg.norm(eg)


def test_training_lowers_the_loss_and_learns_a_rule():
    """A synthetic labelling (best strategy from feature 2 thresholds, card =
    feature 3) is learned by eg_mlp_train."""
    rng = np.random.default_rng(5)
    n = 8192
    x = rng.random((n, 16)).astype(np.float32)
    lab = np.where(x[:, 2] < 0.4, 0, np.where(x[:, 5] < 0.5, 1, 2))
    t = np.eye(3, dtype=np.float32)[lab]
    c = x[:, 3].copy()
    P0 = G.mlp_init(11)
    l0, _ = G.mlp_loss_grad(P0, x, t, c, 1.0)
    P, losses = G.mlp_train(P0, x, t, c, epochs=40, batch=512, lr=3e-3, lam=1.0, seed=2)
    l1, _ = G.mlp_loss_grad(P, x, t, c, 1.0)
    assert l1 < 0.25 * l0 and losses[-1] < losses[0]
    pred = np.argmax(G.mlp_forward(P, x)[:, :3], axis=1)
    assert (pred == lab).mean() > 0.95


def test_dispatch_choice_never_changes_results():
    n = 1 << 20
    keys = O.gen_keys_uniform(n, 70_000, seed=9)
    vals = O.gen_values(n, O.VAL_UNIFORM_I64, np.int64, seed=9, stream_id=2)
    reqs = [(-1, A.COUNT_ALL), (0, A.SUM), (0, A.MAX)]
    exp = O.groupby(keys, [vals], reqs)
    kc, vc = G.make_column_from_host(keys), G.make_column_from_host(vals)
    for bias in range(3):  # a model that always picks HASH, SORT, then PARTITION
        P = np.zeros(G.MLP_PARAMS, np.float32)
        P[5504 + bias] = 5.0
        P[5507] = np.log2(70_000) / 32
        G.mlp_load(P)
        assert G.mlp_loaded()
        algo, card = G.mlp_select(kc, nops=3, value_bytes=8)
        assert algo == [G.groupby_algo.HASH, G.groupby_algo.SORT, G.groupby_algo.PARTITION][bias]
        assert abs(np.log2(card) - np.log2(70_000)) < 0.01
        h = G.groupby_hint()
        t = G.groupby([kc], [vc], reqs, h)
        assert h.groupby_algo == algo
        got = t.to_numpy()
        order = np.argsort(got[0], kind="stable")
        for g, e in zip(got, exp):
            np.testing.assert_array_equal(g[order], e)
        t.free()
    G.mlp_load(None)
    assert not G.mlp_loaded()


def test_batched_selection_equals_single_query_selection():
    rng = np.random.default_rng(21)
    keys = []
    for i, (rows, d) in enumerate([(1000, 10), (50_000, 40_000), (300_000, 1 << 20), (70_000, 3), (0, 1)]):
        k = rng.integers(0, d, rows).astype(np.int32 if i % 2 else np.int64)
        keys.append(G.make_column_from_host(k))
    ra, rn = G.auto_select_groupby_algo_batch(keys)
    for k, a, nd in zip(keys, ra, rn):
        if k.size == 0:
            assert nd == 0 and a == G.groupby_algo.HASH
            continue
        h = G.groupby_hint()
        G.auto_select_groupby_algo([k], h)
        assert h.groupby_algo == a and nd == G.compute_ndv_approx(k)
    P = G.mlp_init(4)
    G.mlp_load(P)
    try:
        live = [k for k in keys if k.size > 0]
        ma, mc = G.mlp_select_batch(live, None, 2, 8)
        for k, a, c in zip(live, ma, mc):
            a1, c1 = G.mlp_select(k, 0, 2, 8)
            assert a1 == a and c1 == c
    finally:
        G.mlp_load(None)
