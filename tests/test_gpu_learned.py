"""Learned delta-stepping: the MLP's GPU kernels against the FP64 oracle, trace recording, and
the trained schedule against fixed delta (identical distances, and no slower)."""
import sys

import numpy as np
import pytest

import graphee_amd as gmp
import oracle
from graphee_amd import learned
from tests import graphs

pytestmark = pytest.mark.gpu


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((n, 16)).astype(np.float32)
    y = rng.integers(0, 8, n).astype(np.int32)
    m = rng.integers(0, 2**63, n, dtype=np.uint64) ^ rng.integers(0, 2**63, n, dtype=np.uint64)
    return X, y, m


def test_mlp_forward_and_gradient_match_fp64(engine):
    rng = np.random.default_rng(4)
    P = rng.normal(0, 0.3, gmp.MLP_PARAMS).astype(np.float32)
    engine.mlp_set(P)
    np.testing.assert_array_equal(engine.mlp_get(), P)
    X, y, m = _batch(300, 5)
    y[::7] = -1  # rows without a class train the shard outputs only
    loss, g = engine.mlp_loss_grad(X, y, m)
    want_loss, want_g = oracle.mlp_loss_grad(P.astype(np.float64), X, y, m)
    assert abs(loss - want_loss) < 1e-5 * max(1.0, abs(want_loss))
    err = np.abs(g - want_g).max() / np.abs(want_g).max()
    assert err < 1e-4, err
    # forward logits: the class part of the output equals the FP64 network's
    Y = engine.mlp_forward(X[:10])
    assert Y.shape == (10, 72)


def test_mlp_training_reduces_loss(engine):
    X, y, m = _batch(512, 6)
    y = (X[:, 0] * 8).astype(np.int32).clip(0, 7)  # learnable labels
    engine.mlp_init(1)
    l0, _ = engine.mlp_loss_grad(X, y, m)
    l1 = engine.mlp_train(X, y, m, epochs=300, lr=5e-3)
    assert l1 < 0.8 * l0, (l0, l1)


def test_trace_recording_keeps_distances(engine):
    engine.generate_rmat(12, 16, 7012, weighted=True)
    g = engine.download(weights=True)
    V = g["outdeg"].size
    src = engine.pick_source(3)
    rows, d, st = engine.sssp_trace(src)
    ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
    np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, ow, src))
    assert 0 < len(rows) <= st.epochs
    assert np.all((rows["best"] >= 0) & (rows["best"] < 8))
    assert np.all(rows["relaxed"] > 0) and np.all(rows["ms"] > 0)
    # the label is the fastest rollout
    np.testing.assert_array_equal(rows["best"], rows["ms"].argmin(axis=1))
    assert np.all(np.isfinite(rows["feat"]))


def _median_ms(engine, src, n=5, **kw):
    t = []
    for _ in range(n):
        _, st = engine.sssp(src, want_dist=False, **kw)
        t.append(st.ms_total)
    return float(np.median(t)), st


def test_learned_schedule_exact_and_not_slower(engine):
    """Trained on traces of R-MAT s14-s16, the learned widths give Dijkstra's distances on s18 and
    run no slower than fixed delta = 64 (median of five runs per source, summed over three
    sources, 5 % allowance for timing noise); the action counts cover every epoch."""
    say = lambda *a: print(*a, flush=True)  # noqa: E731
    say("record traces s14-s16")
    rows = learned.record(engine, [(14, 16, 7100), (15, 16, 7101), (16, 16, 7102)], sources_per_graph=2, seed=1)
    assert len(rows) >= 6
    loss = learned.train(engine, rows, epochs=300)
    say("trained on", len(rows), "rows, loss", loss, learned.accuracy(engine, rows))
    engine.generate_rmat(18, 16, 7200, weighted=True)
    g = engine.download(weights=True)
    V = g["outdeg"].size
    ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
    t_fix = t_lrn = 0.0
    for k in range(3):
        src = engine.pick_source(11 + k)
        say(f"s18 source {src}")
        want = oracle.dijkstra(ooff, oidx, ow, src)
        d_fix, _ = engine.sssp(src, method=gmp.SSSP_DELTA, delta=64)
        d_lrn, st = engine.sssp(src, method=gmp.SSSP_LEARNED)
        np.testing.assert_array_equal(d_fix, want)
        np.testing.assert_array_equal(d_lrn, want)
        assert st.policy == 2
        acts = engine.sssp_actions()
        assert acts.sum() == st.epochs, (acts, st.epochs)
        f, _ = _median_ms(engine, src, method=gmp.SSSP_DELTA, delta=64)
        m, _ = _median_ms(engine, src, method=gmp.SSSP_LEARNED)
        t_fix += f
        t_lrn += m
        say(f"  fixed 64: {f:.3f} ms, learned: {m:.3f} ms, actions {acts.tolist()}")
    say(f"learned vs fixed delta=64 on s18: {t_fix:.3f} / {t_lrn:.3f} ms = {t_fix / t_lrn:.2f}x")
    assert t_lrn <= 1.05 * t_fix, (t_fix, t_lrn)


def test_learned_finish_variant_is_exact(engine):
    """With the learned-finish option, class 7 finishes by Bellman-Ford: traces, training and the
    learned run keep Dijkstra's distances."""
    rows = learned.record(engine, [(12, 16, 7300), (13, 16, 7301)], sources_per_graph=2, seed=2, finish=True)
    learned.train(engine, rows, epochs=200)
    engine.set_option(gmp.OPT_LEARNED_FINISH, 1)
    try:
        engine.generate_rmat(14, 16, 7400, weighted=True)
        g = engine.download(weights=True)
        V = g["outdeg"].size
        ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
        for k in range(2):
            src = engine.pick_source(21 + k)
            d, st = engine.sssp(src, method=gmp.SSSP_LEARNED)
            np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, ow, src))
    finally:
        engine.set_option(gmp.OPT_LEARNED_FINISH, 0)
