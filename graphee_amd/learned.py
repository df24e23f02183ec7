"""Learned delta-stepping schedule: record traces, train the MLP on the GPU, run it.

The engine records a trace row at checkpoint bucket epochs of SSSP (`Engine.sssp_trace`): the
16 features of the far pile, and for every class the time and relaxations of finishing the
solve from that state (one epoch of the class, then fixed width 64), the fastest class and the
shards that became active.  The eight classes are the bucket widths 8, 16, ..., 1024; with
`finish=True` (the engine's learned-finish option) class 7 instead finishes the solve as
Bellman-Ford pushes and strip sweeps.  `train` fits the network with the engine's own HIP
forward/backward/Adam kernels; `Engine.sssp(method=SSSP_LEARNED)` then runs inference on the GPU
at every epoch.  The network only picks how the next epoch relaxes, so distances equal those of
every other method.
"""
import numpy as np

from . import OPT_LEARNED_FINISH, OPT_MLP_NET, SSSP_BF, SSSP_DELTA, SSSP_LEARNED, Engine  # noqa: F401


def record(engine, graphs, sources_per_graph=2, seed=0, finish=False):
    """Traces from SSSP runs on R-MAT graphs given as (scale, edge_factor, seed) tuples; with
    finish=True class 7 is the Bellman-Ford finish instead of width 1024."""
    rows = []
    rng = np.random.default_rng(seed)
    saved = engine.get_option(OPT_LEARNED_FINISH)
    engine.set_option(OPT_LEARNED_FINISH, int(finish))
    try:
        for scale, ef, gseed in graphs:
            engine.generate_rmat(scale, ef, gseed, weighted=True)
            # build the layouts every action uses (out-CSR, hub table, source strips) first
            src0 = engine.pick_source(gseed)
            engine.sssp(src0, method=SSSP_BF, want_dist=False)
            engine.sssp(src0, method=SSSP_DELTA, delta=64, want_dist=False)
            for _ in range(sources_per_graph):
                src = engine.pick_source(int(rng.integers(1 << 30)))
                r, _, _ = engine.sssp_trace(src, want_dist=False)
                rows.append(r)
    finally:
        engine.set_option(OPT_LEARNED_FINISH, saved)
    return np.concatenate(rows) if rows else np.zeros(0)


def labels(rows, tol=0.1):
    """Training label of every row, from the rollout times of all width classes.

    Each row's times are taken relative to its fastest class.  Rows with identical features (the
    first epoch of every run looks the same: the far pile holds only the source) must share one
    decision, so they are pooled and get the class with the smallest summed relative time.  Among
    classes within `tol` of the best, the narrowest is taken: neighbouring widths often tie on
    small graphs, and narrow early buckets are what pays off on larger graphs (where the
    rollouts separate them), so a wider class has to win clearly to become the label."""
    ms = np.asarray(rows["ms"], dtype=np.float64)
    if not len(ms):
        return np.zeros(0, dtype=np.int32)
    rel = ms / ms.min(axis=1, keepdims=True)
    feat = np.ascontiguousarray(rows["feat"], dtype=np.float32)
    _, group = np.unique(feat.view(np.dtype((np.void, feat.dtype.itemsize * feat.shape[1]))).ravel(),
                         return_inverse=True)
    cost = np.zeros((group.max() + 1, ms.shape[1]))
    np.add.at(cost, group, rel)
    cost = cost[group]
    ok = cost <= cost.min(axis=1, keepdims=True) * (1.0 + tol)
    return ok.argmax(axis=1).astype(np.int32)


def dataset(rows, tol=0.1):
    """(X, labels, shard masks) of a trace."""
    return (np.ascontiguousarray(rows["feat"], dtype=np.float32),
            np.ascontiguousarray(labels(rows, tol)),
            np.ascontiguousarray(rows["shard_mask"], dtype=np.uint64))


def train(engine, rows, epochs=400, lr=3e-3, init_seed=7):
    X, y, m = dataset(rows)
    engine.mlp_init(init_seed)
    return engine.mlp_train(X, y, m, epochs=epochs, lr=lr)


def accuracy(engine, rows):
    """Width-class accuracy and shard precision/recall of the trained network on a trace."""
    X, y, m = dataset(rows)
    Y = engine.mlp_forward(X)
    acc = float((Y[:, :8].argmax(axis=1) == y).mean()) if len(y) else 0.0
    pred = Y[:, 8:] > 0
    bits = ((m[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool)
    tp = float((pred & bits).sum())
    prec = tp / max(1.0, float(pred.sum()))
    rec = tp / max(1.0, float(bits.sum()))
    return {"class_accuracy": acc, "shard_precision": prec, "shard_recall": rec}


# ---- the semi-external window predictor (the engine's second network) ----------------------
#
# On a semi-external graph, min sweeps (CC labels, SSSP distances) run only the windows whose
# sources changed in the sweep before; the host learns which ones only after that sweep ends.
# The window predictor sees 16 features of a sweep's exact activation (active fraction per eighth
# of the windows, changed fraction of the source blocks, ...) and scores the 64 shards of the row
# range for the sweep after; with GMP_OPT_WINDOW_PREFETCH the windows of its shards are copied
# ahead and run first.  Its rows carry no class (label -1): only the shard outputs train.


def record_windows(engine, kinds=(0, 2), sources=(0,)):
    """Trace rows of the window predictor on the engine's semi-external graph: for each kind (0
    labels, 1 unit, 2 weighted distances) and source."""
    rows = []
    for kind in kinds:
        for s in (sources if kind else (0,)):
            r, _ = engine.window_trace(kind, int(s))
            rows.append(r)
    return np.concatenate(rows) if rows else np.zeros(0)


def train_windows(engine, rows, epochs=400, lr=3e-3, init_seed=11):
    """Fits the window predictor (network 1) on window trace rows; network 0 stays selected."""
    X = np.ascontiguousarray(rows["feat"], dtype=np.float32)
    y = np.full(len(rows), -1, dtype=np.int32)
    m = np.ascontiguousarray(rows["shard_mask"], dtype=np.uint64)
    engine.set_option(OPT_MLP_NET, 1)
    try:
        engine.mlp_init(init_seed)
        return engine.mlp_train(X, y, m, epochs=epochs, lr=lr)
    finally:
        engine.set_option(OPT_MLP_NET, 0)


def window_accuracy(engine, rows):
    """Shard precision and recall of the window predictor on window trace rows."""
    engine.set_option(OPT_MLP_NET, 1)
    try:
        Y = engine.mlp_forward(np.ascontiguousarray(rows["feat"], dtype=np.float32))
    finally:
        engine.set_option(OPT_MLP_NET, 0)
    pred = Y[:, 8:] > 0
    m = np.asarray(rows["shard_mask"], dtype=np.uint64)
    bits = ((m[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(bool)
    tp = float((pred & bits).sum())
    return {"shard_precision": tp / max(1.0, float(pred.sum())), "shard_recall": tp / max(1.0, float(bits.sum()))}
