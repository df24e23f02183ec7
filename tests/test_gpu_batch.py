"""Batched multi-instance runs (many graphs resident, many SSSP sources and personalised
PageRank seeds in one launch sequence) against the oracle, instance by instance."""
import numpy as np
import pytest

import oracle
from tests import graphs

pytestmark = pytest.mark.gpu


def _batch(engine, specs):
    print("batch", specs, flush=True)
    engine.batch_generate(specs)
    return [engine.batch_download(g) for g in range(len(specs))]


SPECS = [dict(scale=8, edge_factor=8, seed=5000), dict(scale=11, edge_factor=32, seed=5001),
         dict(scale=9, edge_factor=16, seed=5002), dict(scale=12, edge_factor=8, seed=5003),
         dict(scale=6, edge_factor=4, seed=5004)]


def test_batch_graphs_match_single_generation(engine):
    parts = _batch(engine, SPECS)
    from oracle import rmat_ref
    for (off, idx, w), s in zip(parts, SPECS):
        V, roff, ridx = rmat_ref.in_csr(s["scale"], s["edge_factor"], s["seed"])
        np.testing.assert_array_equal(off, roff)
        np.testing.assert_array_equal(idx, ridx)
        rows = np.repeat(np.arange(V, dtype=np.uint64), np.diff(off))
        np.testing.assert_array_equal(w, rmat_ref.weights(idx, rows, s["seed"]))


def test_batch_sssp_matches_dijkstra(engine):
    parts = _batch(engine, SPECS)
    rng = np.random.default_rng(1)
    gsel = np.array([1, 1, 1, 0, 3, 3, 2, 4, 1, 3, 3, 3, 0, 2, 1, 1, 4, 3, 1, 1], dtype=np.uint32)  # uneven
    srcs = np.array([rng.integers(0, parts[g][0].size - 1) for g in gsel], dtype=np.uint32)
    dists, st = engine.batch_sssp(gsel, srcs)
    assert st.iterations > 0 and st.launches > 0
    for i, (g, s) in enumerate(zip(gsel, srcs)):
        off, idx, w = parts[g]
        V = off.size - 1
        ooff, oidx, ow = graphs.transpose(V, off, idx, w)
        np.testing.assert_array_equal(dists[i], oracle.dijkstra(ooff, oidx, ow, int(s)), err_msg=f"instance {i}")


def test_batch_ppr_matches_fp64(engine):
    parts = _batch(engine, SPECS)
    rng = np.random.default_rng(2)
    # 70 instances on graph 2 (more than one 64-wide group), a few elsewhere
    gsel = np.array([2] * 70 + [0, 1, 1, 3, 4], dtype=np.uint32)
    seeds = np.array([rng.integers(0, parts[g][0].size - 1) for g in gsel], dtype=np.uint32)
    ranks, st = engine.batch_ppr(gsel, seeds, damping=0.85, iters=30)
    assert st.iterations == 30
    for i, (g, s) in enumerate(zip(gsel, seeds)):
        off, idx, _ = parts[g]
        V = off.size - 1
        x, _, _ = oracle.pr_f64(off, idx, graphs.out_degree(V, idx), d=0.85, max_iter=30, seed=int(s))
        l1 = np.abs(ranks[i].astype(np.float64) - x).sum()
        assert l1 < 1e-6, (i, l1)


def test_batch_ppr_hub_rows_and_reproducible(engine):
    """Rows above the hub degree (split into pieces and combined) with 2-slot and 1-slot
    instance groups; the task queue hands tasks out in any order, and the results must still
    match FP64 and repeat bit for bit."""
    specs = [dict(scale=15, edge_factor=32, seed=5010), dict(scale=10, edge_factor=16, seed=5011)]
    parts = _batch(engine, specs)
    off0 = parts[0][0]
    assert int(np.diff(off0).max()) > 5 * 1024  # several pieces per hub row
    rng = np.random.default_rng(3)
    gsel = np.array([0] * 23 + [1] * 70, dtype=np.uint32)
    seeds = np.array([rng.integers(0, parts[g][0].size - 1) for g in gsel], dtype=np.uint32)
    seeds[0] = int(np.argmax(np.diff(off0)))  # a seed on the largest hub row
    ranks, _ = engine.batch_ppr(gsel, seeds, damping=0.85, iters=25)
    again, _ = engine.batch_ppr(gsel, seeds, damping=0.85, iters=25)
    np.testing.assert_array_equal(np.concatenate(ranks).view(np.uint32), np.concatenate(again).view(np.uint32))
    for i in list(range(0, 23, 5)) + [23, 60, 87, 92]:
        g, s = gsel[i], seeds[i]
        off, idx, _ = parts[g]
        V = off.size - 1
        x, _, _ = oracle.pr_f64(off, idx, graphs.out_degree(V, idx), d=0.85, max_iter=25, seed=int(s))
        l1 = np.abs(ranks[i].astype(np.float64) - x).sum()
        assert l1 < 1e-6, (i, l1)


def test_batch_layouts_agree(engine):
    """Graphs stored in their degree-ordered layout (the default) or in their own numbering: the
    same rows come back, SSSP distances are identical and personalised ranks agree within the
    FP64 bound; sources and seeds are given in the graphs' own numbering either way."""
    import graphee_amd as gmp
    specs = SPECS[:4]
    rng = np.random.default_rng(4)
    gsel = np.array([0, 1, 1, 2, 3, 3, 1], dtype=np.uint32)
    out = []
    for reorder in (1, 0):
        engine.set_option(gmp.OPT_REORDER, reorder)
        parts = _batch(engine, specs)
        if not out:
            srcs = np.array([rng.integers(0, parts[g][0].size - 1) for g in gsel], dtype=np.uint32)
        d, _ = engine.batch_sssp(gsel, srcs)
        r, _ = engine.batch_ppr(gsel, srcs, damping=0.85, iters=20)
        out.append((parts, d, r))
    engine.set_option(gmp.OPT_REORDER, 1)
    (p1, d1, r1), (p0, d0, r0) = out
    for a, b in zip(p1, p0):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    for i in range(gsel.size):
        np.testing.assert_array_equal(d1[i], d0[i])
        assert np.abs(r1[i].astype(np.float64) - r0[i]).sum() < 1e-6, i
