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


@pytest.mark.parametrize("scan_log,hub_log", [(0, 0), (31, 0), (0, 6), (31, 6), (4, 8)])
def test_batch_sssp_frontier_walk_and_hub_pieces(engine, scan_log, hub_log):
    """The frontier taken from the queue or from the mask array in group-vertex order (scan_log 31:
    every round walks the masks), and entries above 2^hub_log out-edges relaxed in pieces by the
    second launch (hub_log 6: hubs of 65 out-edges and more, many pieces each): every instance
    keeps Dijkstra's distances, alone and in the combined launch with PPR instances."""
    import graphee_amd as gmp
    parts = _batch(engine, SPECS)
    rng = np.random.default_rng(21)
    gsel = np.array([1] * 40 + [3] * 9 + [2, 0, 4], dtype=np.uint32)
    srcs = np.array([rng.integers(0, parts[g][0].size - 1) for g in gsel], dtype=np.uint32)
    keys = (gmp.OPT_BATCH_SSSP_SCAN_LOG, gmp.OPT_BATCH_SSSP_HUB_LOG)
    saved = [engine.get_option(k) for k in keys]
    try:
        engine.set_option(gmp.OPT_BATCH_SSSP_SCAN_LOG, scan_log)
        engine.set_option(gmp.OPT_BATCH_SSSP_HUB_LOG, hub_log)
        dists, st = engine.batch_sssp(gsel, srcs)
        both, _, _, _ = engine.batch_run(gsel, srcs, [1, 2], [3, 4], iters=3)
    finally:
        for k, v in zip(keys, saved):
            engine.set_option(k, v)
    for i, (g, s) in enumerate(zip(gsel, srcs)):
        off, idx, w = parts[g]
        ooff, oidx, ow = graphs.transpose(off.size - 1, off, idx, w)
        ref = oracle.dijkstra(ooff, oidx, ow, int(s))
        np.testing.assert_array_equal(dists[i], ref, err_msg=f"instance {i}")
        np.testing.assert_array_equal(both[i], ref, err_msg=f"combined, instance {i}")


def test_batch_sssp_groups_of_many_instances(engine):
    """Instances of one graph relax together in groups of up to 32 (vertex-major distances,
    frontier entries with instance masks): 70 instances on one graph make groups of 32, 32 and 6,
    some share a source (inside a group and across groups), a 3-instance group is padded to 4, a
    graph has a single instance, and the instances are interleaved in the caller's order."""
    parts = _batch(engine, SPECS)
    rng = np.random.default_rng(11)
    V1 = parts[1][0].size - 1
    many = [(1, int(x)) for x in rng.integers(0, V1, 60)] + [(1, 7)] * 5 + [(1, 0)] * 5
    rest = [(4, 3), (4, 3), (4, int(rng.integers(0, parts[4][0].size - 1))), (0, 1)]
    inst = many + rest
    order = rng.permutation(len(inst))
    gsel = np.array([inst[i][0] for i in order], dtype=np.uint32)
    srcs = np.array([inst[i][1] for i in order], dtype=np.uint32)
    dists, st = engine.batch_sssp(gsel, srcs)
    assert st.iterations > 0
    ref = {}
    for i, (g, s) in enumerate(zip(gsel, srcs)):
        key = (int(g), int(s))
        if key not in ref:
            off, idx, w = parts[g]
            ooff, oidx, ow = graphs.transpose(off.size - 1, off, idx, w)
            ref[key] = oracle.dijkstra(ooff, oidx, ow, int(s))
        np.testing.assert_array_equal(dists[i], ref[key], err_msg=f"instance {i} graph {g} source {s}")


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


def test_batch_run_one_graph_for_both_kinds(engine):
    """SSSP and PPR instances in one replayed graph with two concurrent branches: the distances
    equal Dijkstra's and the ranks equal gmp_batch_ppr's bit for bit (same kernels, order-free
    reductions), for an odd iteration count too; either kind may be absent."""
    parts = _batch(engine, SPECS)
    rng = np.random.default_rng(5)
    gs = np.array([1, 1, 3, 0, 2, 4, 1, 3], dtype=np.uint32)
    srcs = np.array([rng.integers(0, parts[g][0].size - 1) for g in gs], dtype=np.uint32)
    gp = np.array([2] * 9 + [1, 0, 3], dtype=np.uint32)
    seeds = np.array([rng.integers(0, parts[g][0].size - 1) for g in gp], dtype=np.uint32)
    for iters in (7, 30):
        dists, ranks, st_s, st_p = engine.batch_run(gs, srcs, gp, seeds, damping=0.85, iters=iters)
        assert st_p.iterations == iters and st_s.ms_total == st_p.ms_total > 0
        alone, _ = engine.batch_ppr(gp, seeds, damping=0.85, iters=iters)
        for i in range(gp.size):
            np.testing.assert_array_equal(ranks[i].view(np.uint32), alone[i].view(np.uint32), err_msg=f"ppr {i}")
        for i, (g, s) in enumerate(zip(gs, srcs)):
            off, idx, w = parts[g]
            V = off.size - 1
            ooff, oidx, ow = graphs.transpose(V, off, idx, w)
            np.testing.assert_array_equal(dists[i], oracle.dijkstra(ooff, oidx, ow, int(s)), err_msg=f"sssp {i}")
    d_only, r_none, _, _ = engine.batch_run(gs[:2], srcs[:2], [], [], iters=5)
    assert len(r_none) == 0
    off, idx, w = parts[gs[0]]
    ooff, oidx, ow = graphs.transpose(off.size - 1, off, idx, w)
    np.testing.assert_array_equal(d_only[0], oracle.dijkstra(ooff, oidx, ow, int(srcs[0])))
    _, r_only, _, _ = engine.batch_run([], [], gp[:3], seeds[:3], iters=4)
    alone, _ = engine.batch_ppr(gp[:3], seeds[:3], damping=0.85, iters=4)
    np.testing.assert_array_equal(np.concatenate(r_only).view(np.uint32), np.concatenate(alone).view(np.uint32))


def test_batch_autotune_keeps_results(engine):
    """The batch tuner times every knob value on the given instances and applies the best; any
    value gives Dijkstra's distances and ranks within 1e-6 L1 of FP64 (forced extremes too)."""
    import graphee_amd as gmp
    parts = _batch(engine, SPECS)
    rng = np.random.default_rng(6)
    gs = np.array([1, 1, 3, 0, 2], dtype=np.uint32)
    srcs = np.array([rng.integers(0, parts[g][0].size - 1) for g in gs], dtype=np.uint32)
    gp = np.array([2] * 5 + [1, 3], dtype=np.uint32)
    seeds = np.array([rng.integers(0, parts[g][0].size - 1) for g in gp], dtype=np.uint32)
    keys = (gmp.OPT_BATCH_SSSP_UNROLL, gmp.OPT_BATCH_SSSP_GRID, gmp.OPT_BATCH_PPR_PIECE, gmp.OPT_BATCH_PPR_RUN_ROWS,
            gmp.OPT_BATCH_PPR_RUN_EDGES, gmp.OPT_BATCH_PPR_LOADS, gmp.OPT_BATCH_PPR_GRID,
            gmp.OPT_BATCH_SSSP_SCAN_LOG, gmp.OPT_BATCH_SSSP_HUB_LOG)
    saved = [engine.get_option(k) for k in keys]

    def check_results():
        d, _ = engine.batch_sssp(gs, srcs)
        r, _ = engine.batch_ppr(gp, seeds, damping=0.85, iters=15)
        for i, (g, s) in enumerate(zip(gs, srcs)):
            off, idx, w = parts[g]
            ooff, oidx, ow = graphs.transpose(off.size - 1, off, idx, w)
            np.testing.assert_array_equal(d[i], oracle.dijkstra(ooff, oidx, ow, int(s)))
        for i, (g, s) in enumerate(zip(gp, seeds)):
            off, idx, _ = parts[g]
            V = off.size - 1
            x, _, _ = oracle.pr_f64(off, idx, graphs.out_degree(V, idx), d=0.85, max_iter=15, seed=int(s))
            assert np.abs(r[i].astype(np.float64) - x).sum() < 1e-6, i

    try:
        res = engine.batch_autotune(gs, srcs, gp, seeds, iters=15)
        print("batch tune", res, flush=True)
        assert res["trials"] > 10 and res["sssp_ms_after"] <= res["sssp_ms_before"]
        assert [engine.get_option(k) for k in keys] == [res["sssp_unroll"], res["sssp_grid_per_cu"], res["ppr_piece"],
                                                        res["ppr_run_rows"], res["ppr_run_edges"], res["ppr_loads"],
                                                        res["ppr_grid_per_cu"], res["sssp_scan_log"],
                                                        res["sssp_hub_log"]]
        check_results()
        # the queue and no hub list; the mask array every round and hubs above 64 out-edges
        for vals in ((2, 1, 64, 1, 64, 4, 1, 0, 0), (8, 64, 65536, 1024, 65536, 16, 64, 31, 6)):
            for k, v in zip(keys, vals):
                engine.set_option(k, v)
            check_results()
        with pytest.raises(gmp.GmpError):
            engine.set_option(gmp.OPT_BATCH_PPR_LOADS, 5)
        with pytest.raises(gmp.GmpError):
            engine.set_option(gmp.OPT_BATCH_SSSP_HUB_LOG, 5)
    finally:
        for k, v in zip(keys, saved):
            engine.set_option(k, v)
