"""Parity of the HIP engine (through the C ABI) against the CPU oracle.

Integer results (SSSP distances, CC labels, the GraphMP-compatible apps) must match bit
for bit; standard PageRank must be within 1e-6 L1 of the FP64 oracle (the FP32 ranks
carry about 1e-7 relative rounding per vertex)."""
import sys

import numpy as np
import pytest

import graphee_amd as gmp
import oracle
from oracle import rmat_ref
from tests import graphs

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gen(engine, scale, ef=16, seed=None, **kw):
    seed = 1000 + scale if seed is None else seed
    say(f"generate R-MAT scale={scale} ef={ef} seed={seed} {kw}")
    engine.generate_rmat(scale, ef, seed, **kw)
    return engine.download(weights=kw.get("weighted", False))


# ---- graph construction ----------------------------------------------------------------

@pytest.mark.parametrize("scale,undirected,weighted,off64",
                         [(8, False, False, False), (11, False, True, False), (10, True, False, False),
                          (9, False, True, True)])
def test_generator_matches_definition(engine, scale, undirected, weighted, off64):
    g = gen(engine, scale, 16, undirected=undirected, weighted=weighted, off64=off64, num_shards=4)
    V, off, idx = rmat_ref.in_csr(scale, 16, 1000 + scale, symmetric=undirected)
    np.testing.assert_array_equal(g["off"], off)
    np.testing.assert_array_equal(g["idx"], idx)
    np.testing.assert_array_equal(g["outdeg"], graphs.out_degree(V, idx))
    np.testing.assert_array_equal(g["bounds"], oracle.partition_bounds(np.diff(off), 4))
    if weighted:
        rows = np.repeat(np.arange(V, dtype=np.uint64), np.diff(off))
        np.testing.assert_array_equal(g["w"], rmat_ref.weights(idx, rows, 1000 + scale))


def test_pick_source_matches_definition(engine):
    g = gen(engine, 10)
    assert engine.pick_source(1010, 3) == rmat_ref.pick_source(1 << 10, g["outdeg"], 1010, 3)


# ---- standard PageRank -----------------------------------------------------------------

@pytest.mark.parametrize("kahan", [True, False])
def test_pagerank_20_iterations_vs_fp64(engine, kahan):
    g = gen(engine, 14)
    say(f"pagerank 20 iterations kahan={kahan}")
    r, st = engine.pagerank(max_iter=20, kahan=kahan)
    x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=20)
    l1 = np.abs(r.astype(np.float64) - x).sum()
    assert l1 < PR_L1_TOL, l1
    assert st.iterations == 20 and st.edges == 20 * g["idx"].size
    assert abs(st.dangling - x[g["outdeg"] == 0].sum()) < 1e-3


def test_pagerank_tolerance_stop_vs_fp64(engine):
    g = gen(engine, 13)
    say("pagerank until L1 change < 1e-7")
    r, st = engine.pagerank(max_iter=100, tol=1e-7)
    x, it, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=100, tol=1e-7)
    assert abs(st.iterations - it) <= 1 and st.residual < 1e-7
    assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
    assert abs(r.astype(np.float64).sum() - 1.0) < 1e-5


def test_personalised_pagerank_vs_fp64(engine):
    g = gen(engine, 12)
    seed = int(np.argmax(g["outdeg"]))
    say(f"personalised pagerank seed={seed}")
    r, _ = engine.pagerank(max_iter=30, seed_vertex=seed)
    x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=30, seed=seed)
    assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL


def test_pagerank_multichunk_hub_rows(engine):
    """Rows longer than one chunk take the partial + combine path."""
    V = 40000
    rng = np.random.default_rng(5)
    src = np.concatenate([np.arange(1, V), rng.integers(0, V, 100000), np.arange(2, 9000)])
    dst = np.concatenate([np.zeros(V - 1, np.int64), rng.integers(0, V, 100000), np.ones(8998, np.int64)])
    off, idx = graphs.csr_from_edges(V, src, dst)
    say(f"pagerank hub rows: in-degree of 0 = {off[1]}, of 1 = {off[2] - off[1]}")
    engine.load_csr(off, idx)
    r, _ = engine.pagerank(max_iter=25)
    x, _, _ = oracle.pr_f64(off, idx, graphs.out_degree(V, idx), max_iter=25)
    assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL


# ---- GraphMP-compatible apps: bit-exact ------------------------------------------------

def test_compat_pagerank_bit_exact(engine):
    g = gen(engine, 11)
    V = g["outdeg"].size
    vout = g["outdeg"].astype(np.int32)
    msg0 = np.full(V, 1.0 / V, dtype=np.float32)
    say("compat pagerank (pagerank.cpp rule), 200 steps max")
    msg, val, st = engine.run_app(gmp.APP_PAGERANK, msg0, 200, vertex_out=vout)
    want, wval, it, changed = oracle.app_sync(oracle.APP_PAGERANK, g["off"], g["idx"], msg0, 200, vout=vout)
    np.testing.assert_array_equal(msg.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(val.view(np.uint32), wval.view(np.uint32))
    assert st.iterations == it and st.changed == changed[-1]


@pytest.mark.parametrize("app", [gmp.APP_SSSP, gmp.APP_CC, gmp.APP_SSSP_W])
def test_compat_min_apps_bit_exact(engine, app):
    g = gen(engine, 12, weighted=(app == gmp.APP_SSSP_W))
    V = g["outdeg"].size
    if app == gmp.APP_CC:
        msg0 = np.arange(V, dtype=np.int32)
    elif app == gmp.APP_SSSP:
        msg0 = np.full(V, 10000, dtype=np.int32)
        msg0[1] = 0
    else:
        msg0 = np.full(V, 0xFFFFFFFF, dtype=np.uint32)
        msg0[engine.pick_source(1012)] = 0
    say(f"compat app {app}")
    msg, val, st = engine.run_app(app, msg0, 1000)
    want, wval, it, _ = oracle.app_sync(app, g["off"], g["idx"], msg0, 1000, w=g["w"])
    np.testing.assert_array_equal(msg, want)
    np.testing.assert_array_equal(val, wval)
    assert st.iterations == it


# ---- SSSP and CC -----------------------------------------------------------------------

@pytest.mark.parametrize("method,delta", [(gmp.SSSP_BF, 0), (gmp.SSSP_DELTA, 64), (gmp.SSSP_DELTA, 1),
                                          (gmp.SSSP_DELTA, 1000)])
def test_sssp_matches_dijkstra(engine, method, delta):
    g = gen(engine, 13, weighted=True)
    V = g["outdeg"].size
    src = engine.pick_source(1013)
    say(f"sssp method={method} delta={delta} src={src}")
    d, st = engine.sssp(src, method=method, delta=delta)
    ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
    np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, ow, src))
    assert st.edges > 0


@pytest.mark.parametrize("strip_log,bin_log,task_log", [(0, 14, 20), (8, 8, 8), (11, 10, 10), (15, 14, 8)])
def test_sssp_bellman_ford_sweeps_and_frontier_agree(engine, strip_log, bin_log, task_log):
    """Bellman-Ford by sweeps over the strip layout (any strip width, bin size, split size;
    weighted, unit weights and a weight that saturates) and by frontier relaxation (strip_log 0)
    give Dijkstra's distances, and the sweeps skip strips without changes."""
    for k, v in zip((gmp.OPT_STRIP_LOG, gmp.OPT_BIN_LOG, gmp.OPT_BIN_TASK_LOG), (strip_log, bin_log, task_log)):
        engine.set_option(k, v)
    try:
        g = gen(engine, 12, weighted=True)
        V = g["outdeg"].size
        ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
        for s in (engine.pick_source(1), engine.pick_source(2)):
            say(f"bf strips {strip_log}/{bin_log}/{task_log} src={s}")
            d, st = engine.sssp(s, method=gmp.SSSP_BF)
            np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, ow, s))
            d, _ = engine.sssp(s, method=gmp.SSSP_BF, unit_weights=True)
            np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, None, s))
            assert st.edges <= st.iterations * g["idx"].size
            if strip_log and V > (1 << strip_log):
                assert st.edges < st.iterations * g["idx"].size  # the first sweeps skip idle strips
        # a path of 2000 edges with one weight near 2^32 (distances saturate at 2^32 - 2), plus
        # an unreachable part
        V = 5000
        src_ = np.concatenate([np.arange(0, 2000), np.arange(3000, 3999)])
        dst_ = np.concatenate([np.arange(1, 2001), np.arange(3001, 4000)])
        w_ = np.ones(src_.size, np.uint32)
        w_[1500] = 0xFFFFFFF0
        order = np.lexsort((src_, dst_))
        off, idx = graphs.csr_from_edges(V, src_, dst_)
        engine.load_csr(off, idx, w_[order])
        d, st = engine.sssp(0, method=gmp.SSSP_BF)
        ooff, oidx, ow = graphs.transpose(V, off, idx, w_[order])
        np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, ow, 0))
        assert d[2000] == 0xFFFFFFFE and d[3500] == 0xFFFFFFFF
    finally:
        for k, v in STRIP_DEFAULTS:
            engine.set_option(k, v)


@pytest.mark.parametrize("pull_log", [0, 2, 31])
@pytest.mark.parametrize("strip_log,bin_log,task_log", [(15, 14, 20), (8, 8, 8), (11, 10, 10)])
def test_delta_stepping_pull_rounds_keep_distances(engine, pull_log, strip_log, bin_log, task_log):
    """Delta-stepping rounds that pull over the source strips (only the near frontier's sources
    offer, the bins append to the next frontier and the far pile as a push would) in place of
    pushes -- pull_log 0: never, 2: rounds above E/4 frontier out-edges, 31: every round -- give
    Dijkstra's distances for narrow, medium and wide buckets, stored and unit weights, and on
    strip widths, bin sizes and split bins of every kind."""
    keys = (gmp.OPT_DELTA_PULL_LOG, gmp.OPT_STRIP_LOG, gmp.OPT_BIN_LOG, gmp.OPT_BIN_TASK_LOG)
    saved = [engine.get_option(k) for k in keys]
    for k, v in zip(keys, (pull_log, strip_log, bin_log, task_log)):
        engine.set_option(k, v)
    try:
        g = gen(engine, 13, weighted=True)
        V = g["outdeg"].size
        ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
        for s in (engine.pick_source(3), engine.pick_source(4)):
            want = oracle.dijkstra(ooff, oidx, ow, s)
            for delta in (1, 64, 1000):
                say(f"pull_log {pull_log} strips {strip_log}/{bin_log}/{task_log}: delta {delta} src={s}")
                d, st = engine.sssp(s, method=gmp.SSSP_DELTA, delta=delta)
                np.testing.assert_array_equal(d, want, err_msg=f"delta {delta}")
                if pull_log == 0:
                    assert st.changed == 0  # no pull round
                if pull_log == 31:
                    assert st.changed > 0  # pull rounds ran
            d, _ = engine.sssp(s, method=gmp.SSSP_DELTA, delta=8, unit_weights=True)
            np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, None, s))
    finally:
        for k, v in zip(keys, saved):
            engine.set_option(k, v)


def test_sssp_unit_weights_is_bfs(engine):
    g = gen(engine, 12)
    V = g["outdeg"].size
    src = engine.pick_source(1012)
    say(f"sssp unit weights src={src}")
    d, _ = engine.sssp(src, method=gmp.SSSP_DELTA, delta=2)
    ooff, oidx, _ = graphs.transpose(V, g["off"], g["idx"])
    np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, None, src))


@pytest.mark.parametrize("scale", [10, 14])
def test_cc_matches_unionfind(engine, scale):
    g = gen(engine, scale, undirected=True)
    say(f"cc scale={scale}")
    lab, st = engine.cc()
    np.testing.assert_array_equal(lab, oracle.cc_unionfind(g["off"], g["idx"]))
    assert st.iterations >= 1


@pytest.mark.parametrize("strip_log,bin_log,task_log", [(0, 14, 20), (8, 8, 8), (11, 10, 10), (15, 14, 8)])
def test_cc_sweeps_and_frontier_agree(engine, strip_log, bin_log, task_log):
    """CC by min-label sweeps over the strip layout (any strip width, bin size, split size) and
    by frontier relaxation (strip_log 0) give the union-find labels, also on a graph of many
    components with a long path (many sweeps) and with a hub."""
    for k, v in zip((gmp.OPT_STRIP_LOG, gmp.OPT_BIN_LOG, gmp.OPT_BIN_TASK_LOG), (strip_log, bin_log, task_log)):
        engine.set_option(k, v)
    engine.set_option(gmp.OPT_CC_UNION_FIND, 0)  # the sweeps, also on the symmetric R-MAT graph
    try:
        g = gen(engine, 12, undirected=True)
        say(f"cc strips {strip_log}/{bin_log}/{task_log}")
        lab, _ = engine.cc()
        np.testing.assert_array_equal(lab, oracle.cc_unionfind(g["off"], g["idx"]))
        V = 30000
        rng = np.random.default_rng(11)
        a = np.concatenate([np.arange(0, 3000), rng.integers(5000, V, 20000), np.full(4000, 29999)])
        b = np.concatenate([np.arange(1, 3001), rng.integers(5000, V, 20000), rng.integers(5000, V, 4000)])
        src, dst = np.concatenate([a, b]), np.concatenate([b, a])  # symmetric
        off, idx = graphs.csr_from_edges(V, src, dst)
        engine.load_csr(off, idx)
        lab, st = engine.cc()
        np.testing.assert_array_equal(lab, oracle.cc_unionfind(off, idx))
        assert st.iterations >= (3000 if strip_log == 0 else 2)
    finally:
        engine.set_option(gmp.OPT_CC_UNION_FIND, 1)
        for k, v in STRIP_DEFAULTS:
            engine.set_option(k, v)


@pytest.mark.parametrize("scale,seed", [(10, 3), (13, 4), (15, 5)])
def test_cc_union_find_on_symmetric_graphs(engine, tmp_path, scale, seed):
    """CC of a symmetric graph by sampled union-find (the default there) gives the union-find
    oracle's labels -- each component's least original id -- and the sweeps' labels; also for an
    edge list loaded with symmetrize, and with the degree-ordered layout off."""
    g = gen(engine, scale, seed=seed, undirected=True)
    want = oracle.cc_unionfind(g["off"], g["idx"])
    say(f"cc union-find s{scale}")
    lab, st = engine.cc()
    np.testing.assert_array_equal(lab, want)
    assert st.edges < g["idx"].size  # the giant component skips its remaining edges
    engine.set_option(gmp.OPT_CC_UNION_FIND, 0)
    try:
        lab2, _ = engine.cc()
    finally:
        engine.set_option(gmp.OPT_CC_UNION_FIND, 1)
    np.testing.assert_array_equal(lab2, want)
    engine.set_option(gmp.OPT_REORDER, 0)
    try:
        lab3, _ = engine.cc()
    finally:
        engine.set_option(gmp.OPT_REORDER, 1)
    np.testing.assert_array_equal(lab3, want)
    # many small components plus a path, from a symmetrized edge list
    rng = np.random.default_rng(seed)
    V = 5000
    src = np.concatenate([rng.integers(0, V, 3000), np.arange(4000, 4999)])
    dst = np.concatenate([rng.integers(0, V, 3000), np.arange(4001, 5000)])
    p = tmp_path / "e.tsv"
    p.write_text("".join(f"{a} {b}\n" for a, b in zip(src, dst)))
    engine.load_edge_list(p, V=V, symmetrize=True)
    h = engine.download()
    lab4, _ = engine.cc()
    np.testing.assert_array_equal(lab4, oracle.cc_unionfind(h["off"], h["idx"]))


@pytest.mark.parametrize("push_log", [0, 1, 3, 8])
def test_push_sweep_switch_keeps_results(engine, push_log):
    """Bellman-Ford and CC switching between frontier pushes (frontiers of at most V >> push_log
    vertices) and strip sweeps (larger ones) -- push_log 0: sweeps only, 1: nearly all pushes --
    give Dijkstra's distances, BFS levels and the union-find labels, also on a long path that
    switches back and forth."""
    saved = engine.get_option(gmp.OPT_PUSH_LOG)
    engine.set_option(gmp.OPT_PUSH_LOG, push_log)
    try:
        g = gen(engine, 13, weighted=True)
        V = g["outdeg"].size
        ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
        for s in (engine.pick_source(5), engine.pick_source(6)):
            say(f"push_log {push_log}: bf src={s}")
            d, st = engine.sssp(s, method=gmp.SSSP_BF)
            np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, ow, s))
            d, _ = engine.sssp(s, method=gmp.SSSP_BF, unit_weights=True)
            np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, None, s))
            if push_log == 0:
                assert st.epochs == st.iterations  # every step a sweep
            if push_log == 1:
                assert st.epochs < st.iterations  # some steps pushed
        g = gen(engine, 13, undirected=True)
        lab, st = engine.cc()
        np.testing.assert_array_equal(lab, oracle.cc_unionfind(g["off"], g["idx"]))
        # a path with side trees: sparse fronts, dense bursts, back to sparse
        V = 40000
        rng = np.random.default_rng(3)
        a = np.concatenate([np.arange(0, 4000), rng.integers(4000, V, 30000)])
        b = np.concatenate([np.arange(1, 4001), rng.integers(4000, V, 30000)])
        src_, dst_ = np.concatenate([a, b]), np.concatenate([b, a])
        off, idx = graphs.csr_from_edges(V, src_, dst_)
        w = graphs.edge_weights(off, idx, 4)
        engine.load_csr(off, idx, w)
        lab, _ = engine.cc()
        np.testing.assert_array_equal(lab, oracle.cc_unionfind(off, idx))
        ooff, oidx, ow = graphs.transpose(V, off, idx, w)
        for s in (0, 2000, 39999):
            d, _ = engine.sssp(s, method=gmp.SSSP_BF)
            np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, ow, s))
    finally:
        engine.set_option(gmp.OPT_PUSH_LOG, saved)


def test_off64_offsets_path(engine):
    g = gen(engine, 11, weighted=True, off64=True)
    V = g["outdeg"].size
    say("off64: pagerank, sssp")
    r, _ = engine.pagerank(max_iter=15)
    x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=15)
    assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
    src = engine.pick_source(1011)
    d, _ = engine.sssp(src, method=gmp.SSSP_DELTA, delta=64)
    ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
    np.testing.assert_array_equal(d, oracle.dijkstra(ooff, oidx, ow, src))


# ---- layouts and tunables change nothing but speed -------------------------------------

@pytest.mark.parametrize("reorder", [0, 1])
def test_results_independent_of_layout(engine, reorder):
    engine.set_option(gmp.OPT_REORDER, reorder)
    try:
        g = gen(engine, 12, weighted=True, num_shards=8)
        V = g["outdeg"].size
        say(f"layout reorder={reorder}")
        r, _ = engine.pagerank(max_iter=20)
        x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=20)
        assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
        seed = int(np.argmax(g["outdeg"]))
        r, _ = engine.pagerank(max_iter=20, seed_vertex=seed)
        x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=20, seed=seed)
        assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
        src = engine.pick_source(77)
        ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
        want = oracle.dijkstra(ooff, oidx, ow, src)
        for method in (gmp.SSSP_BF, gmp.SSSP_DELTA):
            d, _ = engine.sssp(src, method=method, delta=64)
            np.testing.assert_array_equal(d, want)
        h = gen(engine, 12, undirected=True, num_shards=8)
        lab, _ = engine.cc()
        np.testing.assert_array_equal(lab, oracle.cc_unionfind(h["off"], h["idx"]))
        # the compat apps always run on the loaded layout (reference summation order)
        msg0 = np.arange(V, dtype=np.int32)
        msg, _, _ = engine.run_app(gmp.APP_CC, msg0, 1000)
        want, _, _, _ = oracle.app_sync(oracle.APP_CC, h["off"], h["idx"], msg0, 1000)
        np.testing.assert_array_equal(msg, want)
    finally:
        engine.set_option(gmp.OPT_REORDER, 1)


@pytest.mark.parametrize("cap,rows,chunk,wg", [(64, 1, 64, 1), (4096, 4096, 1 << 20, 64), (256, 16, 512, 3)])
def test_plan_tunables_keep_results(engine, cap, rows, chunk, wg):
    g = gen(engine, 13)
    x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=10)
    keys = (gmp.OPT_STREAM_CAP, gmp.OPT_ROW_CAP, gmp.OPT_CHUNK, gmp.OPT_WG_PER_CU, gmp.OPT_VEC_MIN)
    saved = [engine.get_option(k) for k in keys]
    for k, v in zip(keys, (cap, rows, chunk, wg, max(1, cap // 8))):
        engine.set_option(k, v)
    try:
        say(f"plan cap={cap} rows={rows} chunk={chunk} wg={wg}")
        r, _ = engine.pagerank(max_iter=10)
        assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
    finally:
        for k, v in zip(keys, saved):
            engine.set_option(k, v)


@pytest.mark.parametrize("reorder", [0, 1])
def test_compressed_gather_is_bit_identical(engine, reorder):
    """Delta/varint column ids decoded in the gather give exactly the plain row gather's ranks."""
    engine.set_option(gmp.OPT_REORDER, reorder)
    engine.set_option(gmp.OPT_STRIP_LOG, 0)
    try:
        g = gen(engine, 13, off64=bool(reorder))
        say(f"compressed gather reorder={reorder}")
        r0, s0 = engine.pagerank(max_iter=20)
        r1, s1 = engine.pagerank(max_iter=20, compressed=True)
        np.testing.assert_array_equal(r1.view(np.uint32), r0.view(np.uint32))
        assert s1.index_bytes < s0.index_bytes == 4 * g["idx"].size
        # rows in arbitrary order (restarts), hubs spanning many groups, E not a multiple of 64
        V = 30011
        rng = np.random.default_rng(9)
        src = np.concatenate([rng.integers(0, V, 200001), np.arange(V)])
        dst = np.concatenate([rng.integers(0, 50, 200001), rng.integers(0, V, V)])
        off, idx = graphs.csr_from_edges(V, src, dst)
        for r in range(0, V, 7):  # scramble some rows
            seg = idx[off[r]:off[r + 1]]
            idx[off[r]:off[r + 1]] = seg[::-1]
        engine.load_csr(off, idx)
        r0, _ = engine.pagerank(max_iter=12)
        r1, _ = engine.pagerank(max_iter=12, compressed=True)
        np.testing.assert_array_equal(r1.view(np.uint32), r0.view(np.uint32))
        x, _, _ = oracle.pr_f64(off, idx, graphs.out_degree(V, idx), max_iter=12)
        assert np.abs(r1.astype(np.float64) - x).sum() < PR_L1_TOL
    finally:
        engine.set_option(gmp.OPT_REORDER, 1)
        engine.set_option(gmp.OPT_STRIP_LOG, 15)


STRIP_DEFAULTS = ((gmp.OPT_STRIP_LOG, 15), (gmp.OPT_BIN_LOG, 14), (gmp.OPT_BIN_TASK_LOG, 20),
                  (gmp.OPT_STRIP_TASK_LOG, 6))


@pytest.mark.parametrize("strip_log,bin_log,task_log,gtask_log", [(0, 13, 20, 6), (8, 8, 8, 0), (11, 10, 10, 12),
                                                                  (15, 13, 20, 6), (15, 14, 8, 2), (15, 14, 22, 9)])
def test_source_strips_keep_results(engine, strip_log, bin_log, task_log, gtask_log):
    """The source-strip iteration (LDS-staged strips, per-segment messages, fixed-point bin
    sums, split bins) matches FP64 for every strip width, bin size, split size and workgroup
    task size, and is bit-reproducible; strip_log 0 selects the row gather."""
    for k, v in zip((gmp.OPT_STRIP_LOG, gmp.OPT_BIN_LOG, gmp.OPT_BIN_TASK_LOG, gmp.OPT_STRIP_TASK_LOG),
                    (strip_log, bin_log, task_log, gtask_log)):
        engine.set_option(k, v)
    try:
        for scale, kw in ((13, {}), (12, dict(undirected=True)), (11, dict(off64=True))):
            g = gen(engine, scale, **kw)
            say(f"strips {strip_log}/{bin_log}/{task_log} scale={scale} {kw}")
            r, st = engine.pagerank(max_iter=20)
            # the strip layout streams 2-byte source codes, the row gather 4-byte ids
            assert st.index_bytes == (2 if strip_log else 4) * len(g["idx"])
            r2, _ = engine.pagerank(max_iter=20)
            np.testing.assert_array_equal(r.view(np.uint32), r2.view(np.uint32))
            x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=20)
            assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
            assert abs(float(r.astype(np.float64).sum()) - 1.0) < 1e-5
            seed = int(np.argmax(g["outdeg"]))
            r, _ = engine.pagerank(max_iter=20, seed_vertex=seed)
            x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=20, seed=seed)
            assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
        # a hub with 40k in-edges from every strip, dense pairs, many tasks per strip
        V = 40000
        rng = np.random.default_rng(5)
        src = np.concatenate([np.arange(1, V), rng.integers(0, V, 100000), np.arange(2, 9000)])
        dst = np.concatenate([np.zeros(V - 1, np.int64), rng.integers(0, V, 100000), np.ones(8998, np.int64)])
        off, idx = graphs.csr_from_edges(V, src, dst)
        engine.load_csr(off, idx)
        r, st = engine.pagerank(max_iter=25)
        assert st.index_bytes == (2 if strip_log else 4) * len(idx)
        x, _, _ = oracle.pr_f64(off, idx, graphs.out_degree(V, idx), max_iter=25)
        assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
    finally:
        for k, v in STRIP_DEFAULTS:
            engine.set_option(k, v)


@pytest.mark.parametrize("strip_log,task_log,gtask_log", [(8, 8, 0), (11, 10, 12), (15, 20, 6), (15, 14, 2)])
def test_packed_strip_codes_are_bit_identical(engine, strip_log, task_log, gtask_log):
    """Compressed PageRank on the strip layout (source codes bit-packed per 512-edge round)
    gives exactly the plain strip iteration's ranks, with and without a seed and a tolerance,
    on graphs whose code differences need every width: consecutive sources (width 0), repeated
    edges (a difference of -1, width 15), one-edge segments, strips of a few edges.  So does the
    gather over 4-byte source ids (GMP_PR_STRIP_IDS), whose last rounds read ids of the next
    strip or of the padding past a task's end."""
    for k, v in zip((gmp.OPT_STRIP_LOG, gmp.OPT_BIN_TASK_LOG, gmp.OPT_STRIP_TASK_LOG),
                    (strip_log, task_log, gtask_log)):
        engine.set_option(k, v)
    rng = np.random.default_rng(strip_log)
    try:
        cases = [(f"rmat scale={scale} {kw}", (scale, kw)) for scale, kw in
                 ((13, {}), (12, dict(undirected=True)), (11, dict(off64=True)))]
        V = 40000
        src = np.concatenate([np.arange(1, V), rng.integers(0, V, 100000), np.arange(2, 9000),
                              np.repeat(rng.integers(0, V, 300), 7), np.arange(1000, 31000)])
        dst = np.concatenate([np.zeros(V - 1, np.int64), rng.integers(0, V, 100000), np.ones(8998, np.int64),
                              np.repeat(rng.integers(0, V, 300), 7), np.full(30000, 7, np.int64)])
        off, idx = graphs.csr_from_edges(V, src, dst)
        cases.append(("hub, pairs, repeats, a run of 30000 consecutive sources", (off, idx)))
        V = 70001  # few edges per strip, E not a multiple of 8
        src = rng.integers(0, V, 9997)
        dst = rng.integers(0, V, 9997)
        off, idx = graphs.csr_from_edges(V, src, dst)
        cases.append(("sparse", (off, idx)))
        for V, src, dst in ((3, [0, 1, 2, 0], [1, 2, 0, 2]), (2, [0], [1]), (9, [4] * 9, list(range(9)))):
            off, idx = graphs.csr_from_edges(V, np.array(src), np.array(dst))
            cases.append((f"tiny V={V} E={len(src)}", (off, idx)))
        for name, what in cases:
            if isinstance(what[0], int):
                g = gen(engine, what[0], **what[1])
                off, idx = g["off"], g["idx"]
            else:
                off, idx = what
                engine.load_csr(off, idx)
            outdeg = graphs.out_degree(len(off) - 1, idx)
            say(f"packed strips {strip_log}/{task_log}/{gtask_log}: {name}")
            seed = int(np.argmax(outdeg))
            for kw in (dict(max_iter=20), dict(max_iter=20, seed_vertex=seed), dict(max_iter=100, tol=1e-6)):
                r0, s0 = engine.pagerank(**kw)
                r1, s1 = engine.pagerank(compressed=True, **kw)
                np.testing.assert_array_equal(r1.view(np.uint32), r0.view(np.uint32))
                assert s1.iterations == s0.iterations
                assert s0.index_bytes == 2 * len(idx)
                assert 0 < s1.index_bytes
                r2, s2 = engine.pagerank(strip_ids=True, **kw)
                np.testing.assert_array_equal(r2.view(np.uint32), r0.view(np.uint32))
                assert s2.iterations == s0.iterations
                assert s2.index_bytes == 4 * len(idx)
            x, _, _ = oracle.pr_f64(off, idx, outdeg, max_iter=20)
            r1, _ = engine.pagerank(max_iter=20, compressed=True)
            assert np.abs(r1.astype(np.float64) - x).sum() < PR_L1_TOL
    finally:
        for k, v in STRIP_DEFAULTS:
            engine.set_option(k, v)


def test_selective_scheduling_changes_nothing(engine):
    """Schedule independence: with GraphMP's selective scheduling off (every strip in every sweep)
    CC labels, Bellman-Ford and delta-stepping distances (pull rounds on every round), sweep
    counts and changed counts equal the selective runs', which gather fewer edges."""
    keys = (gmp.OPT_SELECTIVE, gmp.OPT_CC_UNION_FIND, gmp.OPT_PUSH_LOG, gmp.OPT_DELTA_PULL_LOG, gmp.OPT_STRIP_LOG)
    saved = [engine.get_option(k) for k in keys]
    try:
        engine.set_option(gmp.OPT_STRIP_LOG, 10)  # 16 strips at scale 14, so there is something to skip
        engine.set_option(gmp.OPT_CC_UNION_FIND, 0)
        engine.set_option(gmp.OPT_PUSH_LOG, 0)  # sweeps only
        engine.set_option(gmp.OPT_DELTA_PULL_LOG, 31)  # every delta-stepping round pulls
        runs = {}
        for sel in (1, 0):
            engine.set_option(gmp.OPT_SELECTIVE, sel)
            g = gen(engine, 14, undirected=True)
            lab, stc = engine.cc()
            g = gen(engine, 14, weighted=True)
            src = engine.pick_source(5)
            d, stb = engine.sssp(src, method=gmp.SSSP_BF)
            dd, std = engine.sssp(src, method=gmp.SSSP_DELTA, delta=64)
            runs[sel] = (lab, stc, d, stb, dd, std, g, src)
        (l1, c1, d1, b1, e1, s1, g, src), (l0, c0, d0, b0, e0, s0, _, _) = runs[1], runs[0]
        np.testing.assert_array_equal(l1, l0)
        np.testing.assert_array_equal(d1, d0)
        np.testing.assert_array_equal(e1, e0)
        assert (c1.iterations, c1.changed) == (c0.iterations, c0.changed)
        assert (b1.iterations, b1.changed) == (b0.iterations, b0.changed)
        assert c1.edges < c0.edges and b1.edges < b0.edges
        V = g["outdeg"].size
        ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
        np.testing.assert_array_equal(d1, oracle.dijkstra(ooff, oidx, ow, src))
    finally:
        for k, v in zip(keys, saved):
            engine.set_option(k, v)


def test_strip_task_size_changes_nothing(engine):
    """Workgroup tasks only decide which workgroup stages a strip and walks which wave tasks: the
    messages, their sums and the ranks, CC labels and distances are the same bits for any size."""
    g = gen(engine, 14, undirected=True)
    out = []
    engine.set_option(gmp.OPT_CC_UNION_FIND, 0)  # CC by strip sweeps
    try:
        for k in (0, 3, 6, 12):
            engine.set_option(gmp.OPT_STRIP_TASK_LOG, k)
            r, _ = engine.pagerank(max_iter=15)
            lab, _ = engine.cc()
            out.append((r, lab))
    finally:
        engine.set_option(gmp.OPT_STRIP_TASK_LOG, 6)
        engine.set_option(gmp.OPT_CC_UNION_FIND, 1)
    for r, lab in out[1:]:
        np.testing.assert_array_equal(r.view(np.uint32), out[0][0].view(np.uint32))
        np.testing.assert_array_equal(lab, out[0][1])
    assert np.abs(out[0][0].astype(np.float64) - oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=15)[0]).sum() < PR_L1_TOL


def test_autotune_keeps_results_and_round_trips(engine, tmp_path):
    keys = (gmp.OPT_STREAM_CAP, gmp.OPT_ROW_CAP, gmp.OPT_CHUNK, gmp.OPT_WG_PER_CU, gmp.OPT_VEC_MIN,
            gmp.OPT_STRIP_LOG, gmp.OPT_BIN_LOG, gmp.OPT_BIN_TASK_LOG, gmp.OPT_PUSH_LOG, gmp.OPT_STRIP_TASK_LOG)
    saved = [engine.get_option(k) for k in keys]
    try:
        g = gen(engine, 14)
        say("autotune")
        res = engine.autotune(iters=2)
        assert res.trials > 1 and res.gather_ms_after <= res.gather_ms_before
        assert [engine.get_option(k) for k in keys] == [res.stream_cap, res.row_cap, res.chunk, res.wg_per_cu,
                                                         res.vec_min, res.strip_log, res.bin_log, res.bin_task_log,
                                                         res.push_log, res.strip_task_log]
        assert res.push_log in (0, 6, 8, 10, 12, 14)
        assert res.strip_log in (0, 13, 14, 15)
        r, _ = engine.pagerank(max_iter=20)
        x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=20)
        assert np.abs(r.astype(np.float64) - x).sum() < PR_L1_TOL
        p = tmp_path / "tune.json"
        engine.tune_save(p)
        for k, v in zip(keys, saved):
            engine.set_option(k, v)
        engine.tune_load(p)
        assert engine.get_option(gmp.OPT_CHUNK) == res.chunk
        assert engine.get_option(gmp.OPT_STRIP_LOG) == res.strip_log
        assert engine.get_option(gmp.OPT_PUSH_LOG) == res.push_log
        assert engine.get_option(gmp.OPT_STRIP_TASK_LOG) == res.strip_task_log
        p.write_text('{"chunk": 3}')
        with pytest.raises(gmp.GmpError) as ei:
            engine.tune_load(p)
        assert ei.value.code == gmp.EFORMAT
    finally:
        for k, v in zip(keys, saved):
            engine.set_option(k, v)


def test_option_range_checks(engine):
    for k, v in ((gmp.OPT_STREAM_CAP, 8192), (gmp.OPT_REORDER, 2), (gmp.OPT_WG_PER_CU, 0), (99, 1),
                 (gmp.OPT_PUSH_LOG, 32), (gmp.OPT_CC_UNION_FIND, 2), (gmp.OPT_WINDOW_MB, 0),
                 (gmp.OPT_WINDOW_PACKED, 2), (gmp.OPT_EDGE_BUDGET_MB, -1)):
        with pytest.raises(gmp.GmpError) as ei:
            engine.set_option(k, v)
        assert ei.value.code == gmp.EINVAL
    assert engine.get_option(gmp.OPT_CHUNK) == 4096
    # defaults of the traversal and semi-external options
    assert [engine.get_option(k) for k in (gmp.OPT_PUSH_LOG, gmp.OPT_CC_UNION_FIND, gmp.OPT_EDGE_BUDGET_MB,
                                           gmp.OPT_WINDOW_MB, gmp.OPT_WINDOW_PACKED)] == [10, 1, 0, 64, 1]


# ---- shard files and edge cases --------------------------------------------------------

def test_load_shards_roundtrip(engine, tmp_path):
    g = gen(engine, 11, weighted=True, num_shards=8)
    V = g["outdeg"].size
    gmp.write_shards(tmp_path, g["off"], g["idx"], g["bounds"], weights=g["w"])
    r1, _ = engine.pagerank(max_iter=10)
    src = engine.pick_source(5)
    d1, _ = engine.sssp(src)
    say("load shards from", tmp_path)
    engine.load_shards(tmp_path, V, 8)
    h = engine.download(weights=True)
    np.testing.assert_array_equal(h["off"], g["off"])
    np.testing.assert_array_equal(h["idx"], g["idx"])
    np.testing.assert_array_equal(h["w"], g["w"])
    np.testing.assert_array_equal(h["bounds"], g["bounds"])
    r2, _ = engine.pagerank(max_iter=10)
    d2, _ = engine.sssp(src)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(d1, d2)


def test_load_shards_missing_file_is_an_error(engine, tmp_path):
    off, idx = graphs.random_graph(100, 300, seed=1)
    gmp.write_shards(tmp_path, off, idx, np.array([0, 50, 100], dtype=np.uint32))
    (tmp_path / "1.edge.npy").unlink()
    with pytest.raises(gmp.GmpError) as ei:
        engine.load_shards(tmp_path, 100, 2)
    assert ei.value.code == gmp.EIO


def _shard_words(path):
    """The int32 words of a shard file: header [edge_num_raw, nnz, len(indptr), start, end],
    indices, indptr (GraphMP/system/Comp.h:37-42)."""
    return np.load(path).astype(np.int64)


def _corrupt_shard(tmp_path, how):
    """A two-shard directory of a 100-vertex graph whose shard 1 is damaged as `how` says."""
    off, idx = graphs.random_graph(100, 300, seed=2)
    gmp.write_shards(tmp_path, off, idx, np.array([0, 50, 100], dtype=np.uint32))
    p = tmp_path / "1.edge.npy"
    w = _shard_words(p)
    nnz = int(w[1])
    if how == "indptr_length":  # one indptr entry dropped: len(indptr) != end - start + 1
        w = np.concatenate([w[:5], w[5:5 + nnz], w[5 + nnz:-1]])
        w[2] -= 1
        np.save(p, w.astype(np.int32))
    elif how == "source_out_of_range":  # a source id equal to V
        w[5] = 100
        np.save(p, w.astype(np.int32))
    elif how == "truncated_data":  # the file ends 40 bytes before its declared shape
        raw = p.read_bytes()
        p.write_bytes(raw[:-40])
    elif how == "truncated_header":  # the file ends inside the .npy header
        raw = p.read_bytes()
        p.write_bytes(raw[:30])
    elif how == "short_shard_header":  # fewer than the 5 shard-header words
        np.save(p, np.array([3, 0, 1], dtype=np.int32))
    elif how == "big_endian":  # cnpy.h:100-101 refuses big-endian data
        np.save(p, w.astype(">i4"))
    elif how == "two_dimensional":  # cnpy.h:90 reads 1-D arrays only
        np.save(p, np.zeros((2, 3), dtype=np.int32))
    elif how == "gap_between_shards":  # shard 1 starts one row after shard 0 ends
        w[3] += 1
        w = np.concatenate([w[:-1]])
        w[2] -= 1
        np.save(p, w.astype(np.int32))
    else:
        raise AssertionError(how)
    return off, idx


@pytest.mark.parametrize("how,code", [("indptr_length", "EFORMAT"), ("source_out_of_range", "EFORMAT"),
                                      ("truncated_data", "EIO"), ("truncated_header", "EFORMAT"),
                                      ("short_shard_header", "EFORMAT"), ("big_endian", "EFORMAT"),
                                      ("two_dimensional", "EFORMAT"), ("gap_between_shards", "EFORMAT")])
def test_load_shards_malformed_is_an_error(engine, tmp_path, how, code):
    """Damaged shards give a status code through the C ABI (GraphMP aborts or asserts:
    cnpy.h:122-133, Comp.h:37-42), and the engine still loads a good graph afterwards."""
    say(f"malformed shard: {how}")
    off, idx = _corrupt_shard(tmp_path, how)
    with pytest.raises(gmp.GmpError) as ei:
        engine.load_shards(tmp_path, 100, 2)
    assert ei.value.code == getattr(gmp, code), (how, ei.value.code, str(ei.value))
    good = tmp_path / "good"
    good.mkdir()
    gmp.write_shards(good, off, idx, np.array([0, 50, 100], dtype=np.uint32))
    engine.load_shards(good, 100, 2)
    g = engine.download()
    np.testing.assert_array_equal(g["off"], off)
    np.testing.assert_array_equal(g["idx"], idx)


def test_edgeless_and_tiny_graphs(engine):
    say("edgeless graph")
    V = 5
    engine.load_csr(np.zeros(V + 1, dtype=np.uint64), np.zeros(0, dtype=np.uint32))
    r, _ = engine.pagerank(max_iter=3)
    np.testing.assert_allclose(r, 1.0 / V, rtol=1e-6)
    d, _ = engine.sssp(2)
    np.testing.assert_array_equal(d, [0xFFFFFFFF, 0xFFFFFFFF, 0, 0xFFFFFFFF, 0xFFFFFFFF])
    lab, _ = engine.cc()
    np.testing.assert_array_equal(lab, np.arange(V))
    say("single edge")
    engine.load_csr(np.array([0, 0, 1], dtype=np.uint64), np.array([0], dtype=np.uint32))
    d, _ = engine.sssp(0)
    np.testing.assert_array_equal(d, [0, 1])


def test_argument_errors(engine):
    engine.load_csr(np.array([0, 1, 1], dtype=np.uint64), np.array([1], dtype=np.uint32))
    with pytest.raises(gmp.GmpError) as ei:
        engine.sssp(7)
    assert ei.value.code == gmp.EINVAL
    with pytest.raises(gmp.GmpError) as ei:
        engine.load_csr(np.array([0, 1, 1], dtype=np.uint64), np.array([9], dtype=np.uint32))
    assert ei.value.code == gmp.EINVAL
    with pytest.raises(gmp.GmpError) as ei:  # 4-byte ids are the uncompressed form
        engine.pagerank(compressed=True, strip_ids=True)
    assert ei.value.code == gmp.EINVAL
    with gmp.Engine(0) as fresh:
        with pytest.raises(gmp.GmpError) as ei:
            fresh.pagerank()
        assert ei.value.code == gmp.ESTATE


def test_load_edge_list_matches_the_converter(engine, tmp_path):
    """gmp_load_edge_list builds on the device the graph gmp_convert_edge_list writes as shards
    (same CSR, weights and partition), and the apps run on it."""
    rng = np.random.default_rng(8)
    V, E = 5000, 60000
    src, dst, w = rng.integers(0, V, E), rng.integers(0, V, E), rng.integers(1, 200, E)
    p = tmp_path / "g.tsv"
    p.write_text("# edges\n" + "".join(f"{a} {b} {c}\n" for a, b, c in zip(src, dst, w)))
    d = tmp_path / "shards"
    d.mkdir()
    assert gmp.convert_edge_list(p, d, V=V, P=4, weighted=True) == V
    engine.load_edge_list(p, V=V, P=4, weighted=True)
    g = engine.download(weights=True)
    engine.load_shards(d, V, 4)
    h = engine.download(weights=True)
    for k in ("off", "idx", "w", "bounds", "outdeg"):
        np.testing.assert_array_equal(g[k], h[k], err_msg=k)
    off, idx = graphs.csr_from_edges(V, src, dst)
    np.testing.assert_array_equal(g["off"], off)
    np.testing.assert_array_equal(g["idx"], idx)
    s = engine.pick_source(3)
    d_, _ = engine.sssp(s, method=gmp.SSSP_BF)
    ooff, oidx, ow = graphs.transpose(V, g["off"], g["idx"], g["w"])
    np.testing.assert_array_equal(d_, oracle.dijkstra(ooff, oidx, ow, s))
    engine.load_edge_list(p, V=V, P=2, symmetrize=True)
    gs = engine.download()
    lab, _ = engine.cc()
    np.testing.assert_array_equal(lab, oracle.cc_unionfind(gs["off"], gs["idx"]))


def _multigraph(V, seed):
    """Destination-row CSR with repeated edges (adjacent copies), self-loops, a source feeding
    every vertex and a destination fed by every vertex (rows above the hub degree, 1024, in the
    push and the gather) and a tail of isolated vertices; sources ascending within a row."""
    rng = np.random.default_rng(seed)
    n = V - 50  # the last 50 vertices have no edges
    src = rng.integers(0, n, 8 * n)
    dst = rng.integers(0, n, 8 * n)
    loops = rng.integers(0, n, n // 4)
    src = np.concatenate([src, loops, np.full(n, 7), np.arange(n), np.arange(0, n, 3)])
    dst = np.concatenate([dst, loops, np.arange(n), np.full(n, 11), np.full((n + 2) // 3, 11)])
    rep = np.where(rng.random(src.size) < 0.2, 2, 1)
    src, dst = np.repeat(src, rep), np.repeat(dst, rep)
    order = np.lexsort((src, dst))
    src, dst = src[order], dst[order]
    off = np.zeros(V + 1, dtype=np.int64)
    np.add.at(off, dst + 1, 1)
    return np.cumsum(off), src.astype(np.uint32)


def test_multigraph_self_loops_and_hubs(engine):
    """Repeated edges count once per copy (out-degrees, PageRank sums) and the lightest copy
    wins (SSSP); self-loops and isolated vertices take the plain rules; every gather and push
    path agrees with the oracle, the compat apps bit for bit."""
    V = 3000
    off, idx = _multigraph(V, 21)
    assert int(np.diff(off)[11]) > 2048 and int((idx == 7).sum()) > 2048
    w = graphs.edge_weights(off, idx, 22)
    od = graphs.out_degree(V, idx)
    say("multigraph pagerank")
    engine.load_csr(off, idx, weights=w, P=3)
    g = engine.download(weights=True)
    np.testing.assert_array_equal(g["outdeg"], od)
    x, _, _ = oracle.pr_f64(off, idx, od, max_iter=30)
    for strip_log in (15, 0):
        engine.set_option(gmp.OPT_STRIP_LOG, strip_log)
        for compressed in (False, True):
            r, st = engine.pagerank(max_iter=30, compressed=compressed)
            assert st.iterations == 30
            assert np.abs(r - x).sum() < PR_L1_TOL, (strip_log, compressed)
    engine.set_option(gmp.OPT_STRIP_LOG, 15)
    say("multigraph sssp")
    ooff, oidx, ow = graphs.transpose(V, off, idx, w)
    for s in (7, 11, 0, V - 1):
        want = oracle.dijkstra(ooff, oidx, ow, s)
        for method in (gmp.SSSP_BF, gmp.SSSP_DELTA):
            d, _ = engine.sssp(s, method=method)
            np.testing.assert_array_equal(d, want, err_msg=f"source {s} method {method}")
        engine.set_option(gmp.OPT_DELTA_PULL_LOG, 31)  # every delta-stepping round pulls
        d, _ = engine.sssp(s, method=gmp.SSSP_DELTA, delta=32)
        engine.set_option(gmp.OPT_DELTA_PULL_LOG, 3)
        np.testing.assert_array_equal(d, want, err_msg=f"source {s} pulled")
        du, _ = engine.sssp(s, unit_weights=True)
        np.testing.assert_array_equal(du, oracle.dijkstra(ooff, oidx, np.ones_like(ow), s))
    say("multigraph compat apps")
    for app in (gmp.APP_PAGERANK, gmp.APP_SSSP, gmp.APP_CC, gmp.APP_SSSP_W):
        vout = None
        if app == gmp.APP_PAGERANK:
            msg0, vout, cap = np.full(V, 1.0 / V, dtype=np.float32), od.astype(np.int32), 40
        elif app == gmp.APP_CC:
            msg0, cap = np.arange(V, dtype=np.int32), 1000
        elif app == gmp.APP_SSSP:
            msg0, cap = np.full(V, 10000, dtype=np.int32), 1000
            msg0[7] = 0
        else:
            msg0, cap = np.full(V, 0xFFFFFFFF, dtype=np.uint32), 1000
            msg0[7] = 0
        msg, val, st = engine.run_app(app, msg0, cap, vertex_out=vout)
        want, wval, it, changed = oracle.app_sync(app, off, idx, msg0, cap, w=w, vout=vout)
        np.testing.assert_array_equal(msg.view(np.uint32), want.view(np.uint32), err_msg=f"app {app}")
        np.testing.assert_array_equal(val.view(np.uint32), wval.view(np.uint32), err_msg=f"app {app}")
        assert st.iterations == it and st.changed == changed[-1]
    say("multigraph cc")
    dst = np.repeat(np.arange(V), np.diff(off))
    s2 = np.concatenate([idx.astype(np.int64), dst])
    d2 = np.concatenate([dst, idx.astype(np.int64)])
    order = np.lexsort((s2, d2))
    soff = np.zeros(V + 1, dtype=np.int64)
    np.add.at(soff, d2 + 1, 1)
    soff, sidx = np.cumsum(soff), s2[order].astype(np.uint32)
    engine.load_csr(soff, sidx)
    lab, _ = engine.cc()
    np.testing.assert_array_equal(lab, oracle.cc_unionfind(soff, sidx))
    assert (lab[-50:] == np.arange(V - 50, V)).all()


def test_long_runs_of_isolated_vertices(engine):
    """Edges only among ids 1000000..1000299 of 3000000 vertices: the out-CSR, the degree-ordered
    layout and the strip layout are built over a million empty rows before the band and two
    million after it (runs longer than one thread fills, left to the gap kernel)."""
    rng = np.random.default_rng(41)
    V, lo, n = 3_000_000, 1_000_000, 300
    src = lo + rng.integers(0, n, 4000)
    dst = lo + rng.integers(0, n, 4000)
    off, idx = graphs.csr_from_edges(V, src, dst)
    w = graphs.edge_weights(off, idx, 42)
    say("isolated-vertex runs: load")
    engine.load_csr(off, idx, weights=w, P=2)
    g = engine.download(weights=True)
    np.testing.assert_array_equal(g["off"], off)
    np.testing.assert_array_equal(g["outdeg"], graphs.out_degree(V, idx))
    say("isolated-vertex runs: sssp")
    ooff, oidx, ow = graphs.transpose(V, off, idx, w)
    s = int(src[0])
    want = oracle.dijkstra(ooff, oidx, ow, s)
    for method in (gmp.SSSP_BF, gmp.SSSP_DELTA):
        d, _ = engine.sssp(s, method=method)
        np.testing.assert_array_equal(d, want)
    say("isolated-vertex runs: pagerank")
    x, _, _ = oracle.pr_f64(off, idx, graphs.out_degree(V, idx), max_iter=10)
    r, _ = engine.pagerank(max_iter=10)
    assert np.abs(r - x).sum() < PR_L1_TOL
    say("isolated-vertex runs: cc")
    soff, sidx = graphs.symmetrize(V, off, idx)
    engine.load_csr(soff, sidx)
    lab, _ = engine.cc()
    np.testing.assert_array_equal(lab, oracle.cc_unionfind(soff, sidx))
