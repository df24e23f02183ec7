"""Semi-external graphs -- edges in pinned host memory, windows of destination rows streamed
through device slots (GraphMP's sliding window, GraphMP/system/GraphPS.h:256-386, with its
compressed edge cache, Global.h:129-255) -- against the resident engine and the oracle.

Every app must give what it gives on the resident graph: PageRank ranks bit for bit equal to
the resident row gather's over the loaded order (each window runs that gather's plan on its
rows, edge for edge and lane for lane), the same distances and labels, and the compat apps bit
for bit, iteration count and changed count included, even where window activation skips
windows."""
import numpy as np
import pytest

import graphee_amd as gmp
import oracle
from tests import graphs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _engine(budget_mb=0, window_mb=1, packed=True):
    e = gmp.Engine(0)
    if budget_mb:
        e.set_option(gmp.OPT_EDGE_BUDGET_MB, budget_mb)
        e.set_option(gmp.OPT_WINDOW_MB, window_mb)
        e.set_option(gmp.OPT_WINDOW_PACKED, int(packed))
    return e


# (edge budget MiB, window MiB, packed): nothing cached / some windows cached, both streams
CONFIGS = [(1, 1, True), (1, 1, False), (10, 1, True), (10, 1, False)]


@pytest.mark.parametrize("budget,window,packed", CONFIGS)
def test_windows_pagerank_matches_resident(budget, window, packed):
    say(f"semi-external pagerank budget={budget} window={window} packed={packed}")
    with _engine() as r, _engine(budget, window, packed) as s:
        for e in (r, s):
            e.generate_rmat(18, 16, 77, num_shards=4)
        # the resident row gather over the loaded order sums every row like the windows do
        r.set_option(gmp.OPT_REORDER, 0)
        r.set_option(gmp.OPT_STRIP_LOG, 0)
        assert r.windows()["windows"] == 0
        w0 = s.windows()
        assert w0["windows"] >= 4 and w0["packed"] == int(packed)
        if budget == 1:
            assert w0["cached"] == 0
        else:
            assert 0 < w0["cached"] < w0["windows"]
        g, h = r.download(), s.download()
        np.testing.assert_array_equal(g["off"], h["off"])
        np.testing.assert_array_equal(g["idx"], h["idx"])
        np.testing.assert_array_equal(g["outdeg"], h["outdeg"])
        x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=20)
        for seed in (-1, int(np.argmax(g["outdeg"])), g["outdeg"].size - 1):
            for kahan in (True, False):
                a, _ = r.pagerank(max_iter=20, seed_vertex=seed, kahan=kahan)
                b, st = s.pagerank(max_iter=20, seed_vertex=seed, kahan=kahan)
                assert st.iterations == 20
                np.testing.assert_array_equal(a, b, err_msg=f"seed {seed} kahan {kahan}")
                if seed < 0 and kahan:
                    assert np.abs(b - x).sum() < 1e-6
        w1 = s.windows()
        uncached = w1["windows"] - w1["cached"]
        assert w1["windows_run"] - w0["windows_run"] == 6 * 20 * w1["windows"]
        assert (w1["streamed_bytes"] > w0["streamed_bytes"]) == (uncached > 0)


def test_windows_pagerank_stops_on_tolerance():
    with _engine() as r, _engine(1, 1, True) as s:
        for e in (r, s):
            e.generate_rmat(17, 16, 5)
        r.set_option(gmp.OPT_REORDER, 0)
        r.set_option(gmp.OPT_STRIP_LOG, 0)
        a, sa = r.pagerank(max_iter=200, tol=1e-7)
        b, sb = s.pagerank(max_iter=200, tol=1e-7)
        assert sa.iterations == sb.iterations < 200
        assert np.abs(a.astype(np.float64) - b).sum() <= 1e-6


@pytest.mark.parametrize("budget,window,packed", CONFIGS)
def test_windows_cc_and_sssp_match_resident(budget, window, packed):
    say(f"semi-external cc/sssp budget={budget} window={window} packed={packed}")
    with _engine() as r, _engine(budget, window, packed) as s:
        for e in (r, s):
            e.generate_rmat(17, 16, 31, undirected=True)
        assert s.windows()["windows"] >= 4
        a, _ = r.cc()
        b, st = s.cc()
        np.testing.assert_array_equal(a, b)
        assert st.iterations >= 2
        for e in (r, s):
            e.generate_rmat(17, 16, 32, weighted=True)
        assert s.windows()["windows"] >= 4
        src = r.pick_source(32)
        assert s.pick_source(32) == src
        g = r.download(weights=True)
        ooff, oidx, ow = graphs.transpose(g["outdeg"].size, g["off"], g["idx"], g["w"])
        want = oracle.dijkstra(ooff, oidx, ow, src)
        for method in (gmp.SSSP_BF, gmp.SSSP_DELTA):
            d, _ = s.sssp(src, method=method)
            np.testing.assert_array_equal(d, want)
        du, _ = r.sssp(src, unit_weights=True)
        dv, _ = s.sssp(src, unit_weights=True)
        np.testing.assert_array_equal(du, dv)


def _local_graph(V, chain):
    """Pairs in the first half of the ids, chains of `chain` vertices (edges i -> i+1) in the
    second, a self-loop on every vertex: the sources of a window are its own rows, so windows
    of the finished first half go idle while the chains still propagate."""
    src, dst = [np.arange(V)], [np.arange(V)]
    h = V // 2
    a = np.arange(0, h, 2)
    src += [a, a + 1]
    dst += [a + 1, a]
    b = np.arange(h, V)
    b = b[(b - h) % chain != chain - 1]
    src.append(b)
    dst.append(b + 1)
    return graphs.csr_from_edges(V, np.concatenate(src), np.concatenate(dst))


@pytest.mark.parametrize("packed", [False, True])
def test_windows_skip_idle_windows_exactly(packed):
    V = 1 << 20
    off, idx = _local_graph(V, 32)
    w = graphs.edge_weights(off, idx, 9)
    with _engine() as r, _engine(1, 1, packed) as s:
        for e in (r, s):
            e.load_csr(off, idx, weights=w)
        w0 = s.windows()
        assert w0["windows"] >= 3
        a, _ = r.cc()
        b, st = s.cc()
        np.testing.assert_array_equal(a, b)
        w1 = s.windows()
        assert w1["windows_skipped"] > w0["windows_skipped"]
        assert w1["windows_run"] - w0["windows_run"] < st.iterations * w0["windows"]
        d0, _ = r.sssp(V // 2 + 3)
        d1, _ = s.sssp(V // 2 + 3)
        np.testing.assert_array_equal(d0, d1)
        assert s.windows()["windows_skipped"] > w1["windows_skipped"]
        # the compat CC app: activation must leave messages, values and the stop step unchanged
        msg0 = np.arange(V, dtype=np.int32)
        m0, v0, s0 = r.run_app(gmp.APP_CC, msg0, 1000)
        w2 = s.windows()
        m1, v1, s1 = s.run_app(gmp.APP_CC, msg0, 1000)
        np.testing.assert_array_equal(m0, m1)
        np.testing.assert_array_equal(v0, v1)
        assert s0.iterations == s1.iterations and s0.changed == s1.changed
        assert s.windows()["windows_skipped"] > w2["windows_skipped"]


def test_windows_selective_off_runs_every_window():
    """With selective scheduling off every window runs in every sweep and step, and the results,
    stop steps and changed counts equal the selective runs'."""
    V = 1 << 20
    off, idx = _local_graph(V, 32)
    w = graphs.edge_weights(off, idx, 9)
    with _engine(1, 1, True) as s:
        s.load_csr(off, idx, weights=w)
        nwin = s.windows()["windows"]
        out = {}
        for sel in (1, 0):
            s.set_option(gmp.OPT_SELECTIVE, sel)
            w0 = s.windows()
            lab, stc = s.cc()
            d, _ = s.sssp(V // 2 + 3)
            m, v, sa = s.run_app(gmp.APP_CC, np.arange(V, dtype=np.int32), 1000)
            w1 = s.windows()
            out[sel] = (lab, stc, d, m, v, sa, w1["windows_skipped"] - w0["windows_skipped"])
        s.set_option(gmp.OPT_SELECTIVE, 1)
    (l1, c1, d1, m1, v1, a1, k1), (l0, c0, d0, m0, v0, a0, k0) = out[1], out[0]
    np.testing.assert_array_equal(l1, l0)
    np.testing.assert_array_equal(d1, d0)
    np.testing.assert_array_equal(m1, m0)
    np.testing.assert_array_equal(v1, v0)
    assert (a1.iterations, a1.changed) == (a0.iterations, a0.changed)
    assert c1.iterations == c0.iterations
    assert k1 > 0 and k0 == 0 and nwin >= 3


@pytest.mark.parametrize("packed", [False, True])
def test_windows_union_find_cc_streams_what_it_needs(tmp_path, packed):
    """CC of a symmetric semi-external graph by sampled union-find: two large components (the
    second one's rows have more than two edges, so the finish phase streams their windows) and
    isolated vertices, against the union-find oracle and the sweeps."""
    rng = np.random.default_rng(8)
    V = 200000
    h = V // 2
    a = np.concatenate([rng.integers(0, h - 1000, 400000), rng.integers(h, V - 1000, 400000)])
    b = np.concatenate([rng.integers(0, h - 1000, 400000), rng.integers(h, V - 1000, 400000)])
    p = tmp_path / "two.tsv"
    p.write_text("".join(f"{x} {y}\n" for x, y in zip(a, b)))
    with _engine(1, 1, packed) as s:
        s.load_edge_list(p, V=V, symmetrize=True)
        w0 = s.windows()
        assert w0["windows"] >= 3
        g = s.download()
        want = oracle.cc_unionfind(g["off"], g["idx"])
        lab, st = s.cc()
        np.testing.assert_array_equal(lab, want)
        assert st.edges > V * 2  # the other component's remaining edges were linked
        lab2, _ = s.cc()  # the first-ids table is kept
        np.testing.assert_array_equal(lab2, want)
        s.set_option(gmp.OPT_CC_UNION_FIND, 0)
        lab3, _ = s.cc()
        np.testing.assert_array_equal(lab3, want)


@pytest.mark.parametrize("app", [gmp.APP_PAGERANK, gmp.APP_SSSP, gmp.APP_CC, gmp.APP_SSSP_W])
@pytest.mark.parametrize("packed", [False, True])
def test_windows_compat_apps_bit_exact(app, packed):
    say(f"semi-external compat app {app} packed={packed}")
    with _engine(1, 1, packed) as s:
        s.generate_rmat(16, 16, 12, weighted=True, num_shards=3)
        assert s.windows()["windows"] >= 2
        g = s.download(weights=True)
        V = g["outdeg"].size
        vout = None
        if app == gmp.APP_PAGERANK:
            msg0 = np.full(V, 1.0 / V, dtype=np.float32)
            vout = g["outdeg"].astype(np.int32)
            cap = 60
        elif app == gmp.APP_CC:
            msg0 = np.arange(V, dtype=np.int32)
            cap = 1000
        elif app == gmp.APP_SSSP:
            msg0 = np.full(V, 10000, dtype=np.int32)  # sssp.cpp's "infinity"
            msg0[s.pick_source(12)] = 0
            cap = 1000
        else:
            msg0 = np.full(V, 0xFFFFFFFF, dtype=np.uint32)
            msg0[s.pick_source(12)] = 0
            cap = 1000
        msg, val, st = s.run_app(app, msg0, cap, vertex_out=vout)
        want, wval, it, changed = oracle.app_sync(app, g["off"], g["idx"], msg0, cap, w=g["w"], vout=vout)
        np.testing.assert_array_equal(msg.view(np.uint32), want.view(np.uint32))
        np.testing.assert_array_equal(val.view(np.uint32), wval.view(np.uint32))
        assert st.iterations == it and st.changed == changed[-1]


def test_windows_load_paths_and_unsupported_runs(tmp_path):
    rng = np.random.default_rng(3)
    V, E = 50000, 600000
    off, idx = graphs.csr_from_edges(V, rng.integers(0, V, E), rng.integers(0, V, E))
    lines = []
    for v in range(V):
        for k in range(off[v], off[v + 1]):
            lines.append(f"{idx[k]} {v}\n")
    p = tmp_path / "g.tsv"
    p.write_text("".join(lines))
    d = tmp_path / "shards"
    d.mkdir()
    gmp.convert_edge_list(p, d, V=V, P=5)
    with _engine(1, 1, True) as s:
        for load in (lambda: s.load_csr(off, idx, P=5), lambda: s.load_edge_list(p, V=V, P=5),
                     lambda: s.load_shards(d, V, 5)):
            load()
            w = s.windows()
            assert w["windows"] >= 2 and w["host_bytes"] > 0
            h = s.download()
            np.testing.assert_array_equal(h["off"], off)
            np.testing.assert_array_equal(h["idx"], idx)
            np.testing.assert_array_equal(h["outdeg"], graphs.out_degree(V, idx))
        with pytest.raises(gmp.GmpError) as ei:
            s.sssp_trace(0)
        assert ei.value.code == gmp.EUNSUPPORTED
        with pytest.raises(gmp.GmpError) as ei:
            s.autotune(1)
        assert ei.value.code == gmp.EUNSUPPORTED
    # within the budget the graph stays resident
    with _engine(64, 1, True) as s:
        s.load_csr(off, idx)
        assert s.windows()["windows"] == 0


@pytest.mark.parametrize("wmax", [255, 100000])
def test_windows_weights_as_bytes_or_words(wmax):
    """Weights that all fit a byte stream and cache as bytes, wider ones as 32-bit words: the
    host bytes say which, the download gives the loaded weights back, and weighted SSSP keeps
    Dijkstra's distances through streamed and cached windows."""
    say(f"semi-external weights up to {wmax}")
    rng = np.random.default_rng(9)
    V, E = 100000, 3000000
    off, idx = graphs.csr_from_edges(V, rng.integers(0, V, E), rng.integers(0, V, E))
    w = rng.integers(1, wmax + 1, idx.size).astype(np.uint32)
    ooff, oidx, ow = graphs.transpose(V, off, idx, w)
    src = int(np.argmax(graphs.out_degree(V, idx)))
    want = oracle.dijkstra(ooff, oidx, ow, src)
    for budget in (1, 8):  # nothing cached / some windows cached
        with _engine(budget, 1, True) as s:
            s.load_csr(off, idx, weights=w, P=3)
            st = s.windows()
            assert st["windows"] >= 3
            per_edge = 1 if wmax <= 255 else 4
            id_bytes = st["host_bytes"] - per_edge * idx.size
            assert 0 < id_bytes <= 4 * idx.size
            h = s.download(weights=True)
            np.testing.assert_array_equal(h["w"], w)
            for method in (gmp.SSSP_BF, gmp.SSSP_DELTA):
                d, _ = s.sssp(src, method=method)
                np.testing.assert_array_equal(d, want)


@pytest.mark.parametrize("packed", [False, True])
def test_windows_learned_prefetch_keeps_results(packed):
    """The window predictor (the engine's second network), trained on traces of this graph's
    sweeps, picks windows whose copies start before the next sweep's activation is known, and
    they run first: labels and distances stay the resident engine's and Dijkstra's, prefetched
    windows are counted (and those the activation then skips, as wasted bytes), and the trace's
    shard masks cover the windows that ran."""
    from graphee_amd import learned
    V = 1 << 22
    off, idx = _local_graph(V, 32)
    w = graphs.edge_weights(off, idx, 9)
    src = V // 3 + 1
    with _engine() as r, _engine(1, 1, packed) as s:
        for e in (r, s):
            e.load_csr(off, idx, weights=w)
        nwin = s.windows()["windows"]
        assert nwin >= 8
        say(f"window trace: {nwin} windows")
        rows = learned.record_windows(s, kinds=(0, 2), sources=(src, 7))
        assert len(rows) >= 4 and np.all(rows["best"] == -1)
        assert rows["shard_mask"].any()
        loss = learned.train_windows(s, rows, epochs=300)
        acc = learned.window_accuracy(s, rows)
        say(f"window predictor: {len(rows)} rows, loss {loss:.3f}, {acc}")
        a_lab, _ = r.cc()
        a_dist, _ = r.sssp(src)
        ooff, oidx, ow = graphs.transpose(V, off, idx, w)
        want = oracle.dijkstra(ooff, oidx, ow, src)
        np.testing.assert_array_equal(a_dist, want)
        for prefetch in (0, 1):
            s.set_option(gmp.OPT_WINDOW_PREFETCH, prefetch)
            w0 = s.windows()
            b_lab, st1 = s.cc()
            b_dist, st2 = s.sssp(src)
            w1 = s.windows()
            np.testing.assert_array_equal(a_lab, b_lab)
            np.testing.assert_array_equal(want, b_dist)
            say(f"prefetch={prefetch}: prefetched {w1['prefetched'] - w0['prefetched']}, wasted "
                f"{w1['prefetch_wasted_bytes'] - w0['prefetch_wasted_bytes']} B, streamed "
                f"{w1['streamed_bytes'] - w0['streamed_bytes']} B, {st1.ms_total:.2f} + {st2.ms_total:.2f} ms")
            if prefetch:
                assert w1["prefetched"] > w0["prefetched"] and st2.policy == 1
            else:
                assert w1["prefetched"] == w0["prefetched"] and st2.policy == 0
        s.set_option(gmp.OPT_WINDOW_PREFETCH, 0)
