#!/usr/bin/env python3
"""Benchmark of the gfx950 graph engine on directed R-MAT scale 26, edge factor 16.

One step = one PageRank solve (FP32 ranks, Kahan gather, FP64 residual/dangling
reductions) until the L1 change of an iteration drops below 1e-7, at most 100
iterations.  The graph (~1.07e9 samples, deduplicated, seeded relabel) is generated on the
GPU and stays in HBM; the timed region starts with every input resident.

value        whole-job PageRank GTEPS = edges * iterations / wall time of the K solves
roofline     one PageRank iteration, i.e. the launch sequence pr_strip_gather_kernel ->
             pr_bin_kernel -> pr_split_apply_kernel on the engine stream: algorithmic bytes
             (8 E + (o + 16) V, o = offset width) / mean iteration time from HIP events
             recorded on that stream around the three launches, against 8 TB/s; traffic =
             measured HBM bytes of the three kernels per iteration from the rocprofv3
             FETCH_SIZE/WRITE_SIZE summary in profiles/ when present
cpu_baseline GraphMP's OpenMP engine shape (oracle restatement, edges in RAM) on the same
             graph, timed over a bounded number of iterations
extra        SSSP (frontier Bellman-Ford and delta-stepping, delta=64) on the same graph with
             uint32 weights in [1,255], CC on the symmetrised graph, and the FP64 a-posteriori
             error bound of the PageRank result.
Only one GPU is used; --gpus N > 1 exits.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


_T0 = time.perf_counter()


def log(*a):
    print(f"[bench {time.perf_counter() - _T0:6.1f}s]", *a, file=sys.stderr, flush=True)


def cpu_threads():
    """Threads for the CPU baseline: every CPU in this process's affinity set, capped by
    OMP_NUM_THREADS when the environment sets it (a GPU host that grants each GPU job a share
    of its cores says so there)."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(int(env), aff) if env.isdigit() and int(env) > 0 else aff


def host_info():
    """What the CPU baseline ran on: CPU model (lscpu's "Model name", read from /proc/cpuinfo),
    logical CPUs of the host (nproc --all), CPUs this process may use, OMP_NUM_THREADS."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"), "threads_used": cpu_threads()}


def load_traffic(scale, ef):
    """HBM bytes per PageRank iteration from the committed rocprofv3 PMC summary, if any."""
    p = os.path.join(ROOT, "profiles", f"pr_strip_rmat_s{scale}_ef{ef}_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def copy_calibration(nbytes):
    """One hipMemcpy device-to-device of nbytes (reads nbytes, writes nbytes)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes >> 20} MiB device-to-device")


def batched_workload(seed=5000):
    """32 R-MAT graphs with scale uniform in 14..22 and edge factor uniform in 8..32 (weights
    1..255), 256 SSSP sources and 256 PPR seeds placed on graphs with Zipf-skewed odds."""
    rng = np.random.default_rng(seed)
    specs = [dict(scale=int(rng.integers(14, 23)), edge_factor=int(rng.integers(8, 33)), seed=seed + g)
             for g in range(32)]
    odds = 1.0 / np.arange(1, 33) ** 1.1
    odds /= odds.sum()
    gs = rng.choice(32, size=256, p=odds).astype(np.uint32)
    gp = rng.choice(32, size=256, p=odds[::-1]).astype(np.uint32)
    return specs, gs, gp, rng


def run_learned(eng, eval_scale, nsources=8, train_scales=(22, 23, 24, 25, 26), trace_sources=3):
    """Learned delta-stepping on R-MAT eval_scale ef16 (weights 1..255) for nsources seeded
    sources: traces from SSSP runs on R-MAT train_scales, the MLP trained on the GPU, then the
    learned schedule -- the eight bucket widths 8..1024 as classes -- against fixed delta = 64 at
    identical distances; beside it every fixed width, plain Bellman-Ford, and the variant whose
    class 7 finishes by Bellman-Ford (its own traces and network)."""
    import graphee_amd as gmp
    from graphee_amd import learned
    out = {"eval_scale": eval_scale, "sources": nsources, "train_scales": list(train_scales),
           "action_space": "bucket widths 8, 16, ..., 1024"}
    nets = {}
    for variant, finish in (("widths", False), ("finish", True)):
        t0 = time.perf_counter()
        rows = learned.record(eng, [(s, 16, 1000 + s) for s in train_scales], sources_per_graph=trace_sources,
                              seed=27, finish=finish)
        t_rec = time.perf_counter() - t0
        t0 = time.perf_counter()
        loss = learned.train(eng, rows, epochs=400)
        t_train = time.perf_counter() - t0
        acc = learned.accuracy(eng, rows)
        nets[variant] = eng.mlp_get()
        labels = learned.labels(rows)
        out[f"train_{variant}"] = {
            "trace_rows": int(len(rows)), "trace_seconds": round(t_rec, 1), "train_seconds": round(t_train, 2),
            "train_loss": round(loss, 4), "label_histogram": np.bincount(labels, minlength=8).tolist(),
            "class_accuracy": round(acc["class_accuracy"], 3), "shard_precision": round(acc["shard_precision"], 3),
            "shard_recall": round(acc["shard_recall"], 3)}
        log(f"learned ({variant}): {len(rows)} trace rows in {t_rec:.1f} s, trained in {t_train:.2f} s, "
            f"loss {loss:.3f}, labels {np.bincount(labels, minlength=8).tolist()}, {acc}")
    eng.generate_rmat(eval_scale, 16, 1000 + eval_scale, weighted=True, num_shards=16)
    V, E, _, _ = eng.info()
    out.update(vertices=V, edges=E)
    rng = np.random.default_rng(2700 + eval_scale)
    srcs = [eng.pick_source(int(rng.integers(1 << 30))) for _ in range(nsources)]
    eng.sssp(srcs[0], method=gmp.SSSP_DELTA, delta=64, want_dist=False)  # builds the layouts
    eng.sssp(srcs[0], method=gmp.SSSP_BF, want_dist=False)

    def learned_run(variant, s, want):
        eng.mlp_set(nets[variant])
        eng.set_option(gmp.OPT_LEARNED_FINISH, int(variant == "finish"))
        try:
            return eng.sssp(s, method=gmp.SSSP_LEARNED, want_dist=want)
        finally:
            eng.set_option(gmp.OPT_LEARNED_FINISH, 0)

    ms = {k: 0.0 for k in ("fixed_delta64", "learned", "learned_finish", "bf")}
    wall = dict(ms)
    fixed = {8 << c: 0.0 for c in range(8)}
    hist = {"learned": np.zeros(8, np.int64), "learned_finish": np.zeros(8, np.int64)}
    same = True
    for s in srcs:
        d0, _ = eng.sssp(s, method=gmp.SSSP_DELTA, delta=64)
        for variant, key in (("widths", "learned"), ("finish", "learned_finish")):
            d1, _ = learned_run(variant, s, True)
            same = same and bool(np.array_equal(d0, d1))
        db, _ = eng.sssp(s, method=gmp.SSSP_BF)
        same = same and bool(np.array_equal(d0, db))
        del d0, d1, db
        # timed solves (distances stay on the device)
        for key, fn in (("fixed_delta64", lambda: eng.sssp(s, method=gmp.SSSP_DELTA, delta=64, want_dist=False)),
                        ("learned", lambda: learned_run("widths", s, False)),
                        ("learned_finish", lambda: learned_run("finish", s, False)),
                        ("bf", lambda: eng.sssp(s, method=gmp.SSSP_BF, want_dist=False))):
            t0 = time.perf_counter()
            _, st = fn()
            wall[key] += (time.perf_counter() - t0) * 1e3
            ms[key] += st.ms_total
            if key in hist:
                hist[key] += eng.sssp_actions()
        for dl in fixed:
            _, st = eng.sssp(s, method=gmp.SSSP_DELTA, delta=dl, want_dist=False)
            fixed[dl] += st.ms_total
    best = min(fixed, key=fixed.get)
    out.update({
        "fixed_delta64_ms": round(ms["fixed_delta64"], 3), "learned_ms": round(ms["learned"], 3),
        "speedup": round(ms["fixed_delta64"] / ms["learned"], 3),
        "learned_finish_ms": round(ms["learned_finish"], 3),
        "learned_finish_speedup": round(ms["fixed_delta64"] / ms["learned_finish"], 3),
        "bf_ms": round(ms["bf"], 3), "best_fixed_delta": best, "best_fixed_delta_ms": round(fixed[best], 3),
        "fixed_delta_ms": {str(k): round(v, 3) for k, v in fixed.items()},
        "wall_ms": {k: round(v, 3) for k, v in wall.items()},
        "action_histogram": hist["learned"].tolist(), "action_histogram_finish": hist["learned_finish"].tolist(),
        "identical_distances": same})
    log(f"learned s{eval_scale} ({nsources} sources): fixed-64 {ms['fixed_delta64']:.1f} ms, learned widths "
        f"{ms['learned']:.1f} ms ({out['speedup']:.2f}x, actions {hist['learned'].tolist()}), learned+finish "
        f"{ms['learned_finish']:.1f} ms, BF {ms['bf']:.1f} ms, best fixed delta {best}: {fixed[best]:.1f} ms; "
        f"identical distances: {same}")
    return out


def run_pagerank_s20(eng, threads):
    """PageRank on directed R-MAT s20 ef16 in 16 destination-interval shards, exactly 20
    iterations from 1/N, FP32 ranks: GraphMP's CPU engine shape (the oracle's OpenMP
    restatement, timed) beside the GPU, each compared with FP64 -- the same 20 iterations in
    FP64 (arithmetic error) and the FP64 fixed point (arithmetic + truncation) -- plus the
    reference's own update rule (GraphMP/app/pagerank.cpp:42-59, run by gmp_run_app) bit for bit
    against the oracle's restatement."""
    import graphee_amd as gmp
    import oracle
    scale = 20
    eng.generate_rmat(scale, 16, 1000 + scale, num_shards=16)
    V, E, P, _ = eng.info()
    g = eng.download()
    msg0 = np.full(V, 1.0 / V, dtype=np.float32)
    x_cpu, stc = oracle.graphmp_baseline(oracle.APP_PAGERANK, g["off"], g["idx"], g["bounds"], msg0, 20,
                                         outdeg=g["outdeg"], threads=threads)
    x20, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], max_iter=20)
    eng.pagerank(damping=0.85, max_iter=20, want_ranks=False)
    x_gpu, stg = eng.pagerank(damping=0.85, max_iter=20)
    out = {"workload": f"pagerank-rmat-s{scale}-ef16, {P} shards, 20 iterations, FP32 ranks",
           "vertices": V, "edges": E, "shards": P}
    for name, x, secs, it in (("cpu", x_cpu, stc.seconds, stc.iterations),
                              ("gpu", x_gpu, stg.ms_total / 1e3, stg.iterations)):
        d20 = np.abs(x.astype(np.float64) - x20)
        fp = oracle.pr_f64_error(g["off"], g["idx"], g["outdeg"], x, threads=threads, max_iter=200, bound=1e-13)
        out[name] = {"seconds": round(secs, 5), "iterations": it, "gteps": round(E * it / secs / 1e9, 3),
                     "l1_vs_fp64_same_20_iterations": float(d20.sum()),
                     "linf_vs_fp64_same_20_iterations": float(d20.max()),
                     "l1_vs_fp64_fixed_point": fp["l1_error"], "linf_vs_fp64_fixed_point": fp["linf_error"],
                     "fixed_point_reference_bound": fp["reference_l1_bound"]}
    out["cpu"]["threads"] = stc.threads
    out["gpu_speedup"] = round(stc.seconds / (stg.ms_total / 1e3), 2)
    # the reference's own rule: sum in source order, (float)(rel*0.85 + 1.0/N), dangling mass dropped
    vout = g["outdeg"].astype(np.int32)
    m_gpu, _, _ = eng.run_app(gmp.APP_PAGERANK, msg0, 20, vertex_out=vout)
    m_orc, _, _, _ = oracle.app_sync(oracle.APP_PAGERANK, g["off"], g["idx"], msg0, 20, vout=vout)
    out["graphmp_rule_bit_exact"] = bool(np.array_equal(m_gpu.view(np.uint32), m_orc.view(np.uint32)))
    log(f"pagerank s{scale} x20: CPU {stc.seconds * 1e3:.1f} ms on {stc.threads} threads "
        f"(L1/Linf vs FP64 {out['cpu']['l1_vs_fp64_same_20_iterations']:.2e}/"
        f"{out['cpu']['linf_vs_fp64_same_20_iterations']:.2e}), GPU {stg.ms_total:.2f} ms "
        f"({out['gpu']['l1_vs_fp64_same_20_iterations']:.2e}/{out['gpu']['linf_vs_fp64_same_20_iterations']:.2e}); "
        f"GraphMP rule bit-exact: {out['graphmp_rule_bit_exact']}")
    return out


def run_big(scale=28, ef=17, verify=True, threads=16):
    """PageRank on R-MAT s28 ef17 (4.5e9 edges, past 2^32) resident in HBM: the strip layout
    (edge offsets counted from per-strip bases) against the row gather on the same graph, and
    the FP64 one-step bound of the converged strip ranks."""
    import graphee_amd as gmp
    e = gmp.Engine(0)
    try:
        t0 = time.perf_counter()
        e.generate_rmat(scale, ef, 7000 + scale, num_shards=16)
        V, E, _, _ = e.info()
        gen_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        e.pagerank(max_iter=3, want_ranks=False)  # builds the strip layout
        build_s = time.perf_counter() - t0
        r1, st1 = e.pagerank(max_iter=100, tol=1e-7)
        ms1 = st1.ms_kernel / st1.launches
        e.set_option(gmp.OPT_STRIP_LOG, 0)
        e.pagerank(max_iter=3, want_ranks=False)
        r0, st0 = e.pagerank(max_iter=st1.iterations, tol=0.0)
        ms0 = st0.ms_kernel / st0.launches
        out = {"workload": f"pagerank-rmat-s{scale}-ef{ef}-tol1e-07", "vertices": V, "edges": E,
               "edges_over_2e32": round(E / 2**32, 3), "generated_s": round(gen_s, 1),
               "strip_layout_build_s": round(build_s, 1), "iterations": st1.iterations,
               "strip_ms_per_iteration": round(ms1, 3), "strip_gteps": round(E / ms1 / 1e6, 1),
               "strip_index_bytes_per_edge": round(st1.index_bytes / E, 3),
               "row_gather_ms_per_iteration": round(ms0, 3), "strip_speedup": round(ms0 / ms1, 2),
               "l1_strip_vs_row_gather": float(np.abs(r1.astype(np.float64) - r0).sum())}
        if verify:
            import oracle
            g = e.download()
            l1, _ = oracle.pr_step_residual(g["off"], g["idx"], g["outdeg"], r1, threads=threads)
            out["fp64_step_residual_l1"] = l1
            out["fp64_l1_error_bound"] = l1 / 0.15
            del g
        log(f"pagerank s{scale} ef{ef} ({E} edges = {E / 2**32:.2f} x 2^32): strips {ms1:.2f} ms/iteration "
            f"({E / ms1 / 1e6:.0f} GTEPS), row gather {ms0:.2f} ms, L1 between them {out['l1_strip_vs_row_gather']:.2e}"
            + (f", FP64 L1 bound {out['fp64_l1_error_bound']:.2e}" if verify else ""))
        return out
    finally:
        e.close()


def run_sssp_s24(eng, threads):
    """SSSP on directed R-MAT s24 ef16 with uint32 weights in [1,255] from a seeded source of
    out-degree > 0: Bellman-Ford (pushes + strip sweeps, GraphMP's selective scheduling) and
    fixed delta = 64 on the GPU, and the weighted GraphMP rule in the CPU engine restatement;
    the three distance vectors must be identical."""
    import graphee_amd as gmp
    import oracle
    scale = 24
    eng.generate_rmat(scale, 16, 1000 + scale, weighted=True, num_shards=16)
    V, E, _, _ = eng.info()
    src = eng.pick_source(1000 + scale)
    out = {"workload": f"sssp-rmat-s{scale}-ef16, weights 1..255", "vertices": V, "edges": E, "source": src}
    dist = {}
    for name, method in (("bf", gmp.SSSP_BF), ("delta64", gmp.SSSP_DELTA)):
        eng.sssp(src, method=method, delta=64, want_dist=False)
        d, st = eng.sssp(src, method=method, delta=64)
        dist[name] = d
        out[name] = {"ms": round(st.ms_total, 3), "iterations": st.iterations, "edges_examined": st.edges,
                     "work_gteps": round(st.edges / st.ms_total / 1e6, 3), "reached": int((d != 0xFFFFFFFF).sum()),
                     "yardstick_gbs": round(st.bytes / st.ms_total / 1e6, 1),
                     "hbm_frac": round(st.bytes / st.ms_total / 1e6 / HBM_PEAK_GBS, 4)}
    g = eng.download(weights=True)
    # Graph500-style GTEPS counts the edges of the reached vertices
    reach = dist["bf"] != 0xFFFFFFFF
    e_reached = int(np.diff(g["off"].astype(np.int64))[reach].sum())
    for name in ("bf", "delta64"):
        out[name]["graph500_gteps"] = round(e_reached / out[name]["ms"] / 1e6, 3)
    msg0 = np.full(V, 0xFFFFFFFF, dtype=np.uint32)
    msg0[src] = 0
    dc, stc = oracle.graphmp_baseline(oracle.APP_SSSP_W, g["off"], g["idx"], g["bounds"], msg0, 10000, w=g["w"],
                                      threads=threads)
    out["cpu"] = {"seconds": round(stc.seconds, 3), "threads": stc.threads, "sweeps": stc.iterations,
                  "shards_computed": int(stc.shards)}
    # the same engine with GraphMP's per-shard bloom filters choosing the shards of sparse steps
    # (GraphPS.h:349-368): the reference's own activation, false positives included
    db, stb = oracle.graphmp_baseline(oracle.APP_SSSP_W, g["off"], g["idx"], g["bounds"], msg0, 10000, w=g["w"],
                                      threads=threads, bloom=True)
    out["cpu_bloom"] = {"seconds": round(stb.seconds, 3), "sweeps": stb.iterations, "shards_computed": int(stb.shards),
                        "identical": bool(np.array_equal(db, dc))}
    out["identical"] = bool(np.array_equal(dc, dist["bf"]) and np.array_equal(dc, dist["delta64"]))
    out["gpu_bf_speedup"] = round(stc.seconds * 1e3 / out["bf"]["ms"], 2)
    out["gpu_delta64_speedup"] = round(stc.seconds * 1e3 / out["delta64"]["ms"], 2)
    log(f"sssp s{scale}: BF {out['bf']['ms']:.1f} ms, delta-64 {out['delta64']['ms']:.1f} ms, CPU "
        f"{stc.seconds:.2f} s ({stc.threads} threads, {int(stc.shards)} shard computations; with GraphMP's bloom "
        f"filters {stb.seconds:.2f} s, {int(stb.shards)}); identical: {out['identical'] and out['cpu_bloom']['identical']}")
    return out


def gmp_const(name):
    import graphee_amd
    return getattr(graphee_amd, name)


def verify_batched(eng, specs, gs, srcs, dists, gp, seeds, ranks, threads):
    """One SSSP and one personalised-PageRank instance per distinct graph scale (the first
    instance on a graph of that scale) checked against the oracle: distances bit for bit against
    Dijkstra, ranks within 1e-6 L1 of FP64 personalised PageRank (30 iterations, damping 0.85)."""
    import oracle
    from oracle import rmat_ref
    out = {"sssp_checked": 0, "sssp_identical": 0, "ppr_checked": 0, "ppr_max_l1": 0.0, "scales": []}
    cache = {}

    def graph(g):
        if g not in cache:
            cache.clear()
            off, idx, w = eng.batch_download(g)
            cache[g] = (off, idx, w)
        return cache[g]

    for scale in sorted({s["scale"] for s in specs}):
        out["scales"].append(scale)
        on = [i for i, g in enumerate(gs) if specs[g]["scale"] == scale]
        if on:
            i = on[0]
            off, idx, w = graph(int(gs[i]))
            V = off.size - 1
            ooff, oidx, ow = rmat_ref.transpose(V, off, idx, w)
            want = oracle.dijkstra(ooff, oidx, ow, int(srcs[i]))
            out["sssp_checked"] += 1
            out["sssp_identical"] += int(np.array_equal(dists[i], want))
            del ooff, oidx, ow, want
        on = [i for i, g in enumerate(gp) if specs[g]["scale"] == scale]
        if on:
            i = on[0]
            off, idx, w = graph(int(gp[i]))
            V = off.size - 1
            od = np.bincount(idx, minlength=V).astype(np.uint32)
            x, _, _ = oracle.pr_f64(off, idx, od, d=0.85, max_iter=30, seed=int(seeds[i]), threads=threads)
            l1 = float(np.abs(ranks[i].astype(np.float64) - x).sum())
            out["ppr_checked"] += 1
            out["ppr_max_l1"] = max(out["ppr_max_l1"], l1)
    out["verified"] = bool(out["sssp_identical"] == out["sssp_checked"] and out["ppr_max_l1"] <= 1e-6)
    return out


def run_batched(eng, verify=True, tune=True):
    """32 R-MAT graphs of scale 14..22 and edge factor 8..32 (weights 1..255) resident at once;
    256 SSSP sources and 256 personalised-PageRank seeds (30 iterations): the batch autotuner's
    pass over the batched kernels' knobs, each kind alone, and both kinds in one replayed graph
    (gmp_batch_run, whose outputs are the ones verified)."""
    specs, gs, gp, rng = batched_workload()
    t0 = time.perf_counter()
    eng.batch_generate(specs)
    gen_s = time.perf_counter() - t0
    sizes = [eng.batch_info(g) for g in range(32)]
    srcs = np.array([rng.integers(0, sizes[g][0]) for g in gs], dtype=np.uint32)
    seeds = np.array([rng.integers(0, sizes[g][0]) for g in gp], dtype=np.uint32)
    res = {"graphs": 32, "total_edges": int(sum(s[1] for s in sizes))}
    if tune:
        t0 = time.perf_counter()
        res["tuned"] = eng.batch_autotune(gs, srcs, gp, seeds, iters=30)
        res["tuned"]["seconds"] = round(time.perf_counter() - t0, 1)
        log(f"batch autotune: {res['tuned']}")
    eng.batch_sssp(gs[:4], srcs[:4])  # warm the kernels
    _, st = eng.batch_sssp(gs, srcs)
    res["sssp"] = {"instances": 256, "ms": round(st.ms_total, 3), "rounds": st.iterations,
                   "graph_launches": st.launches, "edges_relaxed": st.edges,
                   "work_gteps": round(st.edges / st.ms_total / 1e6, 3)}
    _, st2 = eng.batch_ppr(gp, seeds, damping=0.85, iters=30)
    res["ppr"] = {"instances": 256, "iterations": 30, "ms": round(st2.ms_total, 3),
                  "instance_edges": st2.edges, "gteps": round(st2.edges / st2.ms_total / 1e6, 3)}
    dists, ranks, st3, st4 = eng.batch_run(gs, srcs, gp, seeds, damping=0.85, iters=30)
    res["combined"] = {"ms": round(st3.ms_total, 3), "graph_launches": st3.launches,
                       "separate_ms": round(st.ms_total + st2.ms_total, 3),
                       "speedup_over_separate": round((st.ms_total + st2.ms_total) / st3.ms_total, 3)}
    log(f"batched: {res['total_edges']} edges in 32 graphs (built in {gen_s:.1f} s); SSSP x256 {st.ms_total:.1f} ms, "
        f"PPR x256 {st2.ms_total:.1f} ms, both in one graph {st3.ms_total:.1f} ms")
    if verify:
        t0 = time.perf_counter()
        res["verification"] = verify_batched(eng, specs, gs, srcs, dists, gp, seeds, ranks, cpu_threads())
        res["verification"]["seconds"] = round(time.perf_counter() - t0, 1)
        res["verified"] = res["verification"]["verified"]
        log(f"batched verification: {res['verification']}")
    return res


PCIE5_X16_GBS = 64.0  # MI355X host link, PCIe 5.0 x16: published 128 GB/s both ways, 64 per direction


def pinned_h2d_gbs(nbytes=1 << 30, chunk=64 << 20, reps=4):
    """Host-to-device copy rate from pinned memory: 64 MiB hipMemcpyAsync chunks on two
    non-blocking streams (the shape of the semi-external ring), timed with HIP events."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    h, d = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipHostMalloc(ctypes.byref(h), ctypes.c_size_t(nbytes), ctypes.c_uint(0)) != 0:
        raise RuntimeError("hipHostMalloc failed")
    if hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMalloc failed")
    ctypes.memset(h, 1, nbytes)
    st = [ctypes.c_void_p(), ctypes.c_void_p()]
    for x in st:
        hip.hipStreamCreateWithFlags(ctypes.byref(x), ctypes.c_uint(1))  # hipStreamNonBlocking
    e0, e1, e2 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    for x in (e0, e1, e2):
        hip.hipEventCreate(ctypes.byref(x))

    def one_pass():
        for k in range(nbytes // chunk):
            if hip.hipMemcpyAsync(ctypes.c_void_p(d.value + k * chunk), ctypes.c_void_p(h.value + k * chunk),
                                  ctypes.c_size_t(chunk), 1, st[k % 2]) != 0:  # hipMemcpyHostToDevice
                raise RuntimeError("hipMemcpyAsync failed")

    one_pass()
    hip.hipDeviceSynchronize()
    hip.hipEventRecord(e0, st[0])
    hip.hipStreamWaitEvent(st[1], e0, ctypes.c_uint(0))
    for _ in range(reps):
        one_pass()
    hip.hipEventRecord(e1, st[1])
    hip.hipStreamWaitEvent(st[0], e1, ctypes.c_uint(0))
    hip.hipEventRecord(e2, st[0])
    hip.hipEventSynchronize(e2)
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e2)
    for x in (e0, e1, e2):
        hip.hipEventDestroy(x)
    for x in st:
        hip.hipStreamDestroy(x)
    hip.hipFree(d)
    hip.hipHostFree(h)
    return nbytes * reps / ms.value / 1e6


def run_semi(scale, ef, seed, ref, iters=20):
    """PageRank with the edges in pinned host memory (semi-external windows), streamed packed or
    plain, with and without a device cache; ranks compared with the resident engine's."""
    import graphee_amd as gmp
    out = {"workload": f"pagerank-rmat-s{scale}-ef{ef}, {iters} iterations, edges in pinned host memory",
           "pinned_h2d_gbs": round(pinned_h2d_gbs(), 2), "pcie_peak_gbs": PCIE5_X16_GBS, "runs": {}}
    log(f"pinned host-to-device copy: {out['pinned_h2d_gbs']:.1f} GB/s")
    for name, packed, budget in (("packed_streamed", 1, 512), ("plain_streamed", 0, 512),
                                 ("packed_cache_1gib", 1, 1024 + 512)):
        with gmp.Engine(0) as s:
            s.set_option(gmp.OPT_EDGE_BUDGET_MB, budget)
            s.set_option(gmp.OPT_WINDOW_MB, 64)
            s.set_option(gmp.OPT_WINDOW_PACKED, packed)
            t0 = time.perf_counter()
            s.generate_rmat(scale, ef, seed, num_shards=16)
            setup_s = time.perf_counter() - t0
            E = s.info()[1]
            s.pagerank(damping=0.85, max_iter=2, want_ranks=False)
            runs = []
            for _ in range(2):  # two timed solves; the faster one is reported, both are listed
                w0 = s.windows()
                x, st = s.pagerank(damping=0.85, max_iter=iters)
                w1 = s.windows()
                runs.append((st.ms_total, st, w1["streamed_bytes"] - w0["streamed_bytes"], x))
            runs.sort(key=lambda t: t[0])
            _, st, moved, x = runs[0]
            r = {"windows": w1["windows"], "cached": w1["cached"], "host_bytes": w1["host_bytes"],
                 "device_edge_bytes": w1["device_bytes"], "id_bytes_per_iteration": st.index_bytes,
                 "ms_per_iteration": round(st.ms_total / iters, 3),
                 "ms_per_iteration_runs": sorted(round(t[0] / iters, 3) for t in runs),
                 "gteps": round(E * iters / st.ms_total / 1e6, 3),
                 "streamed_gb_per_iteration": round(moved / iters / 1e9, 3),
                 "h2d_gbs": round(moved / st.ms_total / 1e6, 2),
                 "l1_vs_resident": float(np.abs(x.astype(np.float64) - ref).sum()),
                 "setup_s": round(setup_s, 1)}
            r["h2d_frac_of_peak"] = round(r["h2d_gbs"] / PCIE5_X16_GBS, 3)
            r["h2d_frac_of_pinned_copy"] = round(r["h2d_gbs"] / out["pinned_h2d_gbs"], 3)
            out["runs"][name] = r
            log(f"semi-external {name}: {r['windows']} windows ({r['cached']} cached), "
                f"{r['ms_per_iteration']:.1f} ms/iteration, {r['gteps']:.1f} GTEPS, {r['h2d_gbs']:.1f} GB/s "
                f"host-to-device ({r['h2d_frac_of_peak']:.2f} of PCIe 5.0 x16, {r['h2d_frac_of_pinned_copy']:.2f} of a "
                f"pinned copy), L1 vs resident {r['l1_vs_resident']:.2e}")
            del x, runs
    out["sssp_prefetch"] = run_semi_sssp(scale, ef, seed)
    return out


def run_semi_sssp(scale, ef, seed, budget_mb=512):
    """Weighted SSSP over the windows of a semi-external R-MAT graph (ids packed, weights as
    bytes, both streamed): exact window activation against the learned prefetch, whose window predictor
    is trained on traces of two other sources' solves on the same graph; distances must agree."""
    import graphee_amd as gmp
    from graphee_amd import learned
    out = {"workload": f"sssp-rmat-s{scale}-ef{ef}, weights 1..255, edges in pinned host memory, "
                       f"{budget_mb} MiB device budget"}
    with gmp.Engine(0) as s:
        s.set_option(gmp.OPT_EDGE_BUDGET_MB, budget_mb)
        s.set_option(gmp.OPT_WINDOW_MB, 64)
        s.set_option(gmp.OPT_WINDOW_PACKED, 1)
        s.generate_rmat(scale, ef, seed, weighted=True, num_shards=16)
        rng = np.random.default_rng(seed + 7)
        train_src = [s.pick_source(int(rng.integers(1 << 30))) for _ in range(2)]
        src = s.pick_source(seed)
        t0 = time.perf_counter()
        rows = learned.record_windows(s, kinds=(2,), sources=train_src)
        loss = learned.train_windows(s, rows, epochs=400)
        acc = learned.window_accuracy(s, rows)
        out["predictor"] = {"trace_rows": int(len(rows)), "train_loss": round(loss, 4),
                            "shard_precision": round(acc["shard_precision"], 3),
                            "shard_recall": round(acc["shard_recall"], 3),
                            "seconds": round(time.perf_counter() - t0, 1)}
        dist = {}
        for name, pf in (("exact", 0), ("learned_prefetch", 1)):
            s.set_option(gmp.OPT_WINDOW_PREFETCH, pf)
            w0 = s.windows()
            d, st = s.sssp(src)
            w1 = s.windows()
            dist[name] = d
            out[name] = {"ms": round(st.ms_total, 3), "sweeps": st.iterations,
                         "ms_per_sweep": round(st.ms_total / max(st.iterations, 1), 3),
                         "streamed_gb": round((w1["streamed_bytes"] - w0["streamed_bytes"]) / 1e9, 3),
                         "windows_run": w1["windows_run"] - w0["windows_run"],
                         "windows_skipped": w1["windows_skipped"] - w0["windows_skipped"],
                         "prefetched": w1["prefetched"] - w0["prefetched"],
                         "prefetch_wasted_gb": round((w1["prefetch_wasted_bytes"] - w0["prefetch_wasted_bytes"]) / 1e9, 3)}
        s.set_option(gmp.OPT_WINDOW_PREFETCH, 0)
        out["identical_distances"] = bool(np.array_equal(dist["exact"], dist["learned_prefetch"]))
        out["speedup"] = round(out["exact"]["ms"] / out["learned_prefetch"]["ms"], 3)
    log(f"semi-external sssp: exact {out['exact']['ms']:.1f} ms ({out['exact']['sweeps']} sweeps, "
        f"{out['exact']['streamed_gb']} GB), learned prefetch {out['learned_prefetch']['ms']:.1f} ms "
        f"({out['learned_prefetch']['streamed_gb']} GB, {out['learned_prefetch']['prefetched']} prefetched, "
        f"{out['learned_prefetch']['prefetch_wasted_gb']} GB wasted); identical: {out['identical_distances']}; "
        f"predictor {out['predictor']}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-7)
    ap.add_argument("--max-iter", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bounded CPU baseline sample")
    ap.add_argument("--fp64-steps", type=int, default=40,
                    help="most FP64 CPU power iterations of the reference the PageRank errors are measured against")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and FP64 check")
    ap.add_argument("--no-extra", action="store_true", help="skip the SSSP and CC measurements")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched multi-instance runs")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-schedule comparison")
    ap.add_argument("--no-semi", action="store_true", help="skip the semi-external (host-resident edges) runs")
    ap.add_argument("--no-big", action="store_true", help="skip PageRank on R-MAT s28 ef17 (4.5e9 edges)")
    ap.add_argument("--learn-scale", type=int, default=27, help="R-MAT scale of the learned-SSSP evaluation")
    ap.add_argument("--strip-log", type=int, default=15,
                    help="PageRank source strips of 2^k ids (0: the row gather over the in-CSR)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the autotuner passes (the strip layout before the warm-up, the batched "
                         "kernels' knobs in the batched leg)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size "
                         "dispatch that scales the HBM byte counters under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the engine runs on a "
              "single MI355X (how it would shard over several is a design note, not code)", file=sys.stderr)
        return 2

    import graphee_amd as gmp

    scale, ef = args.scale, args.edge_factor
    seed = 1000 + scale
    eng = gmp.Engine(0)
    eng.set_option(gmp.OPT_STRIP_LOG, args.strip_log)
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)
    t0 = time.perf_counter()
    eng.generate_rmat(scale, ef, seed, weighted=True, num_shards=16)
    V, E, P, _ = eng.info()
    log(f"generated R-MAT s{scale} ef{ef}: V={V} E={E} in {time.perf_counter() - t0:.1f} s")

    tuned = None
    if args.autotune and args.strip_log:
        t0 = time.perf_counter()
        tuned = eng.autotune(iters=5).as_dict()
        tuned["seconds"] = round(time.perf_counter() - t0, 1)
        log(f"autotune: strip 2^{tuned['strip_log']}, bin 2^{tuned['bin_log']}, task 2^{tuned['bin_task_log']}, "
            f"2^{tuned['strip_task_log']} wave tasks per strip task, "
            f"push frontiers <= V >> {tuned['push_log']}, "
            f"{tuned['gather_ms_before']:.3f} -> {tuned['gather_ms_after']:.3f} ms/iteration in {tuned['seconds']} s")

    def solve():
        return eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol, want_ranks=False)

    for i in range(args.warmup):
        _, st = solve()
        log(f"warmup {i}: {st.iterations} iterations, {st.ms_total:.1f} ms")
    iters = 0
    kms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, st = solve()
        iters += st.iterations
        kms += st.ms_kernel
        launches += st.launches
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    gteps = E * iters / wall / 1e9
    o = 4 if E < 2**32 else 8
    bytes_launch = 8 * E + (o + 16) * V
    avg_launch_s = kms / 1e3 / max(launches, 1)
    achieved = bytes_launch / avg_launch_s / 1e9
    traffic = load_traffic(scale, ef)
    log(f"pagerank: {iters / args.steps:.0f} iterations/solve, {ms_step:.1f} ms/solve, {gteps:.1f} GTEPS, "
        f"iteration kernels {avg_launch_s * 1e3:.3f} ms = {achieved:.0f} GB/s")
    result = {
        "metric": "pagerank_gteps_rmat_s26",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic R-MAT generated on the GPU (a=0.57 b=0.19 c=0.19, seeded relabel, dedup)",
        "config": {"workload": f"pagerank-rmat-s{scale}-ef{ef}-tol{args.tol:g}",
                   "vertices": V, "edges": E, "iterations_per_step": iters / args.steps,
                   "damping": 0.85, "max_iter": args.max_iter},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": ("pagerank iteration: pr_strip_gather_kernel + pr_bin_kernel + "
                                "pr_split_apply_kernel" if eng.get_option(gmp.OPT_STRIP_LOG) else
                                "pagerank iteration: pr_gather_kernel + pr_hub_combine_kernel"),
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "avg_launch_ms": round(avg_launch_s * 1e3, 4)},
        "cpu_baseline": None,
        "extra": {"autotune": tuned},
    }

    semi_ref = None
    if not args.no_extra and not args.no_semi:
        semi_ref, _ = eng.pagerank(damping=0.85, max_iter=20)
        semi_ref = semi_ref.astype(np.float64)

    if not args.no_extra:
        # the row-by-row gather over the in-CSR (strip layout off), for comparison
        strip_now = eng.get_option(gmp.OPT_STRIP_LOG)
        eng.set_option(gmp.OPT_STRIP_LOG, 0)
        eng.pagerank(damping=0.85, max_iter=3, want_ranks=False)
        _, str_ = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol, want_ranks=False)
        eng.set_option(gmp.OPT_STRIP_LOG, strip_now)
        kms_r = str_.ms_kernel / max(str_.launches, 1)
        result["extra"]["pagerank_row_gather"] = {
            "ms_per_iteration": round(kms_r, 4), "iterations": str_.iterations,
            "yardstick_gbs": round(bytes_launch / kms_r / 1e6, 1),
            "strip_speedup": round(kms_r / (avg_launch_s * 1e3), 3)}
        log(f"pagerank row gather: {kms_r:.3f} ms/iteration ({kms_r / (avg_launch_s * 1e3):.2f}x the strip path)")
        # compressed ids: on the strip layout its codes bit-packed per gather round (the default
        # with a strip width), in the row gather delta/varint column ids (strip layout off)
        eng.pagerank(damping=0.85, max_iter=3, compressed=True, want_ranks=False)  # packs the codes
        rc, stc = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol, compressed=True)
        _, stc = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol, compressed=True,
                              want_ranks=False)
        rp, stp = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol)
        same = bool(np.array_equal(rc.view(np.uint32), rp.view(np.uint32))) and stc.iterations == stp.iterations
        del rc, rp
        kms_c = stc.ms_kernel / max(stc.launches, 1)
        result["extra"]["pagerank_compressed"] = {
            "layout": "source strips, codes bit-packed per 512-edge round" if strip_now else "row gather, varint ids",
            "ms_per_solve": round(stc.ms_total, 3), "iterations": stc.iterations,
            "gteps": round(E * stc.iterations / stc.ms_total / 1e6, 3),
            "ms_per_iteration": round(kms_c, 4), "yardstick_gbs": round(bytes_launch / kms_c / 1e6, 1),
            "bit_identical_to_uncompressed": same,
            "index_bytes": stc.index_bytes, "uncompressed_index_bytes": stp.index_bytes,
            "index_bytes_per_edge": round(stc.index_bytes / E, 3)}
        log(f"pagerank compressed: {stc.index_bytes / E:.3f} B/edge index ({stp.index_bytes / E:.3f} uncompressed), "
            f"{kms_c:.3f} ms/iteration, bit-identical {same}")
        if strip_now:
            # the strip gather over each edge's 4-byte source id, the index exactly as the in-CSR
            # stores it (the yardstick's 4 B per edge), in place of the 2-byte offset in the strip
            eng.pagerank(damping=0.85, max_iter=3, strip_ids=True, want_ranks=False)  # builds the ids
            ri, sti = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol, strip_ids=True)
            _, sti = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol, strip_ids=True,
                                  want_ranks=False)
            rp, stp = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol)
            same_i = bool(np.array_equal(ri.view(np.uint32), rp.view(np.uint32))) and sti.iterations == stp.iterations
            del ri, rp
            kms_i = sti.ms_kernel / max(sti.launches, 1)
            result["extra"]["pagerank_strip_ids"] = {
                "layout": "source strips, 4-byte source ids", "ms_per_solve": round(sti.ms_total, 3),
                "iterations": sti.iterations, "gteps": round(E * sti.iterations / sti.ms_total / 1e6, 3),
                "ms_per_iteration": round(kms_i, 4), "yardstick_gbs": round(bytes_launch / kms_i / 1e6, 1),
                "yardstick_frac": round(bytes_launch / kms_i / 1e6 / HBM_PEAK_GBS, 4),
                "bit_identical_to_codes": same_i, "index_bytes": sti.index_bytes,
                "index_bytes_per_edge": round(sti.index_bytes / E, 3)}
            log(f"pagerank strip ids: {sti.index_bytes / E:.3f} B/edge index, {kms_i:.3f} ms/iteration = "
                f"{bytes_launch / kms_i / 1e6:.0f} GB/s on the yardstick, bit-identical {same_i}")
        if strip_now:
            eng.set_option(gmp.OPT_STRIP_LOG, 0)
            eng.pagerank(damping=0.85, max_iter=3, compressed=True, want_ranks=False)  # builds the varint lists
            _, stv = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol, compressed=True,
                                  want_ranks=False)
            eng.set_option(gmp.OPT_STRIP_LOG, strip_now)
            kms_v = stv.ms_kernel / max(stv.launches, 1)
            result["extra"]["pagerank_compressed_row_gather"] = {
                "ms_per_iteration": round(kms_v, 4), "iterations": stv.iterations,
                "yardstick_gbs": round(bytes_launch / kms_v / 1e6, 1),
                "index_bytes": stv.index_bytes, "index_bytes_per_edge": round(stv.index_bytes / E, 3)}
            log(f"pagerank compressed row gather: {stv.index_bytes / E:.3f} B/edge varint ids, {kms_v:.3f} ms/iteration")

    if not args.no_cpu:
        import oracle
        ranks, _ = eng.pagerank(damping=0.85, max_iter=args.max_iter, tol=args.tol)
        g = eng.download()
        threads = cpu_threads()
        l1, linf = oracle.pr_step_residual(g["off"], g["idx"], g["outdeg"], ranks, threads=threads)
        t0 = time.perf_counter()
        err = oracle.pr_f64_error(g["off"], g["idx"], g["outdeg"], ranks, threads=threads,
                                  max_iter=args.fp64_steps, bound=1e-11)
        result["extra"]["pagerank_fp64_check"] = {
            "l1_error": err["l1_error"], "linf_error": err["linf_error"],
            "linf_error_upper_bound": err["linf_error"] + err["reference_l1_bound"],
            "reference_l1_bound": err["reference_l1_bound"], "reference_steps": err["reference_steps"],
            "reference_seconds": round(time.perf_counter() - t0, 1),
            "l1_error_bound": l1 / 0.15, "step_residual_l1": l1, "step_residual_linf": linf,
            "method": "errors: |x - x_ref| of the GPU ranks x against an FP64 CPU power iteration x_ref "
                      "started from x and run until its a-posteriori bound ||x_ref - x*||_1 <= "
                      "||x_ref - x_ref_prev||_1 d/(1-d) is reference_l1_bound (each error is exact to within "
                      "it); l1_error_bound = ||T(x)-x||_1/(1-d) from one FP64 step (step_residual_*: T(x)-x)"}
        log(f"FP64 check: L1 error {err['l1_error']:.3e}, Linf error {err['linf_error']:.3e} "
            f"(reference within {err['reference_l1_bound']:.1e} after {err['reference_steps']} steps); "
            f"a-posteriori L1 bound {l1 / 0.15:.3e}")
        msg0 = np.full(V, 1.0 / V, dtype=np.float32)
        _, st1 = oracle.graphmp_baseline(oracle.APP_PAGERANK, g["off"], g["idx"], g["bounds"], msg0, 1,
                                         outdeg=g["outdeg"], threads=threads)
        n = int(max(1, min(args.max_iter, args.cpu_seconds / max(st1.seconds, 1e-3))))
        _, stc = oracle.graphmp_baseline(oracle.APP_PAGERANK, g["off"], g["idx"], g["bounds"], msg0, n,
                                         outdeg=g["outdeg"], threads=threads)
        cpu_gteps = stc.edges / stc.seconds / 1e9
        result["cpu_baseline"] = {
            "value": round(cpu_gteps, 4), "unit": "GTEPS", "cores": stc.threads, "kind": "port",
            "sample": f"{stc.iterations} PageRank iterations of the GraphMP OpenMP engine restatement "
                      f"(16 shards, guided schedule, edges in RAM) on the same s{scale} graph, "
                      f"{stc.seconds:.1f} s",
            "host": host_info()}
        log(f"cpu baseline: {cpu_gteps:.3f} GTEPS on {stc.threads} threads ({stc.iterations} iterations)")
        del g

    if not args.no_extra:
        src = eng.pick_source(seed)
        for name, method in (("sssp_bf", gmp.SSSP_BF), ("sssp_delta64", gmp.SSSP_DELTA)):
            eng.sssp(src, method=method, delta=64, want_dist=False)  # builds the out-CSR once
            t0 = time.perf_counter()
            d, st = eng.sssp(src, method=method, delta=64)
            t = time.perf_counter() - t0
            reached = None
            if d is not None:
                reached = int((d != 0xFFFFFFFF).sum())
            # yardstick: 12 B per relaxed edge (target id, weight, target distance)
            result["extra"][name] = {"ms": round(st.ms_total, 3), "wall_ms": round(t * 1e3, 3),
                                     "edges_relaxed": st.edges, "work_gteps": round(st.edges / st.ms_total / 1e6, 3),
                                     "yardstick_gbs": round(st.bytes / st.ms_total / 1e6, 1),
                                     "hbm_frac": round(st.bytes / st.ms_total / 1e6 / HBM_PEAK_GBS, 4),
                                     "iterations": st.iterations, "source": src, "reached": reached}
            log(f"{name}: {st.ms_total:.1f} ms, {st.iterations} iterations, {st.edges} relaxations")
            if name == "sssp_delta64":
                gpu_dist, gpu_ms = d, st.ms_total
            else:
                bf_dist, bf_ms = d, st.ms_total
        if not args.no_cpu:
            # full GraphMP-engine SSSP on the host (synchronous weighted rule, selective scheduling)
            import oracle
            g = eng.download(weights=True)
            msg0 = np.full(V, 0xFFFFFFFF, dtype=np.uint32)
            msg0[src] = 0
            threads = cpu_threads()
            dist, stc = oracle.graphmp_baseline(oracle.APP_SSSP_W, g["off"], g["idx"], g["bounds"], msg0, 10000,
                                                w=g["w"], threads=threads)
            same = bool(np.array_equal(dist, gpu_dist) and np.array_equal(dist, bf_dist))
            result["extra"]["sssp_cpu_baseline"] = {
                "seconds": round(stc.seconds, 3), "threads": stc.threads, "sweeps": stc.iterations,
                "edges_gathered": stc.edges, "identical_to_gpu": same,
                "gpu_delta64_speedup": round(stc.seconds * 1e3 / gpu_ms, 2),
                "gpu_bf_speedup": round(stc.seconds * 1e3 / bf_ms, 2)}
            log(f"sssp cpu baseline: {stc.seconds:.2f} s ({stc.iterations} sweeps) -> GPU delta-64 is "
                f"{stc.seconds * 1e3 / gpu_ms:.1f}x, GPU Bellman-Ford sweeps {stc.seconds * 1e3 / bf_ms:.1f}x; "
                f"identical: {same}")
            del g, dist
        eng.generate_rmat(scale, ef, seed, undirected=True, num_shards=16)
        Vs, Es, _, _ = eng.info()
        # config 3's CC: min-label propagation (GraphMP/app/cc.cpp:40-56 on the symmetrised graph),
        # sweeps over the source strips with frontier pushes for the sparse tail; yardstick 8 B per
        # gathered edge
        eng.set_option(gmp.OPT_CC_UNION_FIND, 0)
        eng.cc(want_labels=False)
        lab, st = eng.cc()
        eng.set_option(gmp.OPT_CC_UNION_FIND, 1)
        ncomp = int((lab == np.arange(Vs, dtype=np.uint32)).sum())
        result["extra"]["cc"] = {"ms": round(st.ms_total, 3), "edges_symmetric": Es,
                                 "algorithm": "min-label propagation: strip sweeps + frontier pushes",
                                 "iterations": st.iterations, "sweeps": st.epochs,
                                 "gteps": round(Es * st.epochs / st.ms_total / 1e6, 3),
                                 "edges_gathered_or_pushed": st.edges,
                                 "work_gteps": round(st.edges / st.ms_total / 1e6, 3),
                                 "yardstick_gbs": round(st.bytes / st.ms_total / 1e6, 1),
                                 "hbm_frac": round(st.bytes / st.ms_total / 1e6 / HBM_PEAK_GBS, 4),
                                 "components": ncomp}
        log(f"cc (min-label): {st.ms_total:.1f} ms, {st.iterations} iterations ({st.epochs} sweeps), "
            f"{ncomp} components")
        # the same labels by sampled union-find (not the config's algorithm: reported beside it)
        eng.cc(want_labels=False)
        lab2, st2 = eng.cc()
        result["extra"]["cc_union_find"] = {
            "ms": round(st2.ms_total, 3), "algorithm": "sampled union-find (2 neighbours, giant component skipped)",
            "phases": st2.iterations, "edges_examined": st2.edges,
            "gteps_on_edges_examined": round(st2.edges / st2.ms_total / 1e6, 3),
            "identical_to_min_label": bool(np.array_equal(lab, lab2))}
        log(f"cc union-find: {st2.ms_total:.1f} ms, {st2.edges} edges examined, identical: {np.array_equal(lab, lab2)}")
        del lab2
        if not args.no_cpu:
            import oracle
            g = eng.download()
            threads = cpu_threads()
            labc, stc = oracle.graphmp_baseline(oracle.APP_CC, g["off"], g["idx"], g["bounds"],
                                                np.arange(Vs, dtype=np.uint32), 10000, threads=threads)
            result["extra"]["cc_cpu_baseline"] = {
                "seconds": round(stc.seconds, 3), "threads": stc.threads, "sweeps": stc.iterations,
                "edges_gathered": stc.edges, "identical_to_gpu": bool(np.array_equal(labc, lab)),
                "gpu_min_label_speedup": round(stc.seconds * 1e3 / st.ms_total, 2),
                "gpu_union_find_speedup": round(stc.seconds * 1e3 / st2.ms_total, 2)}
            log(f"cc cpu baseline: {stc.seconds:.2f} s ({stc.iterations} sweeps) -> GPU min-label is "
                f"{stc.seconds * 1e3 / st.ms_total:.1f}x; identical: {np.array_equal(labc, lab)}")
            del g, labc

    if not args.no_extra and not args.no_cpu:
        result["extra"]["pagerank_rmat_s20_x20"] = run_pagerank_s20(eng, cpu_threads())
        result["extra"]["sssp_rmat_s24"] = run_sssp_s24(eng, cpu_threads())
    if not args.no_extra and not args.no_batched:
        result["extra"]["batched"] = run_batched(eng, verify=not args.no_cpu, tune=args.autotune)
    if not args.no_extra and not args.no_learned:
        result["extra"]["learned_sssp"] = run_learned(eng, args.learn_scale)

    eng.close()
    if not args.no_extra and not args.no_big:
        result["extra"]["pagerank_beyond_2e32_edges"] = run_big(verify=not args.no_cpu, threads=cpu_threads())
    if semi_ref is not None:
        # one engine in the process, like a semi-external user: its stream and two copy streams fit
        # the device's hardware queues (GPU_MAX_HW_QUEUES, 4 by default), so copies overlap gathers
        result["extra"]["semi_external"] = run_semi(scale, ef, seed, semi_ref)
        del semi_ref
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
