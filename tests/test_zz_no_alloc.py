"""gmp_pagerank_noalloc (no device allocation; working memory in a fixed per-GPU array) against
gmp_pagerank on the same graphs and against the FP64 oracle, through the C ABI: two calls in a
row, the second on a smaller graph, and the size limit."""
import ctypes
import sys

import numpy as np
import pytest

import graphee_amd as gmp
import oracle

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _Hip:
    """Device buffers for the test's inputs and outputs (the caller-owned memory of the entry point)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 4))) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p.value

    def download(self, ptr, n, dtype):
        out = np.zeros(n, dtype=dtype)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr),
                                ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _case(engine, hip, scale, seed, **kw):
    say(f"no-alloc pagerank: R-MAT s{scale} seed {seed} {kw}")
    engine.generate_rmat(scale, 16, seed)
    g = engine.download()
    V, E = g["outdeg"].size, g["idx"].size
    d_off = hip.upload(g["off"].astype(np.uint64))
    d_idx = hip.upload(g["idx"].astype(np.uint32))
    d_od = hip.upload(g["outdeg"].astype(np.uint32))
    d_out = hip.upload(np.full(V, np.nan, dtype=np.float32))  # every entry must be written
    st = engine.pagerank_noalloc(V, E, d_off, d_idx, d_od, d_out, **kw)
    r_na = hip.download(d_out, V, np.float32)
    r_ref, st_ref = engine.pagerank(**kw)
    x, _, _ = oracle.pr_f64(g["off"], g["idx"], g["outdeg"], d=kw.get("damping", 0.85),
                            max_iter=kw.get("max_iter", 20), tol=kw.get("tol", 0.0), seed=kw.get("seed_vertex", -1))
    return g, st, r_na, st_ref, r_ref, x


def test_noalloc_pagerank_matches_gmp_pagerank(engine):
    hip = _Hip()
    try:
        # first call: uniform teleport, fixed iteration count
        g, st, r_na, st_ref, r_ref, x = _case(engine, hip, 14, 7014, max_iter=30)
        assert np.isfinite(r_na).all()
        assert st.iterations == st_ref.iterations == 30
        assert np.abs(r_na.astype(np.float64) - x).sum() < PR_L1_TOL
        assert np.abs(r_na.astype(np.float64) - r_ref).sum() < PR_L1_TOL
        assert abs(st.dangling - st_ref.dangling) < 1e-6
        assert abs(st.residual - st_ref.residual) < 1e-6
        # second call on a smaller graph (the scratch array still holds the first call's data):
        # personalised teleport and the residual stop
        seed = int(np.argmax(g["outdeg"][: 1 << 11]))
        g2, st2, r2, st2_ref, r2_ref, x2 = _case(engine, hip, 11, 7011, max_iter=100, tol=1e-7, seed_vertex=seed)
        assert np.isfinite(r2).all()
        assert st2.residual < 1e-7 and abs(st2.iterations - st2_ref.iterations) <= 1
        assert np.abs(r2.astype(np.float64) - x2).sum() < PR_L1_TOL
        assert np.abs(r2.astype(np.float64) - r2_ref).sum() < PR_L1_TOL
        # and the first graph again, uncompensated: still within the bar
        g3, st3, r3, _, r3_ref, x3 = _case(engine, hip, 12, 7012, max_iter=20, kahan=False)
        assert np.abs(r3.astype(np.float64) - x3).sum() < PR_L1_TOL
    finally:
        hip.free()


def test_noalloc_pagerank_refuses_sizes_beyond_its_array(engine):
    hip = _Hip()
    try:
        d = hip.upload(np.zeros(16, dtype=np.uint64))
        with pytest.raises(gmp.GmpError) as ei:
            engine.pagerank_noalloc(gmp.NOALLOC_MAX_VERTICES + 1, 0, d, d, d, d)
        assert ei.value.code == gmp.EINVAL and "scratch" in str(ei.value)
        with pytest.raises(gmp.GmpError) as ei:
            engine.pagerank_noalloc(4, 8, d, 0, d, d)  # edges but no column ids
        assert ei.value.code == gmp.EINVAL
    finally:
        hip.free()
