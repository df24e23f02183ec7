"""ctypes binding of libgmp.so, the gfx950 vertex-centric graph engine.

The library is the product: HIP kernels plus the C ABI declared in include/gmp.h.  This
module only loads it and converts NumPy arrays to plain pointers.  It never falls back
to a CPU implementation: when libgmp.so is missing, `lib()` raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgmp.so")

OK, EINVAL, EIO, EFORMAT, EHIP, ENOMEM, ESTATE, EUNSUPPORTED = range(8)
APP_PAGERANK, APP_SSSP, APP_CC, APP_SSSP_W = 1, 2, 3, 4
SSSP_BF, SSSP_DELTA, SSSP_LEARNED = 0, 1, 2
PR_NO_KAHAN, PR_COMPRESSED, PR_STRIP_IDS = 1, 2, 4
GRAPH_OFF64 = 1
EDGES_WEIGHTED, EDGES_SYMMETRIZE = 1, 2
OPT_REORDER, OPT_STREAM_CAP, OPT_ROW_CAP, OPT_CHUNK, OPT_WG_PER_CU, OPT_VEC_MIN = 1, 2, 3, 4, 5, 6
OPT_STRIP_LOG, OPT_BIN_LOG, OPT_BIN_TASK_LOG = 7, 8, 9
OPT_EDGE_BUDGET_MB, OPT_WINDOW_MB, OPT_WINDOW_PACKED, OPT_PUSH_LOG, OPT_CC_UNION_FIND = 10, 11, 12, 13, 14
OPT_DELTA_PULL_LOG, OPT_LEARNED_FINISH = 15, 16
NOALLOC_MAX_VERTICES = 1 << 22
(OPT_BATCH_SSSP_UNROLL, OPT_BATCH_SSSP_GRID, OPT_BATCH_PPR_PIECE, OPT_BATCH_PPR_RUN_ROWS, OPT_BATCH_PPR_RUN_EDGES,
 OPT_BATCH_PPR_LOADS, OPT_BATCH_PPR_GRID) = range(17, 24)
OPT_MLP_NET, OPT_WINDOW_PREFETCH = 24, 25
OPT_STRIP_TASK_LOG, OPT_SELECTIVE = 26, 27
OPT_BATCH_SSSP_SCAN_LOG, OPT_BATCH_SSSP_HUB_LOG = 28, 29

# every entry point include/gmp.h declares
SYMBOLS = (
    "gmp_last_error", "gmp_abi_version", "gmp_create", "gmp_destroy", "gmp_load_csr",
    "gmp_load_shards", "gmp_load_edge_list", "gmp_generate_rmat", "gmp_graph_info", "gmp_graph_download",
    "gmp_pick_source", "gmp_run_app", "gmp_pagerank", "gmp_pagerank_noalloc", "gmp_sssp", "gmp_cc",
    "gmp_set_option", "gmp_get_option", "gmp_batch_generate", "gmp_batch_info", "gmp_batch_download",
    "gmp_batch_sssp", "gmp_batch_ppr", "gmp_batch_run", "gmp_batch_autotune", "gmp_autotune", "gmp_tune_save", "gmp_tune_load",
    "gmp_sssp_trace", "gmp_sssp_actions", "gmp_window_trace", "gmp_mlp_init", "gmp_mlp_set", "gmp_mlp_get", "gmp_mlp_forward", "gmp_mlp_loss_grad",
    "gmp_mlp_train", "gmp_windows",
    "gmp_partition_bounds", "gmp_npy_read_u32", "gmp_npy_write_i32", "gmp_write_shards", "gmp_convert_edge_list",
)


class GmpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[gmp error {code}] {msg}")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32), ("launches", ctypes.c_int32),
        ("ms_total", ctypes.c_double), ("ms_kernel", ctypes.c_double),
        ("edges", ctypes.c_uint64), ("vertices", ctypes.c_uint64),
        ("residual", ctypes.c_double), ("dangling", ctypes.c_double),
        ("changed", ctypes.c_uint64), ("bytes", ctypes.c_uint64), ("index_bytes", ctypes.c_uint64),
        ("epochs", ctypes.c_int32), ("policy", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RmatParams(ctypes.Structure):
    _fields_ = [
        ("scale", ctypes.c_uint32), ("edge_factor", ctypes.c_uint32), ("seed", ctypes.c_uint64),
        ("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double),
        ("undirected", ctypes.c_uint32), ("weighted", ctypes.c_uint32),
        ("num_shards", ctypes.c_uint32), ("flags", ctypes.c_uint32),
    ]


class PrParams(ctypes.Structure):
    _fields_ = [
        ("damping", ctypes.c_float), ("max_iter", ctypes.c_int32), ("tol", ctypes.c_double),
        ("seed_vertex", ctypes.c_int64), ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
    ]


class SsspParams(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32), ("delta", ctypes.c_uint32),
        ("unit_weights", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
    ]


MLP_IN, MLP_HIDDEN, MLP_CLASSES, MLP_SHARDS, MLP_PARAMS = 16, 64, 8, 64, 5768

TRACE_DTYPE = np.dtype([("feat", np.float32, (16,)), ("best", np.int32), ("epoch", np.uint32),
                        ("shard_mask", np.uint64), ("ms", np.float32, (8,)), ("relaxed", np.float32, (8,))])


class WindowStats(ctypes.Structure):
    _fields_ = [("windows", ctypes.c_uint32), ("cached", ctypes.c_uint32), ("ring_slots", ctypes.c_uint32),
                ("packed", ctypes.c_uint32), ("host_bytes", ctypes.c_uint64), ("device_bytes", ctypes.c_uint64),
                ("streamed_bytes", ctypes.c_uint64), ("windows_run", ctypes.c_uint64),
                ("windows_skipped", ctypes.c_uint64), ("prefetched", ctypes.c_uint64),
                ("prefetch_wasted_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(ctypes.Structure):
    _fields_ = [("stream_cap", ctypes.c_uint32), ("row_cap", ctypes.c_uint32), ("chunk", ctypes.c_uint32),
                ("vec_min", ctypes.c_uint32), ("wg_per_cu", ctypes.c_uint32), ("trials", ctypes.c_uint32),
                ("gather_ms_before", ctypes.c_double), ("gather_ms_after", ctypes.c_double),
                ("strip_log", ctypes.c_uint32), ("bin_log", ctypes.c_uint32), ("bin_task_log", ctypes.c_uint32),
                ("push_log", ctypes.c_uint32), ("delta_pull_log", ctypes.c_uint32), ("strip_task_log", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BatchTuneResult(ctypes.Structure):
    _fields_ = [("sssp_unroll", ctypes.c_uint32), ("sssp_grid_per_cu", ctypes.c_uint32),
                ("ppr_piece", ctypes.c_uint32), ("ppr_run_rows", ctypes.c_uint32), ("ppr_run_edges", ctypes.c_uint32),
                ("ppr_loads", ctypes.c_uint32), ("ppr_grid_per_cu", ctypes.c_uint32), ("trials", ctypes.c_uint32),
                ("sssp_ms_before", ctypes.c_double), ("sssp_ms_after", ctypes.c_double),
                ("ppr_ms_before", ctypes.c_double), ("ppr_ms_after", ctypes.c_double),
                ("sssp_scan_log", ctypes.c_uint32), ("sssp_hub_log", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_P = ctypes.c_void_p


def lib():
    """Loads libgmp.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "or python graphee_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    u32, u64, i32, i64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int64
    sig = {
        "gmp_last_error": ([], ctypes.c_char_p),
        "gmp_abi_version": ([], ctypes.c_int),
        "gmp_create": ([i32, ctypes.POINTER(_P)], ctypes.c_int),
        "gmp_destroy": ([_P], ctypes.c_int),
        "gmp_load_csr": ([_P, u32, u64, _P, _P, _P, u32, _P], ctypes.c_int),
        "gmp_load_shards": ([_P, ctypes.c_char_p, u32, u32], ctypes.c_int),
        "gmp_load_edge_list": ([_P, ctypes.c_char_p, u32, u32, u32], ctypes.c_int),
        "gmp_generate_rmat": ([_P, ctypes.POINTER(RmatParams)], ctypes.c_int),
        "gmp_graph_info": ([_P, _P, _P, _P, _P], ctypes.c_int),
        "gmp_windows": ([_P, _P], ctypes.c_int),
        "gmp_graph_download": ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
        "gmp_pick_source": ([_P, u64, u32, _P], ctypes.c_int),
        "gmp_set_option": ([_P, i32, i64], ctypes.c_int),
        "gmp_autotune": ([_P, i32, ctypes.POINTER(TuneResult)], ctypes.c_int),
        "gmp_sssp_trace": ([_P, u32, ctypes.POINTER(SsspParams), _P, u32, _P, _P, ctypes.POINTER(Stats)],
                           ctypes.c_int),
        "gmp_sssp_actions": ([_P, _P, u32], ctypes.c_int),
        "gmp_window_trace": ([_P, i32, u32, _P, u32, _P, ctypes.POINTER(Stats)], ctypes.c_int),
        "gmp_mlp_init": ([_P, u64], ctypes.c_int),
        "gmp_mlp_set": ([_P, _P], ctypes.c_int),
        "gmp_mlp_get": ([_P, _P], ctypes.c_int),
        "gmp_mlp_forward": ([_P, _P, u32, _P], ctypes.c_int),
        "gmp_mlp_loss_grad": ([_P, _P, _P, _P, u32, _P, _P], ctypes.c_int),
        "gmp_mlp_train": ([_P, _P, _P, _P, u32, i32, ctypes.c_float, _P], ctypes.c_int),
        "gmp_tune_save": ([_P, ctypes.c_char_p], ctypes.c_int),
        "gmp_tune_load": ([_P, ctypes.c_char_p], ctypes.c_int),
        "gmp_batch_generate": ([_P, u32, _P], ctypes.c_int),
        "gmp_batch_info": ([_P, u32, _P, _P], ctypes.c_int),
        "gmp_batch_download": ([_P, u32, _P, _P, _P], ctypes.c_int),
        "gmp_batch_sssp": ([_P, u32, _P, _P, _P, ctypes.POINTER(Stats)], ctypes.c_int),
        "gmp_batch_ppr": ([_P, u32, _P, _P, ctypes.c_float, i32, _P, ctypes.POINTER(Stats)], ctypes.c_int),
        "gmp_batch_run": ([_P, u32, _P, _P, u32, _P, _P, ctypes.c_float, i32, _P, _P, ctypes.POINTER(Stats),
                           ctypes.POINTER(Stats)], ctypes.c_int),
        "gmp_batch_autotune": ([_P, u32, _P, _P, u32, _P, _P, ctypes.c_float, i32, ctypes.POINTER(BatchTuneResult)],
                               ctypes.c_int),
        "gmp_get_option": ([_P, i32, _P], ctypes.c_int),
        "gmp_run_app": ([_P, i32, i32, _P, _P, _P, _P, ctypes.POINTER(Stats)], ctypes.c_int),
        "gmp_pagerank": ([_P, ctypes.POINTER(PrParams), _P, ctypes.POINTER(Stats)], ctypes.c_int),
        "gmp_pagerank_noalloc": ([_P, u32, u64, _P, _P, _P, ctypes.POINTER(PrParams), _P, ctypes.POINTER(Stats)],
                                 ctypes.c_int),
        "gmp_sssp": ([_P, u32, ctypes.POINTER(SsspParams), _P, ctypes.POINTER(Stats)], ctypes.c_int),
        "gmp_cc": ([_P, _P, ctypes.POINTER(Stats)], ctypes.c_int),
        "gmp_partition_bounds": ([_P, u32, u32, _P], ctypes.c_int),
        "gmp_npy_read_u32": ([ctypes.c_char_p, _P, u64, _P], ctypes.c_int),
        "gmp_npy_write_i32": ([ctypes.c_char_p, _P, u64], ctypes.c_int),
        "gmp_write_shards": ([ctypes.c_char_p, u32, _P, _P, _P, u32, _P, _P], ctypes.c_int),
        "gmp_convert_edge_list": ([ctypes.c_char_p, ctypes.c_char_p, u32, u32, u32, _P], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def check(rc):
    if rc != OK:
        raise GmpError(rc, lib().gmp_last_error().decode(errors="replace"))


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---- host-side format utilities (no GPU) ----------------------------------------------

def partition_bounds(indeg, P):
    indeg = np.ascontiguousarray(indeg, dtype=np.uint64)
    b = np.zeros(P + 1, dtype=np.uint32)
    check(lib().gmp_partition_bounds(_ptr(indeg), indeg.size, P, _ptr(b)))
    return b


def npy_read_u32(path):
    n = ctypes.c_uint64(0)
    check(lib().gmp_npy_read_u32(os.fsencode(path), None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, dtype=np.uint32)
    check(lib().gmp_npy_read_u32(os.fsencode(path), _ptr(out), out.size, ctypes.byref(n)))
    return out


def npy_write_i32(path, data):
    data = np.ascontiguousarray(data, dtype=np.int32)
    check(lib().gmp_npy_write_i32(os.fsencode(path), _ptr(data), data.size))


def write_shards(dir_, off, idx, bounds, weights=None, vertex_out=None):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    V = off.size - 1
    bounds = _u32(bounds)
    w = None if weights is None else _u32(weights)
    vo = None if vertex_out is None else np.ascontiguousarray(vertex_out, dtype=np.int32)
    check(lib().gmp_write_shards(os.fsencode(dir_), V, _ptr(off), _ptr(_u32(idx)), _ptr(w),
                                 bounds.size - 1, _ptr(bounds), _ptr(vo)))


def convert_edge_list(path, dir_, V=0, P=1, weighted=False, symmetrize=False):
    """Text edge list -> GraphMP shard directory (gmp_convert_edge_list, no GPU); returns V."""
    flags = (EDGES_WEIGHTED if weighted else 0) | (EDGES_SYMMETRIZE if symmetrize else 0)
    v = ctypes.c_uint32()
    check(lib().gmp_convert_edge_list(os.fsencode(path), os.fsencode(dir_), V, P, flags, ctypes.byref(v)))
    return v.value


# ---- engine --------------------------------------------------------------------------

class Engine:
    """One engine per GPU: the graph stays in HBM between runs."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(lib().gmp_create(device, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().gmp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # graph
    def load_csr(self, off, idx, weights=None, P=1, bounds=None):
        off = np.ascontiguousarray(off, dtype=np.uint64)
        idx = _u32(idx)
        V = off.size - 1
        w = None if weights is None else _u32(weights)
        b = None if bounds is None else _u32(bounds)
        check(lib().gmp_load_csr(self._h, V, int(off[-1]), _ptr(off), _ptr(idx), _ptr(w), P, _ptr(b)))

    def load_shards(self, dir_, V, P):
        check(lib().gmp_load_shards(self._h, os.fsencode(dir_), V, P))

    def load_edge_list(self, path, V=0, P=1, weighted=False, symmetrize=False):
        """Text edge list ("src dst [weight]" per line) -> the engine's graph (gmp_load_edge_list)."""
        flags = (EDGES_WEIGHTED if weighted else 0) | (EDGES_SYMMETRIZE if symmetrize else 0)
        check(lib().gmp_load_edge_list(self._h, os.fsencode(path), V, P, flags))

    def generate_rmat(self, scale, edge_factor, seed, a=0.57, b=0.19, c=0.19, undirected=False,
                      weighted=False, num_shards=1, off64=False):
        p = RmatParams(scale, edge_factor, seed, a, b, c, int(undirected), int(weighted), num_shards,
                       GRAPH_OFF64 if off64 else 0)
        check(lib().gmp_generate_rmat(self._h, ctypes.byref(p)))

    def info(self):
        V, E, P, w = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().gmp_graph_info(self._h, ctypes.byref(V), ctypes.byref(E), ctypes.byref(P),
                                   ctypes.byref(w)))
        return V.value, E.value, P.value, bool(w.value)

    def download(self, weights=False):
        V, E, P, weighted = self.info()
        off = np.zeros(V + 1, dtype=np.uint64)
        idx = np.zeros(E, dtype=np.uint32)
        w = np.zeros(E, dtype=np.uint32) if (weights and weighted) else None
        bounds = np.zeros(P + 1, dtype=np.uint32)
        od = np.zeros(V, dtype=np.uint32)
        check(lib().gmp_graph_download(self._h, _ptr(off), _ptr(idx), _ptr(w), _ptr(bounds), _ptr(od)))
        return {"off": off.astype(np.int64), "idx": idx, "w": w, "bounds": bounds, "outdeg": od}

    def windows(self):
        """Residency of the loaded graph (windows == 0: resident) and the window counters."""
        r = WindowStats()
        check(lib().gmp_windows(self._h, ctypes.byref(r)))
        return r.as_dict()

    def set_option(self, key, value):
        check(lib().gmp_set_option(self._h, key, int(value)))

    def get_option(self, key):
        v = ctypes.c_int64()
        check(lib().gmp_get_option(self._h, key, ctypes.byref(v)))
        return v.value

    def autotune(self, iters=3):
        r = TuneResult()
        check(lib().gmp_autotune(self._h, iters, ctypes.byref(r)))
        return r

    def tune_save(self, path):
        check(lib().gmp_tune_save(self._h, os.fsencode(path)))

    def tune_load(self, path):
        check(lib().gmp_tune_load(self._h, os.fsencode(path)))

    def pick_source(self, seed, salt=3):
        v = ctypes.c_uint32()
        check(lib().gmp_pick_source(self._h, seed, salt, ctypes.byref(v)))
        return v.value

    # batched multi-instance runs
    def batch_generate(self, specs):
        """specs: list of dicts with scale, edge_factor, seed (weighted graphs)."""
        arr = (RmatParams * len(specs))()
        for i, s in enumerate(specs):
            arr[i] = RmatParams(s["scale"], s["edge_factor"], s["seed"], s.get("a", 0.57), s.get("b", 0.19),
                                s.get("c", 0.19), int(s.get("undirected", False)), int(s.get("weighted", True)),
                                1, 0)
        check(lib().gmp_batch_generate(self._h, len(specs), ctypes.cast(arr, ctypes.c_void_p)))

    def batch_info(self, g):
        V, E = ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().gmp_batch_info(self._h, g, ctypes.byref(V), ctypes.byref(E)))
        return V.value, E.value

    def batch_download(self, g):
        V, E = self.batch_info(g)
        off = np.zeros(V + 1, dtype=np.uint64)
        idx = np.zeros(E, dtype=np.uint32)
        w = np.zeros(E, dtype=np.uint32)
        check(lib().gmp_batch_download(self._h, g, _ptr(off), _ptr(idx), _ptr(w)))
        return off.astype(np.int64), idx, w

    def _split(self, graphs, flat, dtype):
        out, pos = [], 0
        for g in graphs:
            V = self.batch_info(int(g))[0]
            out.append(flat[pos:pos + V])
            pos += V
        return out

    def batch_sssp(self, graphs, sources):
        graphs, sources = _u32(graphs), _u32(sources)
        total = sum(self.batch_info(int(g))[0] for g in graphs)
        out = np.zeros(total, dtype=np.uint32)
        st = Stats()
        check(lib().gmp_batch_sssp(self._h, graphs.size, _ptr(graphs), _ptr(sources), _ptr(out), ctypes.byref(st)))
        return self._split(graphs, out, np.uint32), st

    def batch_ppr(self, graphs, seeds, damping=0.85, iters=30):
        graphs, seeds = _u32(graphs), _u32(seeds)
        total = sum(self.batch_info(int(g))[0] for g in graphs)
        out = np.zeros(total, dtype=np.float32)
        st = Stats()
        check(lib().gmp_batch_ppr(self._h, graphs.size, _ptr(graphs), _ptr(seeds), damping, iters, _ptr(out),
                                  ctypes.byref(st)))
        return self._split(graphs, out, np.float32), st

    def batch_run(self, sssp_graphs, sources, ppr_graphs, seeds, damping=0.85, iters=30):
        """SSSP and PPR instances in one replayed graph (gmp_batch_run); returns
        (distances per SSSP instance, ranks per PPR instance, SSSP stats, PPR stats)."""
        sg, so, pg, se = _u32(sssp_graphs), _u32(sources), _u32(ppr_graphs), _u32(seeds)
        ds = np.zeros(sum(self.batch_info(int(g))[0] for g in sg), dtype=np.uint32)
        rk = np.zeros(sum(self.batch_info(int(g))[0] for g in pg), dtype=np.float32)
        st_s, st_p = Stats(), Stats()
        check(lib().gmp_batch_run(self._h, sg.size, _ptr(sg), _ptr(so), pg.size, _ptr(pg), _ptr(se), damping, iters,
                                  _ptr(ds), _ptr(rk), ctypes.byref(st_s), ctypes.byref(st_p)))
        return self._split(sg, ds, np.uint32), self._split(pg, rk, np.float32), st_s, st_p

    def batch_autotune(self, sssp_graphs, sources, ppr_graphs, seeds, damping=0.85, iters=30):
        sg, so, pg, se = _u32(sssp_graphs), _u32(sources), _u32(ppr_graphs), _u32(seeds)
        r = BatchTuneResult()
        check(lib().gmp_batch_autotune(self._h, sg.size, _ptr(sg), _ptr(so), pg.size, _ptr(pg), _ptr(se), damping,
                                       iters, ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in r._fields_}

    # apps
    def run_app(self, app, msg_init, max_iter, vertex_out=None):
        V = self.info()[0]
        dt = np.float32 if app == APP_PAGERANK else (np.uint32 if app == APP_SSSP_W else np.int32)
        m0 = np.ascontiguousarray(msg_init, dtype=dt)
        if m0.size != V:
            raise ValueError("msg_init must have V entries")
        msg = np.zeros(V, dtype=dt)
        val = np.zeros(V, dtype=dt)
        vo = None if vertex_out is None else np.ascontiguousarray(vertex_out, dtype=np.int32)
        st = Stats()
        check(lib().gmp_run_app(self._h, app, max_iter, _ptr(m0), _ptr(vo), _ptr(msg), _ptr(val),
                                ctypes.byref(st)))
        return msg, val, st

    def pagerank(self, damping=0.85, max_iter=20, tol=0.0, seed_vertex=-1, kahan=True,
                 compressed=False, want_ranks=True, strip_ids=False):
        V = self.info()[0]
        flags = (0 if kahan else PR_NO_KAHAN) | (PR_COMPRESSED if compressed else 0) | (PR_STRIP_IDS if strip_ids else 0)
        p = PrParams(damping, max_iter, tol, seed_vertex, flags, 0)
        out = np.zeros(V, dtype=np.float32) if want_ranks else None
        st = Stats()
        check(lib().gmp_pagerank(self._h, ctypes.byref(p), _ptr(out), ctypes.byref(st)))
        return out, st

    def pagerank_noalloc(self, V, E, d_row_off, d_col_idx, d_out_degree, d_rank_out, damping=0.85, max_iter=20,
                         tol=0.0, seed_vertex=-1, kahan=True):
        """gmp_pagerank_noalloc: the CSR and the rank buffer are device addresses (ints) owned by
        the caller; nothing is allocated."""
        p = PrParams(damping, max_iter, tol, seed_vertex, 0 if kahan else PR_NO_KAHAN, 0)
        st = Stats()
        check(lib().gmp_pagerank_noalloc(self._h, V, E, ctypes.c_void_p(d_row_off), ctypes.c_void_p(d_col_idx),
                                         ctypes.c_void_p(d_out_degree), ctypes.byref(p), ctypes.c_void_p(d_rank_out),
                                         ctypes.byref(st)))
        return st

    def sssp(self, source, method=SSSP_BF, delta=64, unit_weights=False, want_dist=True):
        V = self.info()[0]
        p = SsspParams(method, delta, int(unit_weights), 0)
        out = np.zeros(V, dtype=np.uint32) if want_dist else None
        st = Stats()
        check(lib().gmp_sssp(self._h, source, ctypes.byref(p), _ptr(out), ctypes.byref(st)))
        return out, st

    # learned scheduling
    def sssp_trace(self, source, cap=1 << 16, want_dist=True):
        V = self.info()[0]
        rows = np.zeros(cap, dtype=TRACE_DTYPE)
        n = ctypes.c_uint32()
        out = np.zeros(V, dtype=np.uint32) if want_dist else None
        st = Stats()
        p = SsspParams(SSSP_LEARNED, 0, 0, 0)
        check(lib().gmp_sssp_trace(self._h, source, ctypes.byref(p), _ptr(rows), cap, ctypes.byref(n), _ptr(out),
                                   ctypes.byref(st)))
        return rows[: n.value].copy(), out, st

    def window_trace(self, kind, source=0, cap=1 << 12):
        """gmp_window_trace: min sweeps over a semi-external graph (kind 0 labels, 1 unit, 2
        weighted distances) recording the window predictor's rows."""
        rows = np.zeros(cap, dtype=TRACE_DTYPE)
        n = ctypes.c_uint32()
        st = Stats()
        check(lib().gmp_window_trace(self._h, kind, source, _ptr(rows), cap, ctypes.byref(n), ctypes.byref(st)))
        return rows[: n.value].copy(), st

    def sssp_actions(self):
        """Per-class counts of the bucket epochs the last learned SSSP run chose."""
        c = np.zeros(MLP_CLASSES, dtype=np.uint32)
        check(lib().gmp_sssp_actions(self._h, _ptr(c), MLP_CLASSES))
        return c

    def mlp_init(self, seed):
        check(lib().gmp_mlp_init(self._h, seed))

    def mlp_set(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == MLP_PARAMS
        check(lib().gmp_mlp_set(self._h, _ptr(params)))

    def mlp_get(self):
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        check(lib().gmp_mlp_get(self._h, _ptr(p)))
        return p

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
        Y = np.zeros((X.shape[0], MLP_CLASSES + MLP_SHARDS), dtype=np.float32)
        check(lib().gmp_mlp_forward(self._h, _ptr(X), X.shape[0], _ptr(Y)))
        return Y

    def mlp_loss_grad(self, X, labels, masks):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
        y = np.ascontiguousarray(labels, dtype=np.int32)
        m = np.ascontiguousarray(masks, dtype=np.uint64)
        loss = ctypes.c_float()
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        check(lib().gmp_mlp_loss_grad(self._h, _ptr(X), _ptr(y), _ptr(m), y.size, ctypes.byref(loss), _ptr(g)))
        return loss.value, g

    def mlp_train(self, X, labels, masks, epochs=200, lr=3e-3):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
        y = np.ascontiguousarray(labels, dtype=np.int32)
        m = np.ascontiguousarray(masks, dtype=np.uint64)
        loss = ctypes.c_float()
        check(lib().gmp_mlp_train(self._h, _ptr(X), _ptr(y), _ptr(m), y.size, epochs, lr, ctypes.byref(loss)))
        return loss.value

    def cc(self, want_labels=True):
        V = self.info()[0]
        out = np.zeros(V, dtype=np.uint32) if want_labels else None
        st = Stats()
        check(lib().gmp_cc(self._h, _ptr(out), ctypes.byref(st)))
        return out, st
