"""CPU oracle for the gfx950 graph engine.  TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as
the checker (never as the thing measured or shipped).  gmp_oracle.c restates GraphMP's
per-app update rules and engine loop; rmat_ref.py restates the synthetic-input
definition.  Build with `make -C oracle` (or __graft_entry__.build()).
"""
import ctypes
import os
import subprocess

import numpy as np

from . import rmat_ref  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libgmp_oracle.so")
_lib = None

APP_PAGERANK = 1
APP_SSSP = 2
APP_CC = 3
APP_SSSP_W = 4

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32


class BaselineStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("seconds", ctypes.c_double), ("edges", ctypes.c_double), ("shards", ctypes.c_double)]


class BloomShape(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("m", ctypes.c_uint64), ("salt", ctypes.c_uint32 * 16)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.orc_partition_bounds.argtypes = [_P, _I64, _I32, _P]
        L.orc_app_sync.argtypes = [_I32, _I64, _P, _P, _P, _P, _I32, _P, _P, _P, _P]
        L.orc_pr_f64.argtypes = [_I64, _P, _P, _P, ctypes.c_double, _I32, ctypes.c_double,
                                 _I64, _P, _P, _P, _P, _I32]
        L.orc_sssp_dijkstra.argtypes = [_I64, _P, _P, _P, ctypes.c_uint32, _P]
        L.orc_cc_unionfind.argtypes = [_I64, _P, _P, _P]
        L.orc_graphmp_baseline.argtypes = [_I32, _I64, _I32, _P, _P, _P, _P, _P, ctypes.c_float,
                                           _I32, _I32, _I32, _P, _P]
        L.orc_bf_params.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, _P]
        L.orc_bf_probe.argtypes = [_P, _P, _I64, _P, _I64, _P]
        L.orc_pr_f64_step_residual.argtypes = [_I64, _P, _P, _P, ctypes.c_double, _P, _I32, _P, _P]
        L.orc_mlp_loss_grad.argtypes = [_P, _P, _P, _P, _I64, _P, _P]
        L.orc_pr_f64_refine.argtypes = [_I64, _P, _P, _P, ctypes.c_double, _P, _I32, _I32, ctypes.c_double,
                                        _P, _P, _P]
        for f in ("orc_partition_bounds", "orc_app_sync", "orc_pr_f64", "orc_sssp_dijkstra",
                  "orc_cc_unionfind", "orc_graphmp_baseline", "orc_pr_f64_step_residual",
                  "orc_mlp_loss_grad", "orc_pr_f64_refine", "orc_bf_params", "orc_bf_probe"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def partition_bounds(indeg, P):
    indeg = np.ascontiguousarray(indeg, dtype=np.int64)
    b = np.zeros(P + 1, dtype=np.int64)
    _check(lib().orc_partition_bounds(_ptr(indeg), indeg.size, P, _ptr(b)), "partition")
    return b


def app_sync(app, off, idx, msg_init, max_iter, w=None, vout=None):
    """GraphMP synchronous run(); returns (msg, value, iterations, changed_per_iter)."""
    V = off.size - 1
    off = np.ascontiguousarray(off, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    dt = np.float32 if app == APP_PAGERANK else (np.uint32 if app == APP_SSSP_W else np.int32)
    msg = np.array(msg_init, dtype=dt, copy=True)
    value = np.zeros(V, dtype=dt)
    it = ctypes.c_int32(0)
    changed = np.zeros(max(max_iter, 1), dtype=np.int64)
    wv = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
    vo = None if vout is None else np.ascontiguousarray(vout, dtype=np.int32)
    _check(lib().orc_app_sync(app, V, _ptr(off), _ptr(idx), _ptr(wv), _ptr(vo), max_iter,
                              _ptr(msg), _ptr(value), ctypes.byref(it), _ptr(changed)), "app_sync")
    return msg, value, it.value, changed[: it.value]


def pr_f64(off, idx, outdeg, d=0.85, max_iter=100, tol=0.0, seed=-1, x0=None, threads=1):
    V = off.size - 1
    off = np.ascontiguousarray(off, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    od = np.ascontiguousarray(outdeg, dtype=np.uint32)
    x = np.zeros(V, dtype=np.float64)
    x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    it = ctypes.c_int32(0)
    r = ctypes.c_double(0)
    _check(lib().orc_pr_f64(V, _ptr(off), _ptr(idx), _ptr(od), d, max_iter, tol, seed,
                            _ptr(x0a), _ptr(x), ctypes.byref(it), ctypes.byref(r), threads), "pr_f64")
    return x, it.value, r.value


def dijkstra(ooff, oidx, ow, src):
    V = ooff.size - 1
    ooff = np.ascontiguousarray(ooff, dtype=np.int64)
    oidx = np.ascontiguousarray(oidx, dtype=np.uint32)
    owv = None if ow is None else np.ascontiguousarray(ow, dtype=np.uint32)
    dist = np.zeros(V, dtype=np.uint32)
    _check(lib().orc_sssp_dijkstra(V, _ptr(ooff), _ptr(oidx), _ptr(owv), src, _ptr(dist)),
           "dijkstra")
    return dist


def cc_unionfind(off, idx):
    V = off.size - 1
    off = np.ascontiguousarray(off, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    lab = np.zeros(V, dtype=np.uint32)
    _check(lib().orc_cc_unionfind(V, _ptr(off), _ptr(idx), _ptr(lab)), "unionfind")
    return lab


def graphmp_baseline(app, off, idx, bounds, msg_init, max_iter, w=None, outdeg=None,
                     damping=0.85, threads=0, bloom=False):
    """Timed CPU GraphMP engine; returns (msg, BaselineStats).  bloom: sparse steps activate
    shards by GraphMP's per-shard bloom filters instead of exact source sets."""
    V = off.size - 1
    off = np.ascontiguousarray(off, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    bounds = np.ascontiguousarray(bounds, dtype=np.int64)
    dt = np.float32 if app == APP_PAGERANK else np.uint32
    msg = np.array(msg_init, dtype=dt, copy=True)
    wv = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
    od = None if outdeg is None else np.ascontiguousarray(outdeg, dtype=np.uint32)
    st = BaselineStats()
    _check(lib().orc_graphmp_baseline(app, V, bounds.size - 1, _ptr(bounds), _ptr(off), _ptr(idx),
                                      _ptr(wv), _ptr(od), damping, max_iter, threads, int(bool(bloom)),
                                      _ptr(msg), ctypes.byref(st)), "baseline")
    return msg, st


def bloom_shape(n=1000000, p=0.01, seed=0xA5A5A5A5):
    """GraphMP's bloom-filter sizing and salts for n projected elements at false-positive
    probability p (GraphPS.h:99-108 uses the defaults)."""
    f = BloomShape()
    _check(lib().orc_bf_params(n, p, seed, ctypes.byref(f)), "bloom parameters")
    return f


def bloom_probe(shape, keys, probes):
    """Inserts keys into one filter of `shape` and returns contains() for every probe."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    probes = np.ascontiguousarray(probes, dtype=np.uint32)
    hit = np.zeros(probes.size, dtype=np.uint8)
    _check(lib().orc_bf_probe(ctypes.byref(shape), _ptr(keys), keys.size, _ptr(probes), probes.size, _ptr(hit)),
           "bloom probe")
    return hit.astype(bool)


MLP_PARAMS = 16 * 64 + 64 + 64 * 72 + 72


def mlp_loss_grad(params, X, labels, masks, want_grad=True):
    """FP64 loss and gradient of the scheduling MLP (same layout as the engine's)."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(labels, dtype=np.int32)
    m = np.ascontiguousarray(masks, dtype=np.uint64)
    loss = ctypes.c_double(0)
    g = np.zeros(MLP_PARAMS, dtype=np.float64) if want_grad else None
    _check(lib().orc_mlp_loss_grad(_ptr(P), _ptr(X), _ptr(y), _ptr(m), y.size, ctypes.byref(loss), _ptr(g)),
           "mlp_loss_grad")
    return loss.value, g


def pr_step_residual(off, idx, outdeg, x, d=0.85, threads=0):
    """(||T(x)-x||_1, ||T(x)-x||_inf) of one FP64 step from the FP32 ranks x; the L1 distance
    to the exact FP64 fixed point is at most the first value / (1 - d)."""
    V = off.size - 1
    off = np.ascontiguousarray(off, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    od = np.ascontiguousarray(outdeg, dtype=np.uint32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    l1 = ctypes.c_double(0)
    li = ctypes.c_double(0)
    _check(lib().orc_pr_f64_step_residual(V, _ptr(off), _ptr(idx), _ptr(od), d, _ptr(x), threads,
                                          ctypes.byref(l1), ctypes.byref(li)), "step_residual")
    return l1.value, li.value


def pr_f64_error(off, idx, outdeg, x, d=0.85, threads=0, max_iter=60, bound=1e-11):
    """Error of the FP32 ranks x against the FP64 fixed point of the standard formula: FP64
    power iterations from x until the reference's own a-posteriori L1 bound (last change *
    d/(1-d)) is below `bound`.  Returns a dict with the L1 and L-inf distances of x to that
    reference, the reference's bound (each distance is exact to within it) and the steps run."""
    V = off.size - 1
    off = np.ascontiguousarray(off, dtype=np.int64)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    od = np.ascontiguousarray(outdeg, dtype=np.uint32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    ref = np.zeros(V, dtype=np.float64)
    it = ctypes.c_int32(0)
    last = ctypes.c_double(0)
    _check(lib().orc_pr_f64_refine(V, _ptr(off), _ptr(idx), _ptr(od), d, _ptr(x), threads, max_iter, bound,
                                   _ptr(ref), ctypes.byref(it), ctypes.byref(last)), "pr_f64_refine")
    err = np.abs(x.astype(np.float64) - ref)
    return {"l1_error": float(err.sum()), "linf_error": float(err.max()) if V else 0.0,
            "reference_l1_bound": last.value * d / (1.0 - d), "reference_steps": it.value}
