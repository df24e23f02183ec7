"""NumPy restatement of the synthetic-input definition.  TEST INFRASTRUCTURE ONLY.

The reference (cap-ntu/Graphee) ships no generator -- its inputs were TSV edge lists
converted by GraphMP/tools/format_spark.py.  The engine generates R-MAT graphs on the
GPU; this module states the same definition in vectorised NumPy so the tests can
check the device-built graph edge for edge.

Definition (all arithmetic on uint64, wrapping):
  GAMMA = 0x9E3779B97F4A7C15, mix64 = the SplitMix64 finaliser
  base      = mix64(seed + GAMMA)
  z_k(i)    = mix64((base ^ i) + (k + 1) * GAMMA)          edge sample i, draw k
  level l   = 0 .. scale-1, most significant bit first, uses the 32-bit word
              r = (z_{l // 2}(i) >> (32 * (l % 2))) & 0xFFFFFFFF
  T1, T2, T3 = floor(a * 2^32), floor((a + b) * 2^32), floor((a + b + c) * 2^32)
  r < T1 -> (src bit, dst bit) = (0, 0); r < T2 -> (0, 1); r < T3 -> (1, 0); else (1, 1)
  relabel   = perm(x) with keys K_j = mix64(seed + 1 + (j + 1) * GAMMA), j = 0..3:
              x ^= K_j & mask; x = (x * ((K_j >> 32) | 1)) & mask; x ^= x >> ((scale + 1) // 2)
  self-loops (src == dst) are dropped, duplicates removed, undirected graphs add the
  reverse of every edge before deduplication
  weight(src, dst) = 1 + mix64((seed + 2) ^ (src << 32) ^ dst) % 255
  source t  = mix64(seed + 3 + (t + 1) * GAMMA) % V, first with out-degree > 0
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_MASK32 = np.uint64(0xFFFFFFFF)


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _u64(x):
    return np.uint64(int(x) & 0xFFFFFFFFFFFFFFFF)


def thresholds(a, b, c):
    out = []
    for x in (a, a + b, a + b + c):
        t = int(np.floor(np.float64(x) * 4294967296.0))
        out.append(min(max(t, 0), 0xFFFFFFFF))
    return [np.uint64(t) for t in out]


def perm(x, scale, seed):
    """Seeded bijection of [0, 2**scale)."""
    x = np.asarray(x, dtype=np.uint64).copy()
    mask = _u64((1 << scale) - 1)
    h = np.uint64((scale + 1) // 2)
    with np.errstate(over="ignore"):
        for j in range(4):
            k = mix64(_u64(seed + 1) + _u64(j + 1) * GAMMA)
            x ^= k & mask
            x = (x * ((k >> np.uint64(32)) | np.uint64(1))) & mask
            x ^= x >> h
    return x


def raw_edges(scale, edge_factor, seed, a=0.57, b=0.19, c=0.19, lo=0, hi=None, permute=True):
    """(src, dst) of samples lo..hi-1, self-loops kept, as uint64 arrays; relabelled
    unless permute=False."""
    m = edge_factor << scale
    hi = m if hi is None else hi
    i = np.arange(lo, hi, dtype=np.uint64)
    t1, t2, t3 = thresholds(a, b, c)
    with np.errstate(over="ignore"):
        base = mix64(_u64(seed) + GAMMA)
        st = base ^ i
        src = np.zeros_like(i)
        dst = np.zeros_like(i)
        for k in range((scale + 1) // 2):
            z = mix64(st + _u64(k + 1) * GAMMA)
            for half in range(2):
                lvl = 2 * k + half
                if lvl >= scale:
                    break
                r = (z >> np.uint64(32 * half)) & _MASK32
                sb = (r >= t2).astype(np.uint64)
                db = ((r >= t1) & (r < t2)) | (r >= t3)
                src = (src << np.uint64(1)) | sb
                dst = (dst << np.uint64(1)) | db.astype(np.uint64)
    if not permute:
        return src, dst
    return perm(src, scale, seed), perm(dst, scale, seed)


def weights(src, dst, seed):
    src = np.asarray(src, dtype=np.uint64)
    dst = np.asarray(dst, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = mix64(_u64(seed + 2) ^ (src << np.uint64(32)) ^ dst)
    return (np.uint64(1) + h % np.uint64(255)).astype(np.uint32)


def in_csr(scale, edge_factor, seed, symmetric=False, a=0.57, b=0.19, c=0.19):
    """Destination-row CSR (int64 offsets, uint32 ascending sources), deduplicated."""
    V = 1 << scale
    src, dst = raw_edges(scale, edge_factor, seed, a, b, c)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    if symmetric:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
    key = np.unique((dst << np.uint64(32)) | src)
    rows = (key >> np.uint64(32)).astype(np.int64)
    idx = (key & _MASK32).astype(np.uint32)
    off = np.zeros(V + 1, dtype=np.int64)
    np.add.at(off, rows + 1, 1)
    off = np.cumsum(off)
    return V, off, idx


def transpose(V, off, idx, w=None):
    """Out-CSR (rows = sources) of an in-CSR, rows ascending, with weights carried."""
    dst = np.repeat(np.arange(V, dtype=np.uint64), np.diff(off))
    src = idx.astype(np.uint64)
    order = np.lexsort((dst, src))
    ooff = np.zeros(V + 1, dtype=np.int64)
    np.add.at(ooff, src.astype(np.int64) + 1, 1)
    ooff = np.cumsum(ooff)
    oidx = dst[order].astype(np.uint32)
    ow = None if w is None else np.asarray(w)[order]
    return ooff, oidx, ow


def pick_source(V, outdeg, seed, salt=3, max_tries=1 << 20):
    for t in range(max_tries):
        with np.errstate(over="ignore"):
            v = int(mix64(_u64(seed + salt) + _u64(t + 1) * GAMMA) % np.uint64(V))
        if outdeg[v] > 0:
            return v
    raise ValueError("no vertex with out-degree > 0")
