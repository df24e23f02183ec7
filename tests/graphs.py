"""Small-graph helpers for the tests (NumPy only)."""
import numpy as np
import scipy.sparse as sp


def csr_from_edges(V, src, dst):
    """Destination-row CSR with ascending, deduplicated sources (scipy's COO->CSR rules,
    the same rules GraphMP/tools/format_spark.py:147 relies on)."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    m = sp.csr_matrix((np.ones(src.size, dtype=np.int8), (dst, src)), shape=(V, V))
    m.sum_duplicates()
    m.sort_indices()
    return m.indptr.astype(np.int64), m.indices.astype(np.uint32)


def random_graph(V, E, seed, self_loops=False):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    return csr_from_edges(V, src, dst)


def symmetrize(V, off, idx):
    dst = np.repeat(np.arange(V), np.diff(off))
    src = idx.astype(np.int64)
    return csr_from_edges(V, np.concatenate([src, dst]), np.concatenate([dst, src]))


def transpose(V, off, idx, w=None):
    """Out-CSR (rows = sources), rows ascending, weights carried along."""
    dst = np.repeat(np.arange(V, dtype=np.int64), np.diff(off))
    src = idx.astype(np.int64)
    order = np.lexsort((dst, src))
    ooff = np.zeros(V + 1, dtype=np.int64)
    np.add.at(ooff, src + 1, 1)
    ooff = np.cumsum(ooff)
    return ooff, dst[order].astype(np.uint32), (None if w is None else np.asarray(w)[order])


def scipy_adj(V, off, idx, w=None):
    """Sparse adjacency with rows = sources (A[u, v] = weight of u->v)."""
    dst = np.repeat(np.arange(V, dtype=np.int64), np.diff(off))
    data = np.ones(idx.size) if w is None else np.asarray(w, dtype=np.float64)
    return sp.csr_matrix((data, (idx.astype(np.int64), dst)), shape=(V, V))


def edge_weights(off, idx, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(1, 256, idx.size).astype(np.uint32)


def out_degree(V, idx):
    return np.bincount(idx.astype(np.int64), minlength=V).astype(np.uint32)
