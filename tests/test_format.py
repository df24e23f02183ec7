"""Host-side product logic of libgmp.so that needs no GPU: the .npy container under
cnpy's rules, the shard writer (GraphMP's shard layout), the partitioner, error codes,
and the exported C ABI."""
import os
import re

import numpy as np
import pytest

import graphee_amd as gmp
import oracle
from tests import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    with open(os.path.join(ROOT, "include", "gmp.h")) as f:
        declared = set(re.findall(r"\b(gmp_[a-z0-9_]+)\s*\(", f.read()))
    assert declared == set(gmp.SYMBOLS)
    L = gmp.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.gmp_abi_version() == 4


@pytest.mark.parametrize("dtype", [np.int32, np.uint32])
@pytest.mark.parametrize("n", [0, 1, 5, 1000])
def test_npy_read_accepts_numpy_files(tmp_path, dtype, n):
    a = (np.arange(n) * 7919 % 100003).astype(dtype)
    p = tmp_path / "a.npy"
    np.save(p, a)
    np.testing.assert_array_equal(gmp.npy_read_u32(p), a.astype(np.uint32))


def test_npy_read_version2_header(tmp_path):
    from numpy.lib import format as npf
    a = np.arange(10, dtype=np.int32)
    p = tmp_path / "v2.npy"
    with open(p, "wb") as f:
        npf.write_array(f, a, version=(2, 0))
    np.testing.assert_array_equal(gmp.npy_read_u32(p), a)


def test_npy_write_is_numpy_readable(tmp_path):
    a = np.array([-5, 0, 7, 2**31 - 1], dtype=np.int32)
    p = tmp_path / "w.npy"
    gmp.npy_write_i32(p, a)
    np.testing.assert_array_equal(np.load(p), a)
    with open(p, "rb") as f:
        head = f.read(10)
    assert (10 + head[8] + 256 * head[9]) % 64 == 0


@pytest.mark.parametrize("arr,code", [
    (np.zeros((2, 3), dtype=np.int32), gmp.EFORMAT),   # ndims != 1 (cnpy.h:90)
    (np.zeros(4, dtype=">i4"), gmp.EFORMAT),           # big endian (cnpy.h:100-101)
    (np.zeros(4, dtype=np.float64), gmp.EFORMAT),      # not 4-byte integers
    (np.zeros(4, dtype=np.int16), gmp.EFORMAT),
])
def test_npy_read_rejects_bad_arrays(tmp_path, arr, code):
    p = tmp_path / "bad.npy"
    np.save(p, arr)
    with pytest.raises(gmp.GmpError) as ei:
        gmp.npy_read_u32(p)
    assert ei.value.code == code


def test_npy_read_missing_and_truncated(tmp_path):
    with pytest.raises(gmp.GmpError) as ei:
        gmp.npy_read_u32(tmp_path / "nope.npy")
    assert ei.value.code == gmp.EIO
    p = tmp_path / "t.npy"
    np.save(p, np.arange(100, dtype=np.int32))
    data = p.read_bytes()
    p.write_bytes(data[:-8])
    with pytest.raises(gmp.GmpError) as ei:
        gmp.npy_read_u32(p)
    assert ei.value.code == gmp.EIO
    p.write_bytes(b"not an npy file at all")
    with pytest.raises(gmp.GmpError) as ei:
        gmp.npy_read_u32(p)
    assert ei.value.code == gmp.EFORMAT


def test_partition_bounds_matches_oracle():
    rng = np.random.default_rng(3)
    for V, P in [(1000, 16), (50, 50), (7, 3), (1, 1), (100, 200)]:
        indeg = rng.integers(0, 40, V)
        np.testing.assert_array_equal(gmp.partition_bounds(indeg, P), oracle.partition_bounds(indeg, P))


def test_write_shards_layout(tmp_path):
    """Each shard is int32[edge_num_raw, len(indices), len(indptr), start, end, indices,
    indptr] with rows = destinations of its interval (format_spark.py:137-157)."""
    V = 300
    off, idx = graphs.random_graph(V, 2000, seed=2)
    w = graphs.edge_weights(off, idx, seed=3)
    bounds = oracle.partition_bounds(np.diff(off), 5)
    vout = graphs.out_degree(V, idx).astype(np.int32)
    gmp.write_shards(tmp_path, off, idx, bounds, weights=w, vertex_out=vout)
    np.testing.assert_array_equal(np.load(tmp_path / "vertexout.npy"), vout)
    for p in range(5):
        s, e = int(bounds[p]), int(bounds[p + 1])
        a = np.load(tmp_path / f"{p}.edge.npy")
        nnz = int(off[e] - off[s])
        assert list(a[:5]) == [nnz, nnz, e - s + 1, s, e]
        np.testing.assert_array_equal(a[5:5 + nnz], idx[off[s]:off[e]])
        np.testing.assert_array_equal(a[5 + nnz:], off[s:e + 1] - off[s])
        np.testing.assert_array_equal(np.load(tmp_path / f"{p}.weight.npy"), w[off[s]:off[e]])


def test_write_shards_rejects_bad_bounds(tmp_path):
    off, idx = graphs.random_graph(10, 20, seed=1)
    with pytest.raises(gmp.GmpError) as ei:
        gmp.write_shards(tmp_path, off, idx, np.array([0, 5, 9], dtype=np.uint32))
    assert ei.value.code == gmp.EINVAL


def test_engine_reports_missing_library_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(gmp, "_lib", None)
    monkeypatch.setattr(gmp, "LIB_PATH", str(tmp_path / "libgmp.so"))
    with pytest.raises(ImportError):
        gmp.lib()
