"""Text edge lists -> GraphMP shards (gmp_convert_edge_list, host only) against scipy's CSR
rules and GraphMP's converter rules (GraphMP/tools/format_spark.py: raw in-edge counts for the
partitioner and the header, deduplicated ascending sources, self-loops kept)."""
import numpy as np
import pytest

import graphee_amd as gmp
from tests import graphs


def _write(path, lines):
    path.write_text("".join(lines))
    return str(path)


def _shards(dir_, P):
    out = []
    for p in range(P):
        a = np.load(dir_ / f"{p}.edge.npy")
        raw, nnz, nrow, start, end = (int(x) for x in a[:5])
        idx = a[5:5 + nnz].astype(np.uint32)
        ptr = a[5 + nnz:5 + nnz + nrow].astype(np.int64)
        out.append(dict(raw=raw, nnz=nnz, start=start, end=end, idx=idx, ptr=ptr))
    return out


def _join(shards, V):
    """The shards' rows as one CSR over all V destinations."""
    off = np.zeros(V + 1, np.int64)
    idx, base = [], 0
    for s in shards:
        off[s["start"] + 1:s["end"] + 1] = s["ptr"][1:] + base
        base += s["nnz"]
        idx.append(s["idx"])
    return off, np.concatenate(idx)


def test_convert_small_list_follows_the_converter_rules(tmp_path):
    lines = ["# a comment\n", "% another\n", "\n", "0 1\n", "2\t1\n", "  3   1  \n", "1 1\n",   # self-loop kept
             "0 1\n",                                                                      # duplicate
             "4 2\n", "0 2\n", "3 4\r\n", "4 0"]                                           # no final newline
    p = _write(tmp_path / "g.tsv", lines)
    d = tmp_path / "shards"
    d.mkdir()
    V = gmp.convert_edge_list(p, d, V=0, P=2)
    assert V == 5
    src = np.array([0, 2, 3, 1, 0, 4, 0, 3, 4])
    dst = np.array([1, 1, 1, 1, 1, 2, 2, 4, 0])
    off, idx = graphs.csr_from_edges(V, src, dst)
    sh = _shards(d, 2)
    # raw in-edge counts: v0 1, v1 5 (one duplicate), v2 2, v3 0, v4 1 -> 9 lines, 4 per
    # partition (Python 2 floor) -> boundary 1 = first prefix index reaching 4, i.e. 2
    assert [(s["start"], s["end"]) for s in sh] == [(0, 2), (2, 5)]
    assert [s["raw"] for s in sh] == [6, 3]
    assert [s["nnz"] for s in sh] == [5, 3]
    got_off, got_idx = _join(sh, V)
    np.testing.assert_array_equal(got_off, off)
    np.testing.assert_array_equal(got_idx, idx)
    np.testing.assert_array_equal(np.load(d / "vertexout.npy"), graphs.out_degree(V, idx))


def test_convert_weights_keep_the_least_and_symmetrize(tmp_path):
    p = _write(tmp_path / "w.tsv", ["0 1 7\n", "0 1 3\n", "1 2 5\n", "2 2 9\n"])
    d = tmp_path / "s"
    d.mkdir()
    V = gmp.convert_edge_list(p, d, V=4, P=1, weighted=True, symmetrize=True)
    assert V == 4
    sh = _shards(d, 1)[0]
    w = np.load(d / "0.weight.npy")
    # rows: 0 <- {1 (3)}, 1 <- {0 (3), 2 (5)}, 2 <- {1 (5), 2 (9)}, 3 <- {}
    np.testing.assert_array_equal(sh["ptr"], [0, 1, 3, 5, 5])
    np.testing.assert_array_equal(sh["idx"], [1, 0, 2, 1, 2])
    np.testing.assert_array_equal(w, [3, 3, 5, 5, 9])
    assert sh["raw"] == 8


def test_convert_many_lines_in_parallel_chunks(tmp_path):
    rng = np.random.default_rng(4)
    V, E = 3000, 200000
    src, dst = rng.integers(0, V, E), rng.integers(0, V, E)
    text = "".join(f"{a}\t{b}\n" if i % 3 else f"{a} {b}\n" for i, (a, b) in enumerate(zip(src, dst)))
    p = tmp_path / "big.tsv"
    p.write_text(text)
    d = tmp_path / "b"
    d.mkdir()
    assert gmp.convert_edge_list(p, d, V=V, P=7) == V
    sh = _shards(d, 7)
    off, idx = graphs.csr_from_edges(V, src, dst)
    got_off, got_idx = _join(sh, V)
    np.testing.assert_array_equal(got_off, off)
    np.testing.assert_array_equal(got_idx, idx)
    assert sum(s["raw"] for s in sh) == E
    raw = np.bincount(dst, minlength=V).astype(np.uint64)
    np.testing.assert_array_equal([s["start"] for s in sh] + [V], gmp.partition_bounds(raw, 7))


@pytest.mark.parametrize("text,code", [("0 1\n1 x\n", gmp.EFORMAT), ("0 1\n2\n", gmp.EFORMAT),
                                        ("0 1 2 3\n", gmp.EFORMAT), ("0 99999999999\n", gmp.EFORMAT),
                                        ("0 5\n", gmp.EINVAL)])
def test_convert_rejects_bad_input(tmp_path, text, code):
    p = _write(tmp_path / "bad.tsv", [text])
    d = tmp_path / "o"
    d.mkdir()
    with pytest.raises(gmp.GmpError) as ei:
        gmp.convert_edge_list(p, d, V=5, P=1)
    assert ei.value.code == code


def test_convert_missing_file_is_an_error(tmp_path):
    with pytest.raises(gmp.GmpError) as ei:
        gmp.convert_edge_list(tmp_path / "none.tsv", tmp_path, P=1)
    assert ei.value.code == gmp.EIO
