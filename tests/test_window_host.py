"""The host varint packer of semi-external graphs (window.hip: pack_host / unpack_host), run on
the CPU through a small harness linked against libgmp.so, against an independent restatement of
the stream format the device decoder reads (varint.hpp; the device packer compress.hip writes the
same): 64-edge groups over the global edge index, a restart (absolute id) at a group's first
edge, a row's first edge and wherever ids do not ascend, otherwise id - previous - 1, each value
LEB128; per group its first byte offset and the bit mask of its restarts."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "graphee_amd")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libgmp.so")):
        pytest.fail("libgmp.so is not built (run __graft_entry__.build())")
    exe = str(tmp_path_factory.mktemp("pack") / "pack_host_main")
    subprocess.check_call(["g++", "-O1", "-std=c++17", os.path.join(ROOT, "tests", "host", "pack_host_main.cpp"),
                           f"-L{LIBDIR}", "-lgmp", f"-Wl,-rpath,{LIBDIR}", "-o", exe])
    return exe


def restate(off, idx):
    """gbyte, gmask, bytes of the format, written edge by edge."""
    E = idx.size
    G = (E + 63) // 64
    row_start = np.zeros(E, bool)
    nz = off[:-1][off[:-1] < off[1:]]
    row_start[nz] = True
    out, gbyte, gmask = bytearray(), [], []
    for g in range(G):
        gbyte.append(len(out))
        m = 0
        for lane in range(64):
            e = 64 * g + lane
            if e >= E:
                break
            src, prev = int(idx[e]), (int(idx[e - 1]) if lane else 0)
            restart = lane == 0 or row_start[e] or src <= prev
            v = src if restart else src - prev - 1
            if restart:
                m |= 1 << lane
            while True:
                b = v & 0x7F
                v >>= 7
                out.append(b | (0x80 if v else 0))
                if not v:
                    break
        gmask.append(m)
    gbyte.append(len(out))
    return np.array(gbyte, np.uint64), np.array(gmask, np.uint64), bytes(out)


def run(harness, tmp_path, off, idx, e0, e1):
    V = off.size - 1
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.uint32(V).tobytes() + np.uint64(idx.size).tobytes())
        f.write(off.astype(np.uint64).tobytes() + idx.astype(np.uint32).tobytes())
        f.write(np.uint64(e0).tobytes() + np.uint64(e1).tobytes())
    subprocess.check_call([harness, str(inp), str(outp)])
    raw = outp.read_bytes()
    G = (idx.size + 63) // 64
    nb = int(np.frombuffer(raw, np.uint64, 1)[0])
    p = 8
    gbyte = np.frombuffer(raw, np.uint64, G + 1, p)
    p += 8 * (G + 1)
    gmask = np.frombuffer(raw, np.uint64, G, p)
    p += 8 * G
    data = raw[p:p + nb]
    p += nb
    ids = np.frombuffer(raw, np.uint32, e1 - e0, p)
    return gbyte, gmask, data, ids


def random_csr(rng, V, hub_every, sort_rows, big_ids):
    deg = np.where(rng.random(V) < 1.0 / hub_every, rng.integers(100, 400, V), rng.integers(0, 6, V))
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    hi = 0xFFFFFFFE if big_ids else V
    idx = rng.integers(0, hi, int(off[-1]), dtype=np.uint64).astype(np.uint32)
    if sort_rows:
        for r in range(V):
            idx[off[r]:off[r + 1]].sort()
    return off, idx


@pytest.mark.parametrize("seed,V,hub_every,sort_rows,big_ids",
                         [(1, 300, 20, True, False), (2, 1000, 50, False, False), (3, 57, 5, True, True),
                          (4, 2000, 100, True, False), (5, 1, 1, False, True)])
def test_host_packer_writes_the_decoder_format(harness, tmp_path, seed, V, hub_every, sort_rows, big_ids):
    rng = np.random.default_rng(seed)
    off, idx = random_csr(rng, V, hub_every, sort_rows, big_ids)
    E = idx.size
    e0 = int(rng.integers(0, max(E, 1)))
    e1 = int(rng.integers(e0, E + 1))
    gbyte, gmask, data, ids = run(harness, tmp_path, off, idx, e0, e1)
    rg, rm, rb = restate(off, idx)
    np.testing.assert_array_equal(gbyte, rg)
    np.testing.assert_array_equal(gmask, rm)
    assert data == rb
    np.testing.assert_array_equal(ids, idx[e0:e1])


def test_host_packer_edge_cases(harness, tmp_path):
    # empty rows between rows, a row that crosses groups, a descending row, the largest id
    off = np.array([0, 0, 70, 70, 75, 75, 200], np.uint64)
    idx = np.concatenate([np.arange(70) * 3, [9, 8, 7, 6, 5], np.full(125, 0xFFFFFFFE)]).astype(np.uint32)
    gbyte, gmask, data, ids = run(harness, tmp_path, off, idx, 0, idx.size)
    rg, rm, rb = restate(off, idx)
    np.testing.assert_array_equal(gbyte, rg)
    np.testing.assert_array_equal(gmask, rm)
    assert data == rb
    np.testing.assert_array_equal(ids, idx)
    assert int(gmask[0]) & 1 and (int(gmask[1]) >> (70 - 64)) & 1  # row starts restart
