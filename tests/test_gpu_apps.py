"""GraphMP-style C++ apps (graphee_amd/apps, built on include/gmp_graphps.hpp) run on shard
directories in GraphMP's layout and must reproduce the reference's synchronous rules bit for
bit (checked against the oracle's restatement)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import graphee_amd as gmp
import oracle

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(gmp.__file__), "bin")


def _shards(engine, tmp_path, scale, P):
    engine.generate_rmat(scale, 16, 2000 + scale, num_shards=P)
    g = engine.download()
    gmp.write_shards(tmp_path, g["off"], g["idx"], g["bounds"], vertex_out=g["outdeg"].astype(np.int32))
    return g


def _run(app, *args):
    cmd = [os.path.join(BIN, f"gmp_{app}"), *map(str, args)]
    print("run", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout + r.stderr)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_pagerank_app_bit_exact(engine, tmp_path):
    g = _shards(engine, tmp_path, 11, 6)
    V = g["outdeg"].size
    out = tmp_path / "pr.npy"
    _run("pagerank", str(tmp_path) + "/", V, 6, 200, out)
    got = np.load(out).view(np.float32)
    msg0 = np.full(V, 1.0 / V, dtype=np.float32)
    want, _, it, _ = oracle.app_sync(oracle.APP_PAGERANK, g["off"], g["idx"], msg0, 200,
                                     vout=g["outdeg"].astype(np.int32))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("app,oapp", [("sssp", oracle.APP_SSSP), ("cc", oracle.APP_CC)])
def test_min_apps_bit_exact(engine, tmp_path, app, oapp):
    g = _shards(engine, tmp_path, 12, 5)
    V = g["outdeg"].size
    out = tmp_path / f"{app}.npy"
    if app == "sssp":
        src = int(np.argmax(g["outdeg"]))
        _run(app, tmp_path, V, 5, 1000, out, src)
        msg0 = np.full(V, 10000, dtype=np.int32)
        msg0[src] = 0
    else:
        _run(app, tmp_path, V, 5, 1000, out)
        msg0 = np.arange(V, dtype=np.int32)
    want, _, _, _ = oracle.app_sync(oapp, g["off"], g["idx"], msg0, 1000)
    np.testing.assert_array_equal(np.load(out), want)


def test_app_reports_missing_shards(tmp_path):
    r = subprocess.run([os.path.join(BIN, "gmp_cc"), str(tmp_path), "10", "2", "5"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "error 2" in r.stderr  # GMP_EIO
