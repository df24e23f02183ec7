"""A graph past 2^32 edges resident in HBM: R-MAT s28 ef17 (4.5e9 edges).  The source-strip
layout counts each strip's edges from a 64-bit base, so it runs such graphs; its PageRank must
equal the row gather's (4-byte ids, 64-bit row offsets) to rounding, and stay within 1e-6 L1 of
the FP64 fixed point by a one-step bound.  Its packed codes (compressed ids) and its 4-byte source
ids give the same ranks bit for bit."""
import numpy as np
import pytest

import graphee_amd as gmp
import oracle

pytestmark = pytest.mark.gpu


def test_strips_past_2e32_edges():
    print("R-MAT s28 ef17", flush=True)
    with gmp.Engine(0) as e:
        e.generate_rmat(28, 17, 7028, num_shards=16)
        V, E, _, _ = e.info()
        assert E > 2**32
        r1, st1 = e.pagerank(max_iter=100, tol=1e-7)
        assert st1.index_bytes == 2 * E  # the strip layout ran
        rc, stc = e.pagerank(max_iter=100, tol=1e-7, compressed=True)
        assert 0 < stc.index_bytes < 2 * E  # the packed codes ran
        assert stc.iterations == st1.iterations
        np.testing.assert_array_equal(rc.view(np.uint32), r1.view(np.uint32))
        del rc
        ri, sti = e.pagerank(max_iter=100, tol=1e-7, strip_ids=True)
        assert sti.index_bytes == 4 * E  # the 4-byte source ids ran
        np.testing.assert_array_equal(ri.view(np.uint32), r1.view(np.uint32))
        del ri
        e.set_option(gmp.OPT_STRIP_LOG, 0)
        r0, st0 = e.pagerank(max_iter=st1.iterations, tol=0.0)
        assert st0.index_bytes == 4 * E
        assert np.abs(r1.astype(np.float64) - r0).sum() < 1e-7
        g = e.download()
        l1, _ = oracle.pr_step_residual(g["off"], g["idx"], g["outdeg"], r1, threads=16)
        assert l1 / 0.15 < 1e-6
