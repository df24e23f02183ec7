"""Host-side logic of the learned schedule (no GPU): training labels from rollout times."""
import numpy as np

from graphee_amd import TRACE_DTYPE, learned


def rows_with(ms):
    r = np.zeros(len(ms), dtype=TRACE_DTYPE)
    r["feat"] = np.arange(len(ms), dtype=np.float32)[:, None]  # distinct rows unless set
    r["ms"] = np.asarray(ms, dtype=np.float32)
    r["best"] = r["ms"].argmin(axis=1)
    return r


def test_label_is_narrowest_near_tie():
    r = rows_with([[24.1, 23.9, 24.0, 24.2, 24.1, 25.4, 24.6, 24.1],   # all within 3%: narrowest
                   [85.9, 127.0, 162.8, 200.9, 213.4, 196.5, 175.8, 177.2],  # a clear winner
                   [3.1, 3.1, 3.2, 3.2, 3.1, 2.4, 2.4, 2.3]])  # 2.3 best; 2.4 within 3%? no (4.3%)
    np.testing.assert_array_equal(learned.labels(r, tol=0.03), [0, 0, 7])
    np.testing.assert_array_equal(learned.labels(r, tol=0.05), [0, 0, 5])
    np.testing.assert_array_equal(learned.labels(r), [0, 0, 5])  # default tolerance 10%
    # tol 0 reproduces the fastest class
    np.testing.assert_array_equal(learned.labels(r, tol=0.0), r["best"])


def test_rows_with_equal_features_share_the_cheapest_class():
    r = rows_with([[10.0, 10.5, 20.0, 30.0, 30.0, 30.0, 30.0, 30.0],    # alone: class 0 (5% gap)
                   [30.0, 10.0, 40.0, 60.0, 60.0, 60.0, 60.0, 60.0],    # alone: class 1
                   [5.0, 9.0, 9.0, 9.0, 9.0, 9.0, 9.0, 4.0]])           # other features
    r["feat"][:2] = 0.25
    r["feat"][2] = 0.5
    # pooled relative times of the first two rows: class 0 -> 1 + 3 = 4, class 1 -> 1.05 + 1 = 2.05
    np.testing.assert_array_equal(learned.labels(r), [1, 1, 7])


def test_labels_of_empty_trace():
    assert learned.labels(np.zeros(0, dtype=TRACE_DTYPE)).size == 0
