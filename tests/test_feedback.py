"""The service-time feedback table (fv_gpu_feedback_update, CPU).

Known answers worked by hand from the reference worker's rule
(worker.c:296-333, arbiter.h:50-60 and 221-229): selectivity buckets at
20%, 5%, 1% and 0.2%; an empty field takes its first sample; per bucket the
sample (T / max weight) x (mean weight in the bucket) enters with weight
n_b / (n_b + 64); batch time and size with weight 1/8; the graph weights
queries by want x 1000 / max(sel, 1), brute force weighs them equally.
"""
import ctypes as C

import numpy as np
import pytest

from paves_amd import _lib as L


def _upd(lib, fb, strategy, sel, want, t):
    sel = np.ascontiguousarray(sel, np.int32)
    w = None if want is None else np.ascontiguousarray(want, np.uint32)
    lib.fv_gpu_feedback_update(C.byref(fb), strategy, len(sel), sel.ctypes.data,
                               None if w is None else w.ctypes.data, t)


def test_buckets():
    lib = L.load()
    for p, b in [(1000, 0), (200, 0), (199, 1), (50, 1), (49, 2), (10, 2), (9, 3), (2, 3),
                 (1, 4), (0, 4)]:
        assert lib.fv_gpu_sel_bucket(p) == b, p


def test_brute_and_graph_known_answers():
    lib = L.load()
    fb = L.Feedback()
    # brute: three queries in buckets 0, 2 and 4, T = 100 us: first samples
    _upd(lib, fb, 0, [500, 30, 1], None, 100.0)
    assert [fb.q_us[0][b] for b in range(5)] == [100.0, 0.0, 100.0, 0.0, 100.0]
    assert fb.batch_us[0] == 100.0 and fb.batch_n[0] == 3.0 and fb.batches[0] == 1
    # second brute batch, two queries in bucket 0, T = 50: weight 2 / 66
    _upd(lib, fb, 0, [500, 600], None, 50.0)
    assert fb.q_us[0][0] == pytest.approx(100.0 + 2.0 / 66.0 * (50.0 - 100.0), rel=1e-15)
    assert fb.q_us[0][2] == 100.0
    assert fb.batch_us[0] == pytest.approx(93.75) and fb.batch_n[0] == pytest.approx(2.875)
    # graph: sel 100 (bucket 1) and 1 (bucket 4), want 80: weights 800 and
    # 80000; samples (1000 / 80000) x 800 = 10 and x 80000 = 1000
    _upd(lib, fb, 1, [100, 1], [80, 80], 1000.0)
    assert fb.q_us[1][1] == pytest.approx(10.0) and fb.q_us[1][4] == pytest.approx(1000.0)
    assert fb.q_us[1][0] == 0.0 and fb.batches[1] == 1
    # sel 0 counts as 1 in the weight
    fb2 = L.Feedback()
    _upd(lib, fb2, 1, [0, 1000], [10, 10], 500.0)
    assert fb2.q_us[1][4] == pytest.approx(500.0)
    assert fb2.q_us[1][0] == pytest.approx(500.0 / 10000.0 * 10.0)
