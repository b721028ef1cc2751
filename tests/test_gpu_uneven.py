"""Uneven batches (fv_gpu_batch_stage_ex): per-query k in {1, 10, 100} and
per-query selectivity log-uniform in [1e-4, 1] (price < 1000 s), with the
predicates evaluated inside the scan (inline) or through bitmaps; every
query must equal the oracle's exact top-k for its own k."""
import numpy as np
import pytest

from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

import oracle as O

pytestmark = pytest.mark.gpu


def _uneven(n, nq, seed):
    d = synth.gmm128(n, nq, seed=seed)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    rng = np.random.default_rng(seed)
    kq = rng.choice(np.array([1, 10, 100], np.uint32), nq)
    sel = 10.0 ** rng.uniform(-4, 0, nq)
    preds = [E.cmp(0, E.CMP_LT, float(1000.0 * s)) for s in sel]
    post = E.Postings(rows, vals)
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    return d, cols, rows, kq, b


def _check(out, cnt, trunc, d, cols, b, kq, sel=None):
    idx = np.arange(len(kq)) if sel is None else sel
    for k in np.unique(kq[idx]):
        qs = idx[kq[idx] == k]
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"][qs], b.qd[qs], b.nodes,
                                      b.children, b.invals, int(k))
        for j, q in enumerate(qs):
            c = int(cnt[q])
            assert c == int(oc[j]) and int(trunc[q]) == int(ot[j]), (q, k, c, int(oc[j]))
            assert np.array_equal(out[q, :c]["node"], oh[j, :c]["node"]), (q, k)
            assert np.array_equal(out[q, :c]["dist"], oh[j, :c]["dist"]), (q, k)


@pytest.mark.parametrize("inline,qtw,order", [(1, 0, 0), (0, 0, 0), (1, 1, 0), (0, 2, 0),
                                              (1, 4, 0), (1, 1, 3), (0, 4, 3), (1, 2, 2)])
def test_uneven_k_and_selectivity(inline, qtw, order):
    """query tiles per WG: 0 = auto (1 for this 5-tile batch), 1, 2, 4; tile
    order 0 = auto, 2 = sweep, 3 = query-major grid"""
    n, nq = 200_000, 300
    d, cols, rows, kq, b = _uneven(n, nq, 41)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16)
    try:
        bh = g.stage_ex(d["queries"], b, kq)
        try:
            for route in (1, 0):
                g.run_brute(bh, L.BruteOpts(route=route, inline_pred=inline, qtw=qtw,
                                            order=order))
                g.sync()
                out, cnt, flag, rc = g.fetch(bh, nq, int(kq.max()))
                assert rc == 0
                _check(out, cnt, flag, d, cols, b, kq)
        finally:
            g.free_batch(bh)
    finally:
        g.close()


def test_4096_queries_inline():
    """4096 distinct predicates (beyond the bitmap path's 1024) in one
    batch: the inline path serves it; a sample of queries is checked."""
    n, nq = 100_000, 4096
    d, cols, rows, kq, b = _uneven(n, nq, 43)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16)
    try:
        bh = g.stage_ex(d["queries"], b, kq)
        try:
            with pytest.raises(E.GpuError):
                g.run_brute(bh, L.BruteOpts(route=1, inline_pred=0))
            g.run_brute(bh, L.BruteOpts(route=1, inline_pred=1))
            g.sync()
            out, cnt, flag, rc = g.fetch(bh, nq, int(kq.max()))
            assert rc == 0
            sel = np.random.default_rng(1).choice(nq, 256, replace=False)
            _check(out, cnt, flag, d, cols, b, kq, sel)
        finally:
            g.free_batch(bh)
    finally:
        g.close()


@pytest.mark.parametrize("inline,qtw", [(1, 4), (0, 1)])
def test_graph_replay_equals_eager(inline, qtw):
    """FvGpuBruteOpts.graph: the first run is eager, the second captures a
    hipGraph, later runs replay it. Every replay, with other eager work and
    synchronisations between replays, equals the eager result."""
    n, nq = 120_000, 600
    d, cols, rows, kq, b = _uneven(n, nq, 47)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16)
    try:
        bh = g.stage_ex(d["queries"], b, kq)
        other = g.stage_ex(d["queries"][:64], E.concat([E.serialize(None, n)] * 64, n),
                           np.full(64, 10, np.uint32))
        try:
            g.run_brute(bh, L.BruteOpts(route=1, inline_pred=inline, qtw=qtw))
            g.sync()
            ref = g.fetch(bh, nq, int(kq.max()))
            assert ref[3] == 0
            _check(ref[0], ref[1], ref[2], d, cols, b, kq,
                   np.random.default_rng(2).choice(nq, 96, replace=False))
            opts = L.BruteOpts(route=1, inline_pred=inline, qtw=qtw, graph=1)
            for it in range(5):
                g.run_brute(bh, opts)
                g.sync()
                out, cnt, flag, rc = g.fetch(bh, nq, int(kq.max()))
                assert rc == 0, it
                assert np.array_equal(cnt, ref[1]) and np.array_equal(flag, ref[2]), it
                assert np.array_equal(out, ref[0]), it
                g.run_brute(other, L.BruteOpts(route=1))
                g.sync()
        finally:
            g.free_batch(other)
            g.free_batch(bh)
    finally:
        g.close()


def _uneven768(n, nq, seed):
    d = synth.gmm768(n, nq, seed=seed)
    cols = synth.columns(d)
    rng = np.random.default_rng(seed)
    kq = rng.choice(np.array([1, 10, 100], np.uint32), nq)
    sel = 10.0 ** rng.uniform(-4, 0, nq)
    preds = [E.cmp(0, E.CMP_LT, float(1000.0 * s)) for s in sel]
    b = E.concat([E.serialize(p, n) for p in preds], n)
    return d, cols, kq, b


@pytest.mark.parametrize("dim", [128, 768])
def test_autotuned_uneven_batch_and_replay(dim):
    """The uneven workload at both vector shapes (4096 queries at 128 dims,
    1024 at 768; k in {1, 10, 100}, selectivity 10^U(-4, 0) inline): the
    autotuner's chosen launch shape certifies every query, the eager run
    with that shape equals the oracle on a sample of queries (distances
    within 1e-5 relative of FP64 at 768 dims, where sums are not exact),
    and hipGraph replays of the tuned shape equal the eager run."""
    if dim == 128:
        n, nq = 100_000, 4096
        d, cols, rows, kq, b = _uneven(n, nq, 53)
    else:
        n, nq = 50_000, 1024
        d, cols, kq, b = _uneven768(n, nq, 57)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_BF16)
    try:
        bh = g.stage_ex(d["queries"], b, kq)
        try:
            print(f"case tune dim={dim}", flush=True)
            best, us, ntried = g.tune(bh, L.BruteOpts(route=1, inline_pred=1), reps=2)
            assert ntried > 1 and np.isfinite(us) and us > 0
            assert best.inline_pred == 1
            g.run_brute(bh, best)
            g.sync()
            ref = g.fetch(bh, nq, int(kq.max()))
            assert ref[3] == 0
            sel = np.random.default_rng(3).choice(nq, 128, replace=False)
            _check(ref[0], ref[1], ref[2], d, cols, b, kq, sel)
            if dim == 768:
                for q in sel[:16]:
                    for k in range(int(ref[1][q])):
                        f64 = O.l2sq_f64(d["queries"][q], d["base"][int(ref[0][q, k]["node"])])
                        assert abs(float(ref[0][q, k]["dist"]) - f64) <= 1e-5 * max(f64, 1e-30)
            opts = L.BruteOpts(**{f: getattr(best, f) for f, _ in L.BruteOpts._fields_})
            opts.graph = 1
            for it in range(4):
                g.run_brute(bh, opts)
                g.sync()
                out, cnt, flag, rc = g.fetch(bh, nq, int(kq.max()))
                assert rc == 0, it
                assert np.array_equal(cnt, ref[1]) and np.array_equal(out, ref[0]), it
        finally:
            g.free_batch(bh)
    finally:
        g.close()
