"""Regression tests for fixed defects of the batch API.

* A captured brute run (FvGpuBruteOpts.graph) must not replay into buffers
  that a graph search on the same batch reallocated: stage, brute through
  the hipGraph twice (eager, capture), a graph search with ef > topb (which
  grows the output buffer and its stride), then another graph-mode brute
  run; its fetched results equal the oracle.
* Batches with more distinct predicates than the bitmap path holds (1024)
  return a clean error from every entry point that builds or reads
  predicate bitmaps (graph search, features, routed runs), never a launch
  on a missing bitmap.
* A routed (learned) run with k above the graph search's result capacity
  (FV_GPU_MAX_EF) sends every query to brute force (mode 0) or fails
  cleanly (graph-only modes).
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import hnsw, learned, synth

pytestmark = pytest.mark.gpu

N = 30_000


@pytest.fixture(scope="module")
def world():
    d = synth.gmm128(N, 1100, seed=808)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    f = hnsw.build(d["base"], M=8, ef_construction=48, threads=8)
    ent = hnsw.entries(f, d["base"], d["queries"])
    g = E.GpuIndex(d["base"], cols, rows, l0=f["l0"], precision=L.FV_GPU_PREC_BF16, device=0)
    yield {"d": d, "cols": cols, "post": E.Postings(rows, vals), "f": f, "ent": ent, "g": g}
    g.close()


def test_capture_survives_graph_run_on_same_batch(world):
    w = world
    d, g = w["d"], w["g"]
    nq, topb = 64, 10
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, N, w["post"]) for p in preds], N, entries=w["ent"][:nq])
    qs = d["queries"][:nq]
    oh, oc, ot, _ = O.brute_batch(d["base"], w["cols"], qs, b.qd, b.nodes, b.children, b.invals,
                                  topb)
    bh = g.stage(qs, b, topb)
    try:
        opts = L.BruteOpts(route=1, graph=1)
        for _ in range(2):
            g.run_brute(bh, opts)
        g.sync()
        g.run_graph(bh, 300, 200000)   # grows the output buffer: stride 300
        g.sync()
        for it in range(3):
            g.run_brute(bh, opts)
            g.sync()
            out, cnt, flag, rc = g.fetch(bh, nq, topb)
            assert rc == 0, it
            assert np.array_equal(cnt, oc) and np.array_equal(flag, ot), it
            for q in range(nq):
                c = int(cnt[q])
                assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), (it, q)
                assert np.array_equal(out[q, :c]["dist"], oh[q, :c]["dist"]), (it, q)
    finally:
        g.free_batch(bh)


def test_too_many_predicates_rejected(world):
    w = world
    d, g = w["d"], w["g"]
    nq = 1100
    preds = [E.cmp(0, E.CMP_LT, 10.0 + i) for i in range(nq)]   # 1100 distinct programs
    b = E.concat([E.serialize(p, N) for p in preds], N, entries=w["ent"][:nq])
    bh = g.stage_ex(d["queries"][:nq], b, np.full(nq, 10, np.uint32))
    m = learned.Model(128, d["centers"].astype(np.float32))
    try:
        with pytest.raises(E.GpuError, match="distinct predicates"):
            g.run_graph(bh, 32, 200000)
        with pytest.raises(E.GpuError, match="distinct predicates"):
            g.run_brute(bh, L.BruteOpts(route=1, inline_pred=0))
        with pytest.raises(E.GpuError, match="distinct predicates"):
            m.batch_features(g, bh, nq)
        with pytest.raises(E.GpuError, match="distinct predicates"):
            g.run_learned(bh, m, L.LearnedOpts(k=10, mode=0))
        # the inline path still serves the batch
        g.run_brute(bh, L.BruteOpts(route=1, inline_pred=1))
        g.sync()
        out, cnt, flag, rc = g.fetch(bh, nq, 10)
        assert rc == 0
        sel = np.arange(0, nq, 97)
        oh, oc, ot, _ = O.brute_batch(d["base"], w["cols"], d["queries"][sel], b.qd[sel], b.nodes,
                                      b.children, b.invals, 10)
        for j, q in enumerate(sel):
            c = int(cnt[q])
            assert c == int(oc[j])
            assert np.array_equal(out[q, :c]["node"], oh[j, :c]["node"]), q
    finally:
        m.close()
        g.free_batch(bh)


def test_learned_k_above_graph_capacity(world):
    w = world
    d, g = w["d"], w["g"]
    nq, k = 16, 600
    b = E.concat([E.serialize(None, N)] * nq, N, entries=w["ent"][:nq])
    qs = d["queries"][:nq]
    bh = g.stage(qs, b, k)
    m = learned.Model(128, d["centers"].astype(np.float32))
    try:
        for mode in (1, 2):
            with pytest.raises(E.GpuError, match="exceeds"):
                g.run_learned(bh, m, L.LearnedOpts(k=k, mode=mode, ef_max=512))
        g.run_learned(bh, m, L.LearnedOpts(k=k, mode=0, ef_max=512))
        g.sync()
        route, ef = g.routes(bh, nq)
        assert not route.any()
        out, cnt, flag, rc = g.fetch(bh, nq, k)
        assert rc == 0
        oh, oc, ot, _ = O.brute_batch(d["base"], w["cols"], qs, b.qd, b.nodes, b.children,
                                      b.invals, k)
        assert np.array_equal(cnt, oc)
        for q in range(nq):
            assert np.array_equal(out[q, :k]["node"], oh[q, :k]["node"]), q
    finally:
        m.close()
        g.free_batch(bh)
