"""GPU parity of the filtered level-0 beam search (k_graph) against the CPU
oracle's restatement of the reference HnswCursor (engine.cpp:1582-1731),
plus recall against exact brute force and the ABI's output contract.

Graphs come from this library's host builder (paves_amd/hnsw.py, pinned
bit-for-bit against the reference builder in test_hnsw_build.py), in the
reference level-0 layout [cnt, neighbours...] with stride 2M+1; entry points
from its greedy descent (pinned against the oracle).
"""
import os

import numpy as np
import pytest

from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import hnsw, synth

import oracle as O

pytestmark = pytest.mark.gpu

N = 20_000
NQ = 48
EF = 32


@pytest.fixture(scope="module")
def world():
    d = synth.gmm128(N, NQ, seed=31)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    f = hnsw.build(d["base"], M=8, ef_construction=64, threads=8)
    f["vectors"] = d["base"]
    entries = hnsw.entries(f, d["base"], d["queries"])
    g = E.GpuIndex(d["base"], cols, rows, l0=f["l0"], precision=L.FV_GPU_PREC_BF16, device=0)
    yield {"d": d, "cols": cols, "post": E.Postings(rows, vals), "f": f, "entries": entries,
           "g": g}
    g.close()


def _batch(w, preds):
    return E.concat([E.serialize(p, N, w["post"]) for p in preds], N, entries=w["entries"])


PRED_SETS = {
    "none": lambda qc: None,
    "price<300": lambda qc: E.cmp(0, E.CMP_LT, 300.0),
    "cat_in": lambda qc: E.isin(1, [1.0, 2.0, 3.0]),
    "cat_eq_centre": lambda qc: synth.filter_pred("cat_eq_centre", qc),
    "or_not": lambda qc: E.or_(E.cmp(0, E.CMP_GE, 900.0), E.not_(E.cmp(1, E.CMP_NE, 4.0))),
}


@pytest.mark.parametrize("name", list(PRED_SETS))
def test_w1_equals_cpu_cursor(world, name):
    """One frontier head per iteration = the CPU cursor's pop order: the
    result lists (ids and FP32 distances) and exhausted flags are equal."""
    w = world
    d = w["d"]
    preds = [PRED_SETS[name](d["qcent"][i]) for i in range(NQ)]
    b = _batch(w, preds)
    mv = max(200000, EF * 400)
    out, cnt, exh = w["g"].graph(d["queries"], b, EF, mv, wmax=1, fcap=1024)
    oh, oc, oe = O.hnsw_batch(w["f"]["vectors"], w["f"]["l0"], w["cols"], d["queries"], b.qd,
                              b.nodes, b.children, b.invals, w["entries"], EF, mv)
    for q in range(NQ):
        c = int(cnt[q])
        assert c == int(oc[q]), (name, q, c, int(oc[q]))
        assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), (name, q)
        assert np.array_equal(out[q, :c]["dist"], oh[q, :c]["dist"]), (name, q)
        assert int(exh[q]) == int(oe[q]), (name, q)


@pytest.mark.parametrize("name", list(PRED_SETS))
def test_contract_and_recall(world, name):
    """Default search (several frontier heads per iteration): results pass
    the predicate, are unique, ascending by (dist, node) with exact FP32
    distances, and reach recall@10 >= 0.9 against exact brute force."""
    w = world
    d = w["d"]
    preds = [PRED_SETS[name](d["qcent"][i]) for i in range(NQ)]
    b = _batch(w, preds)
    ef = 64
    out, cnt, exh = w["g"].graph(d["queries"], b, ef, max(200000, ef * 400))
    th, tc, _, _ = O.brute_batch(d["base"], w["cols"], d["queries"], b.qd, b.nodes, b.children,
                                 b.invals, 10)
    rec = []
    for q in range(NQ):
        c = int(cnt[q])
        ids = out[q, :c]["node"].astype(np.int64)
        dist = out[q, :c]["dist"]
        assert len(set(ids.tolist())) == c
        if preds[q] is not None:
            assert E.eval_pred(preds[q], w["cols"])[ids].all()
        keys = [(float(x), int(i)) for x, i in zip(dist, ids)]
        assert keys == sorted(keys)
        for k in range(c):
            assert dist[k] == O.l2sq(d["queries"][q], d["base"][ids[k]])
        truth = th[q, : int(tc[q])]["node"]
        rec.append(O.recall(ids[:10], truth))
    assert np.mean(rec) >= 0.9, (name, np.mean(rec))


def test_exhausted_and_max_visited(world):
    """max_visited caps the search and reports exhausted; a filter nothing
    passes drains the frontier (exhausted, zero hits)."""
    w = world
    d = w["d"]
    preds = [None] * 8 + [E.cmp(0, E.CMP_LT, -5.0)] * 8
    sel = np.r_[np.arange(8), np.arange(8)]
    w2 = dict(w, entries=w["entries"][sel])
    b = _batch(w2, preds)
    out, cnt, exh = w["g"].graph(d["queries"][sel], b, EF, 50)
    oh, oc, oe = O.hnsw_batch(w["f"]["vectors"], w["f"]["l0"], w["cols"], d["queries"][sel],
                              b.qd, b.nodes, b.children, b.invals, w2["entries"], EF, 50)
    assert (exh[:8] == 1).all() and (oe[:8] == 1).all()
    assert (cnt[8:] == 0).all() and (exh[8:] == 1).all()
    # max_visited = 0 stops after the first iteration (gpu_micro.cpp:116-119)
    out0, cnt0, exh0 = w["g"].graph(d["queries"][sel], b, EF, 0)
    assert (exh0 == 1).all()


def test_device_resident_graph_batch(world):
    w = world
    d = w["d"]
    preds = [PRED_SETS["price<300"](0)] * NQ
    b = _batch(w, preds)
    g = w["g"]
    out, cnt, exh = g.graph(d["queries"], b, EF, 200000)
    bh = g.stage(d["queries"], b, 1)
    try:
        for _ in range(3):
            g.run_graph(bh, EF, 200000)
        g.sync()
        o2, c2, f2, rc = g.fetch(bh, NQ, EF)
        assert rc == 0
        assert np.array_equal(c2, cnt) and np.array_equal(f2, exh)
        for q in range(NQ):
            c = int(cnt[q])
            assert np.array_equal(o2[q, :c]["node"], out[q, :c]["node"])
    finally:
        g.free_batch(bh)


def test_graph_abi_errors(world):
    w = world
    d = w["d"]
    b = _batch(w, [None] * 4)
    with pytest.raises(E.GpuError):
        w["g"].graph(d["queries"][:4], b, 0, 1000)          # ef = 0
    with pytest.raises(E.GpuError):
        w["g"].graph(d["queries"][:4], b, 513, 1000)        # ef above FV_GPU_MAX_EF
    bad = E.concat([E.serialize(None, N)] * 4, N, entries=[N + 5] * 4)
    with pytest.raises(E.GpuError):
        w["g"].graph(d["queries"][:4], bad, 16, 1000)       # entry out of range
