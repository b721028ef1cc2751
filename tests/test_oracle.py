"""Pins the C oracle (oracle/paves_oracle.c) before it is trusted.

Checks, each chosen so that a plausible mistake in the restatement fails:
* or_l2sq is bit-identical to the compiled reference l2sq (engine.cpp:43-63)
  on random non-integer data of awkward dimensions (tails, single blocks),
  observed through the reference BruteCursor's returned distances;
* on integer data every distance equals the FP64 closed form;
* bf16 rounding matches a direct round-to-nearest-even computation and is
  exact on integers up to 256;
* predicate known-answer tests (NULL/NaN, NE, IN, AND/OR/NOT, empty AND/OR);
* or_brute_topb equals the reference BruteCursor (ids and distances) with
  and without predicates, and a pure-Python FP64 brute force on tiny input;
* or_hnsw_search equals the reference HnswCursor on a graph the reference
  builder wrote, and or_graph_entry equals fv_index_graph_entry.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from conftest import have_reference_build
from paves_amd import engine as E
from paves_amd import synth

needs_ref = pytest.mark.skipif(not have_reference_build(), reason="oracle/_ref not built")


def _rand(n, d, seed, integer=False):
    r = np.random.default_rng(seed)
    x = r.normal(0, 3, size=(n, d)).astype(np.float32)
    if integer:
        x = np.rint(np.abs(x) * 20).astype(np.float32)
    return x


# ----------------------------------------------------------------- distances

def test_l2sq_integer_equals_fp64():
    x = _rand(200, 128, 1, integer=True)
    q = _rand(1, 128, 2, integer=True)[0]
    for i in range(200):
        assert float(O.l2sq(q, x[i])) == O.l2sq_f64(q, x[i])


def test_l2sq_tail_order_differs_from_naive():
    # the restated order must not silently equal a naive left-to-right sum on
    # all inputs (otherwise the bit-exact reference test below pins little)
    x = _rand(500, 77, 3)
    q = _rand(1, 77, 4)[0]
    naive = []
    for i in range(500):
        s = np.float32(0)
        for j in range(77):
            d = np.float32(q[j] - x[i, j])
            s = np.float32(s + np.float32(d * d))
        naive.append(s)
    ours = [O.l2sq(q, x[i]) for i in range(500)]
    assert any(a != b for a, b in zip(ours, naive))


@needs_ref
@pytest.mark.parametrize("dim", [8, 16, 17, 20, 23, 24, 28, 31, 32, 40, 45, 48, 77, 128, 130, 768])
def test_l2sq_bit_identical_to_reference(tmp_path, dim):
    ref = O.RefEngine()
    x = _rand(300, dim, 10 + dim)
    path = str(tmp_path / f"d{dim}.idx")
    ref.build(path, x, None, M=8, efc=32, threads=1)
    idx = ref.open(path)
    try:
        for qi in range(3):
            q = _rand(1, dim, 100 + qi)[0]
            hits = ref.brute(idx, q, None, 300, threads=1)
            assert len(hits) == 300
            for node, dist in hits:
                assert np.float32(dist) == O.l2sq(q, x[node]), (dim, node)
    finally:
        ref.close(idx)


def test_bf16_rounding():
    def rne(v):
        u = np.float32(v).view(np.uint32)
        u = (int(u) + 0x7FFF + ((int(u) >> 16) & 1)) & 0xFFFFFFFF
        return u >> 16
    r = np.random.default_rng(5)
    for v in list(r.normal(0, 100, 500).astype(np.float32)) + [0.0, -0.0, 1.0, 255.0, 256.0, 257.0]:
        assert O.bf16_bits(float(v)) == rne(v)
    for k in range(0, 257):
        b = O.bf16_bits(float(k))
        back = np.array([b << 16], dtype=np.uint32).view(np.float32)[0]
        assert back == k


# ----------------------------------------------------------------- predicates

def _ser(p, count=4):
    s = E.serialize(p, count)
    return s.nodes, s.children, s.invals


def test_predicate_kats():
    nan = float("nan")
    cols = np.array([[1.0, 2.0, nan, 5.0], [3.0, 7.0, 3.0, nan]])
    cases = [
        (E.cmp(0, E.CMP_LT, 2.0), [1, 0, 0, 0]),
        (E.cmp(0, E.CMP_LE, 2.0), [1, 1, 0, 0]),
        (E.cmp(0, E.CMP_EQ, 5.0), [0, 0, 0, 1]),
        (E.cmp(0, E.CMP_GE, 2.0), [0, 1, 0, 1]),
        (E.cmp(0, E.CMP_GT, 1.0), [0, 1, 0, 1]),
        (E.cmp(0, E.CMP_NE, 2.0), [1, 0, 0, 1]),   # NaN is not "!= 2"
        (E.cmp(0, E.CMP_EQ, nan), [0, 0, 0, 0]),
        (E.isin(1, [3.0, 9.0]), [1, 0, 1, 0]),
        (E.isin(1, []), [0, 0, 0, 0]),
        (E.and_(E.cmp(0, E.CMP_GE, 1.0), E.isin(1, [3.0])), [1, 0, 0, 0]),
        (E.or_(E.cmp(0, E.CMP_GT, 4.0), E.isin(1, [7.0])), [0, 1, 0, 1]),
        (E.not_(E.cmp(0, E.CMP_LT, 2.0)), [0, 1, 1, 1]),  # NOT(NaN<2) is true
        (E.and_(), [1, 1, 1, 1]),
        (E.or_(), [0, 0, 0, 0]),
        (E.not_(E.and_(E.cmp(0, E.CMP_LT, 3.0), E.not_(E.isin(1, [7.0])))), [0, 1, 1, 1]),
    ]
    for p, want in cases:
        nodes, children, invals = _ser(p)
        got = [O.pred_eval(nodes, children, invals, cols, 0, r) for r in range(4)]
        assert got == [bool(w) for w in want], p
        assert list(E.eval_pred(p, cols)) == [bool(w) for w in want], p


# ----------------------------------------------------------------- brute force

def _small(n=3000, nq=12, seed=7):
    d = synth.gmm128(n, nq, seed=seed)
    return d, synth.columns(d)


def _batch(preds, n, post=None):
    return E.concat([E.serialize(p, n, post) for p in preds], n)


PREDS = [None, E.cmp(0, E.CMP_LT, 100.0), E.isin(1, [3.0, 7.0]),
         E.and_(E.cmp(0, E.CMP_GE, 100.0), E.cmp(0, E.CMP_LT, 300.0), E.cmp(1, E.CMP_NE, 4.0)),
         E.or_(E.cmp(0, E.CMP_LT, 20.0), E.cmp(1, E.CMP_EQ, 5.0)),
         E.not_(E.cmp(0, E.CMP_LT, 900.0)), E.cmp(0, E.CMP_LT, -1.0)]


def test_brute_tiny_fp64_bruteforce():
    r = np.random.default_rng(11)
    x = r.integers(0, 6, size=(40, 16)).astype(np.float32)   # many exact ties
    cols = np.vstack([r.random(40) * 10, r.integers(0, 3, 40)]).astype(np.float64)
    qs = r.integers(0, 6, size=(5, 16)).astype(np.float32)
    for p in [None, E.cmp(0, E.CMP_LT, 5.0), E.isin(1, [1.0])]:
        b = _batch([p] * 5, 40)
        hits, cnt, trunc, passing = O.brute_batch(x, cols, qs, b.qd, b.nodes, b.children, b.invals, 7)
        mask = np.ones(40, bool) if p is None else E.eval_pred(p, cols)
        for qi in range(5):
            pairs = sorted((sum((float(qs[qi, j]) - float(x[i, j])) ** 2 for j in range(16)), i)
                           for i in range(40) if mask[i])
            want = pairs[:7]
            assert cnt[qi] == len(want)
            assert trunc[qi] == (1 if len(want) == 7 else 0)
            assert passing[qi] == mask.sum()
            for k, (d, i) in enumerate(want):
                assert hits[qi, k]["node"] == i and float(hits[qi, k]["dist"]) == d


@needs_ref
def test_brute_equals_reference_cursor(tmp_path):
    d, cols = _small()
    ref = O.RefEngine()
    path = str(tmp_path / "b.idx")
    ref.build(path, d["base"], cols, M=8, efc=40, threads=4)
    idx = ref.open(path)
    try:
        for p in PREDS:
            ph, keep = ref.compile(idx, p)
            b = _batch([p] * len(d["queries"]), len(d["base"]))
            hits, cnt, _, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                            b.children, b.invals, 50)
            for qi, q in enumerate(d["queries"]):
                got = ref.brute(idx, q, ph, 50, threads=4)
                mine = [(int(h["node"]), float(h["dist"])) for h in hits[qi, : cnt[qi]]]
                assert mine == got, (p, qi)
            if ph:
                ref.L.fv_pred_free(ph)
    finally:
        ref.close(idx)


@needs_ref
def test_serialize_matches_reference(tmp_path):
    """E.serialize / make_driver produce the reference's spans and flags."""
    d, cols = _small(n=2000)
    ref = O.RefEngine()
    path = str(tmp_path / "s.idx")
    ref.build(path, d["base"], cols, M=8, efc=32, threads=4)
    idx = ref.open(path)
    f = O.read_index_file(path)
    post = E.Postings(f["post_rows"], f["post_vals"])
    try:
        extra = [E.and_(E.cmp(0, E.CMP_GE, 100.0), E.cmp(0, E.CMP_LT, 200.0)),
                 E.or_(E.isin(1, [1.0, 2.0]), E.cmp(0, E.CMP_GT, 990.0)),
                 E.and_(E.isin(1, [3.0]), E.cmp(0, E.CMP_LT, 500.0))]
        for p in PREDS[1:] + extra:
            ph, keep = ref.compile(idx, p)
            nodes = np.zeros(64, dtype=O.PRED_DT)
            children = np.zeros(64, np.int32)
            invals = np.zeros(64, np.float64)
            spans = np.zeros(64, dtype=E.SPAN_DT)
            bp = O.FvGpuBatchPred()
            tot = O.C.c_uint64()
            rc = ref.L.fv_pred_serialize(idx, ph, nodes.ctypes.data, 64, children.ctypes.data, 64,
                                         invals.ctypes.data, 64, spans.ctypes.data, 64,
                                         O.C.byref(bp), O.C.byref(tot))
            assert rc == 0
            s = E.serialize(p, 2000, post)
            assert bp.nnodes == len(s.nodes) and bp.need_eval == s.need_eval, p
            assert bp.dedup == s.dedup and tot.value == s.total, p
            assert bp.nspans == len(s.spans)
            for i in range(bp.nspans):
                assert tuple(spans[i]) == tuple(s.spans[i]), p
            for i in range(bp.nnodes):
                assert tuple(nodes[i])[:7] == tuple(s.nodes[i])[:7]
                assert nodes[i]["value"] == s.nodes[i]["value"] or (
                    math.isnan(nodes[i]["value"]) and math.isnan(s.nodes[i]["value"]))
            ref.L.fv_pred_free(ph)
        # posting lists of the reference builder equal synth.postings
        rows, vals = synth.postings(cols)
        assert np.array_equal(rows, f["post_rows"]) and np.array_equal(vals, f["post_vals"])
    finally:
        ref.close(idx)


# ----------------------------------------------------------------- HNSW

@needs_ref
def test_hnsw_and_entry_equal_reference(tmp_path):
    d, cols = _small(n=4000, nq=16, seed=9)
    ref = O.RefEngine()
    path = str(tmp_path / "h.idx")
    ref.build(path, d["base"], cols, M=8, efc=64, threads=1)
    idx = ref.open(path)
    f = O.read_index_file(path)
    try:
        for p in [None, E.cmp(0, E.CMP_LT, 300.0), E.isin(1, [1.0, 2.0, 3.0])]:
            ph, keep = ref.compile(idx, p)
            s = E.serialize(p, 4000)
            proot = 0 if len(s.nodes) else -1
            for qi, q in enumerate(d["queries"]):
                entry = ref.graph_entry(idx, q)
                mine_entry = O.graph_entry(f["vectors"], f["upper_offs"], f["upper"], f["M"],
                                           f["entry_node"], f["max_level"], q)
                assert entry == mine_entry
                ef = 24
                got = ref.hnsw(idx, q, ph, ef, ef)   # first phase only
                hits, exh, vis = O.hnsw_one(f["vectors"], f["l0"], cols, q, s.nodes, s.children,
                                            s.invals, proot, entry, ef, max(200000, ef * 400))
                mine = [(int(h["node"]), float(h["dist"])) for h in hits]
                assert mine[: len(got)] == got, (p, qi)
            if ph:
                ref.L.fv_pred_free(ph)
    finally:
        ref.close(idx)
