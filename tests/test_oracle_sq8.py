"""Pins the oracle's restatement of the quantized SQ8 path
(or_sq8_params / or_sq8_codes / or_sq8_brute_batch) before the GPU is
checked against it.

* Quantizer known answers, worked by hand: midrange centres, the one scale
  127 / largest half-range, round-half-even, clamping, constant dims.
* Code distances and the candidate set against an independent numpy
  restatement in int64 (codes from numpy's own float32 arithmetic), with
  predicates: the same R = over x k candidates by (code distance, row), the
  same exact re-rank.
* Reduction to the exact search: with over x k at least the number of
  passing rows every passing row is re-ranked exactly, so the result must
  equal the exact brute force (or_brute_batch) bit for bit.
* Recall against the exact search is high and never drops when the
  overfetch grows (the candidate sets are nested).
"""
import numpy as np

import oracle as O
from paves_amd import engine as E
from paves_amd import synth


def test_quantizer_known_answers():
    v = np.array([[0.0, 10.0, 3.0], [2.0, -10.0, 3.0], [1.5, 5.0, 3.0]], np.float32)
    mid, inv = O.sq8_params(v)
    # lo/hi per dim: [0, 2], [-10, 10], [3, 3] -> mid [1, 0, 3]; half-ranges
    # [1, 10, 0] -> scale 127 / 10
    assert mid.tolist() == [1.0, 0.0, 3.0]
    assert inv == np.float32(12.7)
    c = O.sq8_codes(v, mid, inv)
    # (0-1)*12.7 = -12.7 -> -13; 10*12.7 = 127; (1.5-1)*12.7 = 6.35 -> 6;
    # 5*12.7 = 63.5 -> 64 (half to even); the constant dim -> 0
    assert c.tolist() == [[-13, 127, 0], [13, -127, 0], [6, 64, 0]]
    # beyond the rows' range a query clamps
    q = O.sq8_codes(np.array([[20.0, -40.0, 3.0]], np.float32), mid, inv)
    assert q.tolist() == [[127, -127, 0]]
    # every dim constant: scale 1
    m2, i2 = O.sq8_params(np.full((4, 2), 7.0, np.float32))
    assert m2.tolist() == [7.0, 7.0] and i2 == np.float32(1.0)


def _numpy_sq8(base, cols, queries, preds, topb, over):
    mid = ((base.min(0) + base.max(0)) * np.float32(0.5)).astype(np.float32)
    half = ((base.max(0) - base.min(0)) * np.float32(0.5)).astype(np.float32)
    hmax = half.max()
    inv = np.float32(127.0) / hmax if hmax > 0 else np.float32(1.0)

    def codes(x):
        c = np.rint(((x - mid).astype(np.float32) * inv).astype(np.float32))
        return np.clip(c, -127, 127).astype(np.int64)

    xc = codes(base)
    xn = (xc * xc).sum(1)
    out = []
    for qi, q in enumerate(queries):
        qc = codes(q[None, :])[0]
        mask = np.ones(len(base), bool) if preds[qi] is None else E.eval_pred(preds[qi], cols)
        rows = np.flatnonzero(mask)
        dq = (xn[rows] + (qc * qc).sum() - 2 * (xc[rows] @ qc)).astype(np.float32)
        order = np.lexsort((rows, dq))
        R = min(topb * over, 4096, len(rows))
        cand = rows[order[:R]]
        ex = np.array([O.l2sq(q, base[r]) for r in cand], np.float32)
        o2 = np.lexsort((cand, ex))
        out.append((cand[o2][:topb], ex[o2][:topb]))
    return out


def _world(n, dim, nq, seed):
    d = synth.gmm768(n, nq, seed=seed, ncent=16, dim=dim)
    cols = synth.columns(d)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n) for p in preds], n)
    return d, cols, preds, b


def test_sq8_restatement_matches_numpy():
    n, dim, nq = 3000, 40, 12
    d, cols, preds, b = _world(n, dim, nq, 11)
    for topb, over in ((10, 4), (5, 1), (25, 3)):
        oh, oc, ot = O.sq8_brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                       b.children, b.invals, topb, over)
        ref = _numpy_sq8(d["base"], cols, d["queries"], preds, topb, over)
        for q in range(nq):
            ids, ds = ref[q]
            c = int(oc[q])
            assert c == len(ids) and int(ot[q]) == int(c == topb), (q, topb, over)
            assert np.array_equal(oh[q, :c]["node"], ids), (q, topb, over)
            assert np.array_equal(oh[q, :c]["dist"].view(np.uint32), ds.view(np.uint32))


def test_sq8_full_overfetch_is_exact():
    n, dim, nq = 600, 24, 12
    d, cols, preds, b = _world(n, dim, nq, 12)
    eh, ec, et, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                  b.invals, 10)
    sh, sc, st = O.sq8_brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                   b.invals, 10, 64)     # 640 >= 600 passing rows
    assert np.array_equal(ec, sc) and np.array_equal(et, st)
    for q in range(nq):
        c = int(ec[q])
        assert np.array_equal(eh[q, :c], sh[q, :c]), q


def test_sq8_recall_high_and_monotone():
    n, dim, nq = 8000, 64, 24
    d, cols, preds, b = _world(n, dim, nq, 13)
    eh, ec, _, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                 b.invals, 10)
    prev = None
    for over in (1, 2, 4, 8):
        sh, sc, _ = O.sq8_brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                      b.children, b.invals, 10, over)
        per = [O.recall(sh[q, : sc[q]]["node"], eh[q, : ec[q]]["node"]) for q in range(nq)]
        if prev is not None:
            assert all(a >= p for a, p in zip(per, prev)), over
        prev = per
        if over >= 4:
            assert np.mean(per) >= 0.95, (over, np.mean(per))
