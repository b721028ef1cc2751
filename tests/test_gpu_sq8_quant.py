"""The quantized SQ8 path: SQ8 indexes over non-integer data (or integer
data beyond the exact INT8 window).

The index keeps int8 codes of the per-dim midrange-centred rows under one
scale, zero-padded to 64, 128 or a multiple of 256 dims; the INT8 MFMA scan
(v_mfma_i32_32x32x32_i8) ranks rows by the exact integer code distance and
the select stage re-ranks the top overfetch x k rows with the exact FP32
distance. Every result must equal the oracle's restatement
(or_sq8_brute_batch, pinned in test_oracle_sq8.py) bit for bit — ids, FP32
distances, counts, truncated flags — at dims 100, 128, 200 and 768, six
filter classes, topb 10 and 64, overfetch 2, 4 (default) and 8, through
predicate bitmaps, inline predicates and per-query k; recall@10 against
the exact search is reported and must reach 0.9 at the default overfetch.
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

pytestmark = pytest.mark.gpu


def _check(out, cnt, trunc, oh, oc, ot, ctx):
    assert np.array_equal(cnt, oc), f"{ctx}: counts {cnt[:8]} vs {oc[:8]}"
    assert np.array_equal(trunc, ot), f"{ctx}: truncated"
    for q in range(len(cnt)):
        c = int(cnt[q])
        assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), (ctx, q)
        assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                              oh[q, :c]["dist"].view(np.uint32)), (ctx, q)


def _world(n, dim, nq, seed):
    d = synth.gmm768(n, nq, seed=seed, ncent=64, dim=dim)
    cols = synth.columns(d)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n) for p in preds], n)
    return d, cols, preds, b


@pytest.mark.parametrize("dim", [100, 128, 200, 768])
def test_quantized_sq8_equals_oracle(dim):
    n, nq = (12_000 if dim < 768 else 8_000), 48
    d, cols, preds, b = _world(n, dim, nq, 600 + dim)
    qs = d["queries"]
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_SQ8, device=0)
    try:
        for topb, over in ((10, 0), (10, 8), (64, 2)):
            print(f"case sq8q dim={dim} topb={topb} over={over}", flush=True)
            oh, oc, ot = O.sq8_brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children,
                                           b.invals, topb, over or 4)
            out, cnt, trunc = g.brute(qs, b, topb, L.BruteOpts(route=1, overfetch=over))
            assert g.stats().precision == L.FV_GPU_PREC_SQ8, dim
            _check(out, cnt, trunc, oh, oc, ot, f"sq8q/{dim}/{topb}/{over}")
            if topb == 10 and over == 0:
                eh, ec, _, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children,
                                             b.invals, 10)
                rec = np.mean([O.recall(out[q, : cnt[q]]["node"], eh[q, : ec[q]]["node"])
                               for q in range(nq)])
                print(f"sq8 quantized recall@10 dim={dim}: {rec:.4f}", flush=True)
                assert rec >= 0.9, (dim, rec)
    finally:
        g.close()


def test_quantized_sq8_inline_uneven():
    """Per-query k in {1, 10, 100} and one-comparison predicates evaluated in
    the INT8 scan's epilogue: each query equals the oracle at its own k."""
    n, nq, dim = 20_000, 128, 128
    d = synth.gmm768(n, nq, seed=777, ncent=64, dim=dim)
    cols = synth.columns(d)
    rng = np.random.default_rng(778)
    kq = rng.choice(np.array([1, 10, 100], np.uint32), nq)
    sel = 10.0 ** rng.uniform(-2, 0, nq)
    preds = [E.cmp(0, E.CMP_LT, float(1000.0 * s)) for s in sel]
    b = E.concat([E.serialize(p, n) for p in preds], n)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_SQ8, device=0)
    try:
        bh = g.stage_ex(d["queries"], b, kq)
        try:
            print("case sq8q inline uneven", flush=True)
            g.run_brute(bh, L.BruteOpts(route=1, inline_pred=1))
            g.sync()
            assert g.stats().precision == L.FV_GPU_PREC_SQ8
            out, cnt, flag, rc = g.fetch(bh, nq, int(kq.max()))
            assert rc == 0
            for k in (1, 10, 100):
                qsel = np.flatnonzero(kq == k)
                oh, oc, ot = O.sq8_brute_batch(d["base"], cols, d["queries"][qsel], b.qd[qsel],
                                               b.nodes, b.children, b.invals, int(k), 4)
                for j, q in enumerate(qsel):
                    c = int(cnt[q])
                    assert c == int(oc[j]) and int(flag[q]) == int(ot[j]), (q, k)
                    assert np.array_equal(out[q, :c]["node"], oh[j, :c]["node"]), (q, k)
                    assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                                          oh[j, :c]["dist"].view(np.uint32)), (q, k)
        finally:
            g.free_batch(bh)
    finally:
        g.close()
