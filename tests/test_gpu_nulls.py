"""NULL attributes (stored as NaN, engine.cpp:397-430) through every GPU
predicate path, against the oracle, bit-exact.

Columns carry 5% NULLs in both price and category. Comparisons with a NULL
row value are false (NE included: it passes v == v && v != c); NOT of a
comparison on a NULL row is true; IN never matches NULL; a NaN constant
matches nothing except under NE. Paths: the predicate bitmap (full scan),
predicates evaluated inline in the scan (one comparison per query on one
column), the span-gather route, the FP32, BF16 and exact INT8 scans, and the
level-0 beam search (one frontier head per iteration equals the CPU cursor).
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import hnsw, synth

pytestmark = pytest.mark.gpu

N = 20_000
NAN = float("nan")


@pytest.fixture(scope="module")
def world():
    d = synth.gmm128(N, 48, seed=505)
    cols = synth.columns(d)
    r = np.random.default_rng(506)
    cols[0, r.random(N) < 0.05] = np.nan
    cols[1, r.random(N) < 0.05] = np.nan
    rows, vals = synth.postings(cols)
    return d, cols, E.Postings(rows, vals)


def mixed_preds():
    """Every comparison op, IN, NOT/OR/AND and NaN constants."""
    return [
        E.cmp(0, E.CMP_LT, 300.0), E.cmp(0, E.CMP_LE, 300.0), E.cmp(1, E.CMP_EQ, 5.0),
        E.cmp(0, E.CMP_GE, 700.0), E.cmp(0, E.CMP_GT, 700.0), E.cmp(1, E.CMP_NE, 5.0),
        E.isin(1, [3.0, 7.0]), E.isin(1, [NAN, 2.0]), E.not_(E.cmp(0, E.CMP_LT, 300.0)),
        E.or_(E.cmp(0, E.CMP_LT, 100.0), E.not_(E.cmp(1, E.CMP_NE, 4.0))),
        E.and_(E.cmp(0, E.CMP_GE, 200.0), E.isin(1, [1.0, 2.0, 3.0])),
        E.cmp(0, E.CMP_EQ, NAN), E.cmp(0, E.CMP_NE, NAN), E.not_(E.isin(1, [3.0])),
        None, E.cmp(1, E.CMP_LT, 8.0),
    ]


def inline_preds():
    """One comparison per query, all on the price column (inline-eligible)."""
    return [E.cmp(0, E.CMP_LT, 300.0), E.cmp(0, E.CMP_LE, 300.0), E.cmp(0, E.CMP_EQ, 500.0),
            E.cmp(0, E.CMP_GE, 700.0), E.cmp(0, E.CMP_GT, 700.0), E.cmp(0, E.CMP_NE, 250.0),
            E.cmp(0, E.CMP_EQ, NAN), E.cmp(0, E.CMP_NE, NAN)]


def check(out, cnt, trunc, oh, oc, ot, ctx):
    assert np.array_equal(cnt, oc), f"{ctx}: counts {cnt} vs {oc}"
    assert np.array_equal(trunc, ot), f"{ctx}: truncated"
    for q in range(len(cnt)):
        c = int(cnt[q])
        assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), (ctx, q)
        assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                              oh[q, :c]["dist"].view(np.uint32)), (ctx, q)


def _batch(preds, post, nq):
    ps = [preds[i % len(preds)] for i in range(nq)]
    return E.concat([E.serialize(p, N, post) for p in ps], N)


@pytest.mark.parametrize("prec", [L.FV_GPU_PREC_FP32, L.FV_GPU_PREC_BF16, L.FV_GPU_PREC_SQ8])
def test_nulls_bitmap_and_gather(world, prec):
    d, cols, post = world
    qs = d["queries"][:48]
    b = _batch(mixed_preds(), post, 48)
    oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children, b.invals, 15)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=prec, device=0)
    try:
        for route in (1, 2):
            print(f"case nulls prec={prec} route={route}", flush=True)
            out, cnt, trunc = g.brute(qs, b, 15, L.BruteOpts(route=route))
            if route == 1 and prec == L.FV_GPU_PREC_SQ8:
                assert g.stats().precision == L.FV_GPU_PREC_SQ8
            check(out, cnt, trunc, oh, oc, ot, f"nulls/{prec}/{route}")
        # a NaN constant under EQ admits nothing
        assert int(oc[11]) == 0 and int(cnt[11]) == 0
    finally:
        g.close()


@pytest.mark.parametrize("prec", [L.FV_GPU_PREC_BF16, L.FV_GPU_PREC_SQ8])
def test_nulls_inline(world, prec):
    d, cols, post = world
    qs = d["queries"][:32]
    b = _batch(inline_preds(), post, 32)
    oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children, b.invals, 12)
    g = E.GpuIndex(d["base"], cols, None, precision=prec, device=0)
    try:
        print(f"case nulls inline prec={prec}", flush=True)
        out, cnt, trunc = g.brute(qs, b, 12, L.BruteOpts(route=1, inline_pred=1))
        check(out, cnt, trunc, oh, oc, ot, f"nulls-inline/{prec}")
        # the same batch staged with per-query k (uneven path, inline)
        kq = np.array([1, 5, 12, 3] * 8, np.uint32)
        bh = g.stage_ex(qs, b, kq)
        try:
            g.run_brute(bh, L.BruteOpts(route=1, inline_pred=1))
            g.sync()
            out2, cnt2, fl2, rc = g.fetch(bh, 32, 12)
            assert rc == 0
            for q in range(32):
                k = int(kq[q])
                c = min(k, int(oc[q]))
                assert int(cnt2[q]) == c, q
                assert np.array_equal(out2[q, :c]["node"], oh[q, :c]["node"]), q
                assert int(fl2[q]) == int(c == k), q
        finally:
            g.free_batch(bh)
    finally:
        g.close()


def test_nulls_graph(world):
    d, cols, post = world
    f = hnsw.build(d["base"], M=8, ef_construction=64, threads=8)
    qs = d["queries"][:48]
    entries = hnsw.entries(f, d["base"], qs)
    preds = [mixed_preds()[i % len(mixed_preds())] for i in range(48)]
    b = E.concat([E.serialize(p, N) for p in preds], N, entries=entries)
    ef = 24
    mv = max(200000, 400 * ef)
    oh, oc, oe = O.hnsw_batch(d["base"], f["l0"], cols, qs, b.qd, b.nodes, b.children, b.invals,
                              entries, ef, mv)
    g = E.GpuIndex(d["base"], cols, None, l0=f["l0"], precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        print("case nulls graph", flush=True)
        out, cnt, exh, dbg = g.graph(qs, b, ef, mv, wmax=1, fcap=1024, debug=True)
        assert not dbg[:, 3].any(), "frontier entries lost"
        for q in range(48):
            c = int(cnt[q])
            assert c == int(oc[q]), q
            assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), q
            assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                                  oh[q, :c]["dist"].view(np.uint32)), q
            assert int(exh[q]) == int(oe[q]), q
            mask = None if preds[q] is None else E.eval_pred(preds[q], cols)
            ids = out[q, : cnt[q]]["node"].astype(np.int64)
            if mask is not None:
                assert mask[ids].all(), q
    finally:
        g.close()

