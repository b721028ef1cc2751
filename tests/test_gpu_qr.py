"""GPU parity of the query-resident multi-tile scan (k_scan_bf16_qr).

Batches of at least eight 64-query tiles at 128 (or fewer) dims with
predicate bitmaps take the query-resident kernel by default. Each case
compares the batch's results with the C oracle bit for bit (ids, distances,
counts, truncated):

* six filter classes (at most 16 programs per wave: the penalty-MFMA
  filter form), topb 10 and 64, row counts off the 32-row tile grid;
* one distinct predicate per query plus NULL attribute values (more than
  16 programs per wave: the per-element seed form);
* non-integer rows at 100 dims (zero-padded to 128; certified re-rank);
* a batch that is not a multiple of the WG's eight query tiles;
* the query-major grid, several WGs per CU and hipGraph replays, each equal
  to the default launch;
* inline one-comparison predicates as rank ranges of the column's posting
  order (<, <=, ==, >=, > with ties and NULLs, per-query k), and a batch
  with != (two ranges) that takes the double-precision inline scan.
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

pytestmark = pytest.mark.gpu

QR = L.BruteOpts(route=1, qtw=8)


def assert_same(out, cnt, trunc, ohits, ocnt, otrunc, ctx=""):
    assert np.array_equal(cnt, ocnt), f"{ctx}: counts differ"
    assert np.array_equal(trunc, otrunc), f"{ctx}: truncated differs"
    for q in range(len(cnt)):
        c = int(cnt[q])
        ids, oids = out[q, :c]["node"], ohits[q, :c]["node"]
        if not np.array_equal(ids, oids):
            bad = int(np.argmax(ids != oids))
            raise AssertionError(f"{ctx}: query {q} id mismatch at {bad}: "
                                 f"{ids[max(0, bad - 2):bad + 3]} vs {oids[max(0, bad - 2):bad + 3]}")
        assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                              ohits[q, :c]["dist"].view(np.uint32)), f"{ctx}: query {q} dist bits"


def _setup(n, nq, seed, dim=128, nan_frac=0.0, gmm=synth.gmm128):
    d = gmm(n, nq, seed=seed, dim=dim) if dim != 128 else gmm(n, nq, seed=seed)
    cols = synth.columns(d)
    if nan_frac:
        rng = np.random.default_rng(seed)
        cols[synth.COL_PRICE, rng.random(n) < nan_frac] = np.nan
    rows, vals = synth.postings(cols)
    return d, cols, E.Postings(rows, vals)


@pytest.fixture(scope="module")
def six():
    n, nq = 30_011, 1024
    d, cols, post = _setup(n, nq, seed=31)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16)
    yield d, cols, b, g
    g.close()


@pytest.mark.parametrize("topb", [10, 64])
def test_six_classes(six, topb):
    d, cols, b, g = six
    out, cnt, trunc = g.brute(d["queries"], b, topb, QR)
    oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                  b.invals, topb)
    assert_same(out, cnt, trunc, oh, oc, ot, f"six classes topb={topb}")
    # the library's default shape for this batch is the same kernel
    out2, cnt2, trunc2 = g.brute(d["queries"], b, topb, L.BruteOpts(route=1))
    assert_same(out2, cnt2, trunc2, oh, oc, ot, f"default shape topb={topb}")


def test_grid_orders_and_replay(six):
    d, cols, b, g = six
    ref = g.brute(d["queries"], b, 10, QR)
    for opts in (L.BruteOpts(route=1, qtw=8, order=3), L.BruteOpts(route=1, qtw=8, wg_per_cu=2),
                 L.BruteOpts(route=1, qtw=8, sample_stride=4),
                 L.BruteOpts(route=1, qtw=8, sample_stride=64)):
        got = g.brute(d["queries"], b, 10, opts)
        for x, y in zip(got, ref):
            assert np.array_equal(x, y)
    bh = g.stage(d["queries"], b, 10)
    try:
        eager = None
        for rep in range(4):
            g.run_brute(bh, L.BruteOpts(route=1, qtw=8, graph=1))
            g.sync()
            r = g.fetch(bh, len(d["queries"]), 10)
            assert r[3] == 0
            if eager is None:
                eager = r
            else:
                for x, y in zip(r[:3], eager[:3]):
                    assert np.array_equal(x, y), f"replay {rep}"
        for x, y in zip(eager[:3], ref):
            assert np.array_equal(x, y)
    finally:
        g.free_batch(bh)


def test_many_programs_and_nulls():
    """One predicate per query (640 programs, 64 per wave) over a price
    column with 5% NULLs: comparisons, NE, IN and NOT with NaN semantics."""
    n, nq = 20_003, 640
    d, cols, post = _setup(n, nq, seed=32, nan_frac=0.05)
    preds = []
    for i in range(nq):
        t = float((i * 997) % 1000)
        kind = i % 5
        if kind == 0:
            preds.append(E.cmp(synth.COL_PRICE, E.CMP_LT, t))
        elif kind == 1:
            preds.append(E.cmp(synth.COL_PRICE, E.CMP_GE, t))
        elif kind == 2:
            preds.append(E.not_(E.cmp(synth.COL_PRICE, E.CMP_LT, t)))
        elif kind == 3:
            preds.append(E.and_(E.cmp(synth.COL_PRICE, E.CMP_NE, t),
                                E.isin(synth.COL_CATEGORY, [i % 32, (i * 7) % 32])))
        else:
            preds.append(E.or_(E.cmp(synth.COL_PRICE, E.CMP_LT, t / 10), E.cmp(synth.COL_CATEGORY, E.CMP_EQ, i % 32)))
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16)
    try:
        out, cnt, trunc = g.brute(d["queries"], b, 16, QR)
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                      b.invals, 16)
        assert_same(out, cnt, trunc, oh, oc, ot, "many programs")
    finally:
        g.close()


def test_float_rows_dim100_partial_group():
    """Non-integer rows at 100 dims (eps > 0: certified re-rank) and 520
    queries (nine query tiles: the second WG group has one busy wave)."""
    n, nq = 24_001, 520
    d, cols, post = _setup(n, nq, seed=33, dim=100, gmm=synth.gmm768)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16)
    try:
        for topb in (1, 10):
            out, cnt, trunc = g.brute(d["queries"], b, topb, QR)
            oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                          b.children, b.invals, topb)
            assert_same(out, cnt, trunc, oh, oc, ot, f"dim 100 topb={topb}")
    finally:
        g.close()


def _inline_world(n, nq, seed, col, nan_frac):
    d, cols, post = _setup(n, nq, seed=seed, nan_frac=nan_frac)
    rng = np.random.default_rng(seed + 1)
    kq = rng.choice(np.array([1, 10, 100], np.uint32), nq)
    ops = [E.CMP_LT, E.CMP_LE, E.CMP_EQ, E.CMP_GE, E.CMP_GT]
    vals = cols[col][np.isfinite(cols[col])]
    preds = []
    for i in range(nq):
        if i % 11 == 0:
            preds.append(None)
            continue
        op = ops[i % 5]
        # constants taken from the column (ties at the boundary) or between
        # its values; == only on existing values
        c = float(vals[rng.integers(len(vals))])
        if op != E.CMP_EQ and i % 3 == 0:
            c = float(rng.uniform(0, 1000)) if col == synth.COL_PRICE else c + 0.5
        preds.append(E.cmp(col, op, c))
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    return d, cols, post, kq, b


def _check_k(out, cnt, trunc, d, cols, b, kq):
    for k in np.unique(kq):
        qs = np.flatnonzero(kq == k)
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"][qs], b.qd[qs], b.nodes,
                                      b.children, b.invals, int(k))
        for j, q in enumerate(qs):
            c = int(cnt[q])
            assert c == int(oc[j]) and int(trunc[q]) == int(ot[j]), (q, k, c, int(oc[j]))
            assert np.array_equal(out[q, :c]["node"], oh[j, :c]["node"]), (q, k)
            assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                                  oh[j, :c]["dist"].view(np.uint32)), (q, k)


@pytest.mark.parametrize("col,nan_frac", [(synth.COL_PRICE, 0.05), (synth.COL_CATEGORY, 0.0)])
def test_inline_rank_ranges(col, nan_frac):
    """One comparison per query (<, <=, ==, >=, >; constants on and between
    the column's values; NULLs; no predicate) evaluated as rank ranges in the
    query-resident scan, per-query k in {1, 10, 100}, eager and captured."""
    n, nq = 20_011, 1000
    d, cols, post, kq, b = _inline_world(n, nq, 60 + col, col, nan_frac)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16)
    try:
        bh = g.stage_ex(d["queries"], b, kq)
        try:
            ref = None
            for graph in (0, 1, 1):
                g.run_brute(bh, L.BruteOpts(route=1, inline_pred=1, qtw=8, graph=graph))
                g.sync()
                out, cnt, flag, rc = g.fetch(bh, nq, int(kq.max()))
                assert rc == 0
                if ref is None:
                    _check_k(out, cnt, flag, d, cols, b, kq)
                    ref = (out, cnt, flag)
                else:
                    for x, y in zip((out, cnt, flag), ref):
                        assert np.array_equal(x, y)
        finally:
            g.free_batch(bh)
    finally:
        g.close()


def test_inline_not_equal_falls_back():
    """!= is two rank ranges: such a batch takes the double-precision inline
    scan and still equals the oracle."""
    n, nq = 12_007, 600
    d, cols, post = _setup(n, nq, seed=71, nan_frac=0.05)
    rng = np.random.default_rng(72)
    kq = np.full(nq, 10, np.uint32)
    preds = [E.cmp(synth.COL_PRICE, E.CMP_NE if i % 2 else E.CMP_LT, float(rng.uniform(0, 1000)))
             for i in range(nq)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16)
    try:
        bh = g.stage_ex(d["queries"], b, kq)
        try:
            g.run_brute(bh, L.BruteOpts(route=1, inline_pred=1, qtw=8))
            g.sync()
            out, cnt, flag, rc = g.fetch(bh, nq, 10)
            assert rc == 0
            _check_k(out, cnt, flag, d, cols, b, kq)
        finally:
            g.free_batch(bh)
    finally:
        g.close()


def test_staged_batch_of_4096():
    """A device-resident batch of 4096 queries (uniform k): the
    query-resident scan over 64 query tiles equals the oracle, and a
    routed run in brute mode returns the same."""
    n, nq = 20_011, 4096
    d, cols, post = _setup(n, nq, seed=81)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16)
    bh = g.stage(d["queries"], b, 10)
    try:
        g.run_brute(bh, L.BruteOpts(route=1))
        g.sync()
        out, cnt, trunc, rc = g.fetch(bh, nq, 10)
        assert rc == 0
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                      b.invals, 10)
        assert_same(out, cnt, trunc, oh, oc, ot, "4096 staged")
        g.run_learned(bh, None, L.LearnedOpts(k=10, mode=3))
        g.sync()
        r = g.fetch(bh, nq, 10)
        for x, y in zip(r[:3], (out, cnt, trunc)):
            assert np.array_equal(x, y)
    finally:
        g.free_batch(bh)
        g.close()


def test_staged_batch_of_4096_worker_top64():
    """The reference worker's request, top-64 (max(4k, 64), exec.c:286), on
    a 4096-query batch: the default shape (pre-pass stride 6 at B = 64 with
    32 row chunks, about 12 candidates per query per WG against its 16
    slots) equals the oracle, and so does a run at stride 1."""
    n, nq, k = 20_011, 4096, 64
    d, cols, post = _setup(n, nq, seed=83)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16)
    bh = g.stage(d["queries"], b, k)
    try:
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                      b.invals, k)
        for o in (L.BruteOpts(route=1), L.BruteOpts(route=1, sample_stride=1)):
            g.run_brute(bh, o)
            g.sync()
            out, cnt, trunc, rc = g.fetch(bh, nq, k)
            assert rc == 0
            assert_same(out, cnt, trunc, oh, oc, ot, f"4096 staged top-{k} stride {o.sample_stride}")
    finally:
        g.free_batch(bh)
        g.close()
