"""The SQ8 precision: an exact INT8 MFMA scan for integer data.

An index uploaded with FV_GPU_PREC_SQ8 over integer values spanning at most
255 (and at most 256 dims) keeps an int8 copy of x - centre; batches whose
queries are integers inside the same window scan it with
v_mfma_i32_32x32x32_i8 (exact dot products, exact distances). Every result
must equal the oracle's exact top-k bit for bit, like the BF16 and FP32
scans. Batches whose queries fall outside the window take the BF16 scan
(reported precision BF16) and are still exact. Data outside the window
gets the quantized copy instead (test_gpu_sq8_quant.py).
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

from test_gpu_brute import assert_same, class_preds

pytestmark = pytest.mark.gpu


def _data(n, nq, dim, seed):
    if dim == 128:
        d = synth.gmm128(n, nq, seed=seed)
    else:
        # integer GMM at 256 dims: the 128-d mixture twice over, same range
        a = synth.gmm128(n, nq, seed=seed)
        b = synth.gmm128(n, nq, seed=seed + 1)
        d = dict(a)
        d["base"] = np.ascontiguousarray(np.concatenate([a["base"], b["base"]], 1))
        d["queries"] = np.ascontiguousarray(np.concatenate([a["queries"], b["queries"]], 1))
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    return d, cols, E.Postings(rows, vals)


@pytest.mark.parametrize("dim", [128, 256])
def test_int8_scan_equals_oracle(dim):
    n, nq = 60_000, 64
    d, cols, post = _data(n, nq, dim, 211 + dim)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_SQ8)
    try:
        for cls in synth.FILTER_CLASSES:
            preds = class_preds(cls, d["qcent"][:nq])
            b = E.concat([E.serialize(p, n, post) for p in preds], n)
            for topb in (10, 100):
                out, cnt, trunc = g.brute(d["queries"], b, topb, L.BruteOpts(route=1))
                assert g.stats().precision == L.FV_GPU_PREC_SQ8, (cls, "INT8 scan not taken")
                oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                              b.children, b.invals, topb)
                assert_same(out, cnt, trunc, oh, oc, ot, f"sq8 dim={dim} {cls} topb={topb}")
            # the shared-chip shape (4-wave workgroups) gives the same answer
            out, cnt, trunc = g.brute(d["queries"], b, 10, L.BruteOpts(route=1, waves=4,
                                                                     wg_per_cu=1))
            oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                          b.children, b.invals, 10)
            assert_same(out, cnt, trunc, oh, oc, ot, f"sq8 dim={dim} {cls} 4 waves")
    finally:
        g.close()


def test_queries_outside_window_take_bf16():
    """A non-integer query, or one outside [centre-128, centre+127], sends
    its batch to the BF16 scan; results stay exact."""
    n, nq = 40_000, 32
    d, cols, post = _data(n, nq, 128, 307)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_SQ8)
    try:
        b = E.concat([E.serialize(None, n)] * nq, n)
        for tweak in ("half", "far"):
            q = d["queries"].copy()
            if tweak == "half":
                q[3, 7] += 0.5
            else:
                q[5, 0] = 400.0
            out, cnt, trunc = g.brute(q, b, 10, L.BruteOpts(route=1))
            assert g.stats().precision == L.FV_GPU_PREC_BF16, tweak
            oh, oc, ot, _ = O.brute_batch(d["base"], cols, q, b.qd, b.nodes, b.children,
                                          b.invals, 10)
            assert_same(out, cnt, trunc, oh, oc, ot, f"fallback {tweak}")
        out, cnt, trunc = g.brute(d["queries"], b, 10, L.BruteOpts(route=1))
        assert g.stats().precision == L.FV_GPU_PREC_SQ8
    finally:
        g.close()


def test_wide_range_index_takes_quantized_codes():
    """Values spanning more than 255 cannot shift into int8 exactly: the
    index keeps quantized codes and SQ8 requests equal the oracle's
    quantized path; BF16 requested on the same index stays exact."""
    n, nq = 30_000, 16
    d, cols, post = _data(n, nq, 128, 401)
    base = d["base"].copy()
    base[:, 0] *= 3.0          # column 0 now spans up to 654
    g = E.GpuIndex(base, cols, post.rows, precision=L.FV_GPU_PREC_SQ8)
    try:
        b = E.concat([E.serialize(None, n)] * nq, n)
        out, cnt, trunc = g.brute(d["queries"], b, 10, L.BruteOpts(route=1))
        assert g.stats().precision == L.FV_GPU_PREC_SQ8
        oh, oc, ot = O.sq8_brute_batch(base, cols, d["queries"], b.qd, b.nodes, b.children,
                                       b.invals, 10, 4)
        assert_same(out, cnt, trunc, oh, oc, ot, "wide range, quantized")
        out, cnt, trunc = g.brute(d["queries"], b, 10,
                                  L.BruteOpts(route=1, precision=L.FV_GPU_PREC_BF16))
        assert g.stats().precision == L.FV_GPU_PREC_BF16
        eh, ec, et, _ = O.brute_batch(base, cols, d["queries"], b.qd, b.nodes, b.children,
                                      b.invals, 10)
        assert_same(out, cnt, trunc, eh, ec, et, "wide range, BF16")
    finally:
        g.close()


def test_int8_inline_uneven_batch():
    """Uneven batch (per-query k, one-comparison predicates evaluated inside
    the INT8 scan) against the oracle per query."""
    n, nq = 80_000, 200
    d, cols, post = _data(n, nq, 128, 503)
    rng = np.random.default_rng(5)
    kq = rng.choice(np.array([1, 10, 100], np.uint32), nq)
    sel = 10.0 ** rng.uniform(-4, 0, nq)
    preds = [E.cmp(0, E.CMP_LT, float(1000.0 * s)) for s in sel]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_SQ8)
    try:
        bh = g.stage_ex(d["queries"], b, kq)
        try:
            g.run_brute(bh, L.BruteOpts(route=1, inline_pred=1))
            g.sync()
            assert g.stats().precision == L.FV_GPU_PREC_SQ8
            out, cnt, flag, rc = g.fetch(bh, nq, int(kq.max()))
            assert rc == 0
            for k in (1, 10, 100):
                qs = np.flatnonzero(kq == k)
                oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"][qs], b.qd[qs],
                                              b.nodes, b.children, b.invals, int(k))
                for j, q in enumerate(qs):
                    c = int(cnt[q])
                    assert c == int(oc[j]) and int(flag[q]) == int(ot[j]), (q, k)
                    assert np.array_equal(out[q, :c], oh[j, :c]), (q, k)
        finally:
            g.free_batch(bh)
    finally:
        g.close()


@pytest.mark.parametrize("float_rows", [False, True])
def test_large_batch_takes_exact_query_resident_scan(float_rows):
    """An SQ8 index served a batch of 16 query tiles without a precision
    asked: the query-resident BF16 scan runs (reported BF16) and the result
    is the exact top-k bit for bit — on integer rows (exact INT8 copy) and
    on float rows (quantized copy). Asking for SQ8 keeps the code scan."""
    n, nq = 30_011, 1024
    d, cols, post = _data(n, nq, 128, 611)
    base = d["base"]
    if float_rows:
        base = (base + np.random.default_rng(7).random(base.shape, dtype=np.float32) * 0.5)
        base = np.ascontiguousarray(base.astype(np.float32))
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    g = E.GpuIndex(base, cols, post.rows, precision=L.FV_GPU_PREC_SQ8, device=0)
    try:
        print(f"case sq8 large batch float_rows={float_rows}", flush=True)
        out, cnt, trunc = g.brute(d["queries"], b, 10, L.BruteOpts(route=1))
        assert g.stats().precision == L.FV_GPU_PREC_BF16
        oh, oc, ot, _ = O.brute_batch(base, cols, d["queries"], b.qd, b.nodes, b.children,
                                      b.invals, 10)
        assert_same(out, cnt, trunc, oh, oc, ot, f"sq8 index, {nq} queries")
        # the synchronous reference entry point gets the same
        o2, c2, t2 = g.brute(d["queries"], b, 10)
        for x, y in zip((o2, c2, t2), (out, cnt, trunc)):
            assert np.array_equal(x, y)
        g.brute(d["queries"], b, 10, L.BruteOpts(route=1, precision=L.FV_GPU_PREC_SQ8))
        assert g.stats().precision == L.FV_GPU_PREC_SQ8
    finally:
        g.close()
