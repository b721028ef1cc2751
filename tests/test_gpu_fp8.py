"""The FP8 index (FV_GPU_PREC_FP8): E4M3 codes scanned with
v_mfma_f32_32x32x16_fp8_fp8, top overfetch x k by code distance re-ranked
with the exact FP32 distance.

The oracle computes the code distance exactly (FP64; every product of two
E4M3 values and their sums fit its mantissa), the MFMA in FP32, so the
candidate sets may differ only among rows whose code distance lies within
that rounding of the R-th one (R = overfetch x k). Each query's result is
checked against that characterisation of the oracle (test_oracle_fp8.py
pins it): every returned row lies within the R nearest codes plus the
rounding band, every row certainly inside them whose exact distance beats
the k-th result is returned, the returned distances equal the exact FP32
distance bit for bit, and counts and truncated flags equal the exact
search's. The scan also decodes the codes as the oracle does: a wrong FP8
interpretation in the MFMA would rank by another distance and fail the
containment. With overfetch x k at least the passing rows the result must
equal the exact brute force bit for bit. Dims 100, 128 and 768 (the wide
code kernel), six filter classes; recall@10 is reported.
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

pytestmark = pytest.mark.gpu


def _world(n, dim, nq, seed):
    d = synth.gmm768(n, nq, seed=seed, ncent=64, dim=dim)
    cols = synth.columns(d)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n) for p in preds], n)
    return d, cols, preds, b


def _check_fp8(d, cols, preds, b, out, cnt, trunc, k, over, ctx):
    base, qs = d["base"], d["queries"]
    n, dim = base.shape
    mid, sc = O.fp8_params(base)
    tab = O.fp8_decode_table().astype(np.float64)
    X = tab[O.fp8_codes(base, mid, sc)]
    Q = tab[O.fp8_codes(qs, mid, sc)]
    xn = (X * X).sum(1)
    qn = (Q * Q).sum(1)
    D = xn[None, :] + qn[:, None] - 2.0 * (Q @ X.T)
    # every passing row in exact (dist, id) order
    eh, ec, et, _ = O.brute_batch(base, cols, qs, b.qd, b.nodes, b.children, b.invals, n)
    assert np.array_equal(cnt, np.minimum(ec, k)), f"{ctx}: counts"
    assert np.array_equal(trunc, (np.minimum(ec, k) == k).astype(np.uint8)), f"{ctx}: truncated"
    xmax = float(xn.max())
    for q in range(len(qs)):
        npass = int(ec[q])
        ids_all = eh[q, :npass]["node"].astype(np.int64)
        ex = dict(zip(ids_all.tolist(), eh[q, :npass]["dist"].tolist()))
        c = int(cnt[q])
        got = out[q, :c]["node"].astype(np.int64)
        gd = out[q, :c]["dist"]
        for j in range(c):
            assert np.float32(ex[int(got[j])]).view(np.uint32) == gd[j].view(np.uint32), (ctx, q, j)
        keys = list(zip(gd.tolist(), got.tolist()))
        assert keys == sorted(keys), (ctx, q, "order")
        R = min(k * over, 4096, npass)
        if R == 0:
            continue
        dq = D[q, ids_all]
        dR = np.sort(dq)[R - 1]
        delta = 2.0 ** -20 * (dim + 4) * (xmax + qn[q] + 2.0 * np.sqrt(xmax * qn[q]))
        hi = set(ids_all[dq <= dR + delta].tolist())
        lo = ids_all[dq < dR - delta] if R < npass else ids_all
        assert set(got.tolist()) <= hi, (ctx, q, "row beyond the R nearest codes")
        if c == k:
            kth = (float(gd[-1]), int(got[-1]))
            gs = set(got.tolist())
            for r in lo.tolist():
                if (ex[r], r) < kth:
                    assert r in gs, (ctx, q, r, "missed a certain candidate")
    hits = sum(len(set(out[q, :cnt[q]]["node"].tolist()) & set(eh[q, :min(ec[q], k)]["node"].tolist()))
               for q in range(len(qs)))
    return hits / max(1, int(np.minimum(ec, k).sum()))


@pytest.mark.parametrize("dim", [100, 128, 768])
def test_fp8_matches_oracle(dim):
    n, nq = (10_000 if dim < 768 else 6_000), 40
    d, cols, preds, b = _world(n, dim, nq, 900 + dim)
    qs = d["queries"]
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_FP8, device=0)
    try:
        assert g.precision == L.FV_GPU_PREC_FP8
        for topb, over in ((10, 0), (10, 8), (32, 2)):
            print(f"case fp8 dim={dim} topb={topb} over={over}", flush=True)
            out, cnt, trunc = g.brute(qs, b, topb, L.BruteOpts(route=1, overfetch=over))
            assert g.stats().precision == L.FV_GPU_PREC_FP8
            rec = _check_fp8(d, cols, preds, b, out, cnt, trunc, topb, over or 4,
                             f"fp8/{dim}/{topb}/{over}")
            print(f"fp8 recall@{topb} dim={dim} over={over or 4}: {rec:.4f}", flush=True)
            if topb == 10 and over == 0:
                assert rec >= 0.9, (dim, rec)
        # asking for SQ8 on an FP8 index scans the same codes
        o1 = g.brute(qs, b, 10, L.BruteOpts(route=1))
        o2 = g.brute(qs, b, 10, L.BruteOpts(route=1, precision=L.FV_GPU_PREC_SQ8))
        for x, y in zip(o1, o2):
            assert np.array_equal(x, y)
    finally:
        g.close()


def test_fp8_full_overfetch_is_exact():
    """overfetch x k >= every query's passing rows: every passing row is
    re-ranked, so the result is the exact search's bit for bit."""
    n, nq, dim = 3_000, 24, 128
    d, cols, preds, b = _world(n, dim, nq, 951)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_FP8, device=0)
    try:
        out, cnt, trunc = g.brute(d["queries"], b, 10, L.BruteOpts(route=1, overfetch=64))
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                      b.invals, 10)
        assert np.array_equal(cnt, oc) and np.array_equal(trunc, ot)
        for q in range(nq):
            c = int(cnt[q])
            assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), q
            assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                                  oh[q, :c]["dist"].view(np.uint32)), q
    finally:
        g.close()


def test_fp8_asked_on_bf16_index_is_exact():
    n, nq, dim = 5_000, 16, 128
    d, cols, preds, b = _world(n, dim, nq, 952)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        out, cnt, trunc = g.brute(d["queries"], b, 10,
                                  L.BruteOpts(route=1, precision=L.FV_GPU_PREC_FP8))
        assert g.stats().precision == L.FV_GPU_PREC_BF16
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                      b.invals, 10)
        assert np.array_equal(cnt, oc)
        for q in range(nq):
            assert np.array_equal(out[q, :cnt[q]]["node"], oh[q, :cnt[q]]["node"]), q
    finally:
        g.close()
