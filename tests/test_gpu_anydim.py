"""The BF16 MFMA scan at dims that are not multiples of 128.

The tiled BF16 copy and the BF16 query operands are zero-padded to the next
multiple of 128 dims (96 -> 128, 100 -> 128, 200 -> 256, 960 -> 1024); the
padding adds exactly zero to every dot product, and the certified re-rank
restores the FP32 distances in the reference summation order. Checked
bit-exact against the oracle's brute force (ids, FP32 distances, counts,
truncated flags) on non-integer unit-norm-mixture data and on integer data
(the exact-scan case, eps = 0), for the six filter classes, topb 10 and 64,
with the scan precision actually taken reported as BF16. The plain FP32
MFMA scan (dims that are multiples of 8 up to 512) and the FP32 VALU scan
(every other dim) are checked at the same dims.
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

pytestmark = pytest.mark.gpu

N = 12_000
NQ = 48
CLASSES = synth.FILTER_CLASSES


def _check(out, cnt, trunc, oh, oc, ot, ctx):
    assert np.array_equal(cnt, oc), f"{ctx}: counts {cnt[:8]} vs {oc[:8]}"
    assert np.array_equal(trunc, ot), f"{ctx}: truncated"
    for q in range(len(cnt)):
        c = int(cnt[q])
        assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), (ctx, q)
        assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                              oh[q, :c]["dist"].view(np.uint32)), (ctx, q)


def _world(dim, integer):
    if integer:
        d = synth.gmm128(N, NQ, seed=900 + dim, ncent=64, dim=dim)
    else:
        d = synth.gmm768(N, NQ, seed=700 + dim, ncent=64, dim=dim)
    cols = synth.columns(d)
    preds = [synth.filter_pred(CLASSES[i % 6], d["qcent"][i]) for i in range(NQ)]
    b = E.concat([E.serialize(p, N) for p in preds], N)
    return d, cols, b


@pytest.mark.parametrize("integer", [False, True], ids=["float", "int"])
@pytest.mark.parametrize("dim", [96, 100, 200, 960])
def test_bf16_scan_padded_dims(dim, integer):
    d, cols, b = _world(dim, integer)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        for topb in (10, 64):
            print(f"case anydim dim={dim} int={integer} topb={topb}", flush=True)
            oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                          b.children, b.invals, topb)
            out, cnt, trunc = g.brute(d["queries"], b, topb, L.BruteOpts(route=1))
            assert g.stats().precision == L.FV_GPU_PREC_BF16, dim
            _check(out, cnt, trunc, oh, oc, ot, f"bf16/{dim}/{integer}/{topb}")
    finally:
        g.close()


@pytest.mark.parametrize("dim", [96, 100, 200])
def test_fp32_index_and_plain_fp32_at_padded_dims(dim):
    """An FP32 index is served by the padded BF16 scan; FP32 asked for takes
    the FP32 MFMA scan (dim % 8 == 0) or the VALU scan (100): all equal."""
    d, cols, b = _world(dim, False)
    oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                  b.invals, 10)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_FP32, device=0)
    try:
        print(f"case anydim fp32 index dim={dim}", flush=True)
        out, cnt, trunc = g.brute(d["queries"], b, 10, L.BruteOpts(route=1))
        assert g.stats().precision == L.FV_GPU_PREC_BF16
        _check(out, cnt, trunc, oh, oc, ot, f"fp32-index/{dim}")
        out, cnt, trunc = g.brute(d["queries"], b, 10,
                                  L.BruteOpts(route=1, precision=L.FV_GPU_PREC_FP32))
        want = L.FV_GPU_PREC_FP32 if dim % 8 == 0 else L.FV_GPU_PREC_FP32_VALU
        assert g.stats().precision == want, (dim, g.stats().precision)
        _check(out, cnt, trunc, oh, oc, ot, f"fp32/{dim}")
    finally:
        g.close()


def test_padded_dim_batched_and_uneven():
    """dim 100 (padded to 128) through the multi-tile scan (a batch of 256
    queries, two query tiles per workgroup) and a staged uneven batch with
    per-query k, against the oracle."""
    dim, nq = 100, 256
    d = synth.gmm768(N, nq, seed=1100, ncent=64, dim=dim)
    cols = synth.columns(d)
    preds = [synth.filter_pred(CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, N) for p in preds], N)
    oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                  b.invals, 32)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        print("case anydim batched dim=100", flush=True)
        out, cnt, trunc = g.brute(d["queries"], b, 32, L.BruteOpts(route=1))
        _check(out, cnt, trunc, oh, oc, ot, "batched/100")
        kq = np.array([1, 10, 32, 7] * (nq // 4), np.uint32)
        bh = g.stage_ex(d["queries"], b, kq)
        try:
            g.run_brute(bh, L.BruteOpts(route=1))
            g.sync()
            o2, c2, f2, rc = g.fetch(bh, nq, 32)
            assert rc == 0
            for q in range(nq):
                k = int(kq[q])
                c = min(k, int(oc[q]))
                assert int(c2[q]) == c, q
                assert np.array_equal(o2[q, :c]["node"], oh[q, :c]["node"]), q
                assert np.array_equal(o2[q, :c]["dist"].view(np.uint32),
                                      oh[q, :c]["dist"].view(np.uint32)), q
        finally:
            g.free_batch(bh)
    finally:
        g.close()
