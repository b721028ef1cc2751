"""fv_gpu_brute_batch_noalloc (the allocation-free variant of
fv_gpu_brute_batch, buffers carved from the library's static device
scratch) against fv_gpu_brute_batch on the same arguments: two calls in a
row, the second with a smaller batch (the scratch keeps the first call's
contents), plus every scan precision; and the clean error for a batch whose
buffers exceed the scratch."""
import numpy as np
import pytest

from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

pytestmark = pytest.mark.gpu

N = 100_000


def noalloc(g, queries, b, topb):
    lib = L.load()
    q = np.ascontiguousarray(queries, dtype=np.float32)
    out = np.zeros((b.nq, topb), dtype=L.HIT_DT)
    cnt = np.zeros(b.nq, dtype=np.uint32)
    trunc = np.zeros(b.nq, dtype=np.uint8)
    eb = L.errbuf()
    rc = lib.fv_gpu_brute_batch_noalloc(g.h, b.nq, L.ptr(q), L.ptr(b.qd), L.ptr(b.dedup),
                                        L.ptr(b.nodes), len(b.nodes), L.ptr(b.children),
                                        len(b.children), L.ptr(b.invals), len(b.invals),
                                        L.ptr(b.spans), len(b.spans), L.ptr(b.spanoff),
                                        len(b.spanoff), topb, L.ptr(out), L.ptr(cnt),
                                        L.ptr(trunc), eb, len(eb))
    if rc != 0:
        raise E.GpuError(eb.value.decode())
    return out, cnt, trunc


def same(a, b, ctx):
    out, cnt, tr = a
    out2, cnt2, tr2 = b
    assert np.array_equal(cnt, cnt2) and np.array_equal(tr, tr2), ctx
    for q in range(len(cnt)):
        c = int(cnt[q])
        assert np.array_equal(out[q, :c]["node"], out2[q, :c]["node"]), (ctx, q)
        assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                              out2[q, :c]["dist"].view(np.uint32)), (ctx, q)


@pytest.fixture(scope="module")
def world():
    d = synth.gmm128(N, 96, seed=909)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    return d, cols, E.Postings(rows, vals)


@pytest.mark.parametrize("prec", [L.FV_GPU_PREC_BF16, L.FV_GPU_PREC_FP32, L.FV_GPU_PREC_SQ8])
def test_noalloc_equals_brute_batch(world, prec):
    d, cols, post = world
    g = E.GpuIndex(d["base"], cols, post.rows, precision=prec, device=0)
    try:
        for nq, topb, seed in ((64, 10, 0), (17, 64, 1)):
            print(f"case noalloc prec={prec} nq={nq} topb={topb}", flush=True)
            cls = np.random.default_rng(seed).choice(synth.FILTER_CLASSES, nq)
            preds = [synth.filter_pred(c, d["qcent"][i]) for i, c in enumerate(cls)]
            b = E.concat([E.serialize(p, N, post) for p in preds], N)
            qs = d["queries"][:nq]
            want = g.brute(qs, b, topb)
            got = noalloc(g, qs, b, topb)
            same(want, got, (prec, nq))
    finally:
        g.close()


def test_noalloc_rejects_oversized_batch(world):
    d, cols, post = world
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        qs = np.repeat(d["queries"], 11, axis=0)[:1024]
        b = E.concat([E.serialize(None, N)] * 1024, N)
        with pytest.raises(E.GpuError, match="no-alloc"):
            noalloc(g, qs, b, 4096)
        # the scratch is usable again after the rejected call
        b2 = E.concat([E.serialize(None, N)] * 8, N)
        same(g.brute(qs[:8], b2, 10), noalloc(g, qs[:8], b2, 10), "after reject")
    finally:
        g.close()
