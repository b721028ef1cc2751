"""GPU parity of the brute-force path against the oracle (bit-exact).

Every test compares fv_gpu_brute_batch (through the C ABI) with the C
oracle on the same seeded inputs: ids, distances, out_cnt and truncated must
be identical. Covers both scan precisions, the six filter classes, the
span-gather route, uneven batch sizes, edge cases (no passing rows, fewer
passing rows than topb, exact ties, row counts that are not multiples of the
tile), mixed-sign integer data at the exactness bounds (no re-rank at eps =
0), non-integer 768-d data (certified re-rank), the device-resident batch
entry points, ABI error returns, and the reference engine's own host glue
(engine.cpp built with -DFV_GPU) driving this library as a drop-in.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from conftest import have_reference_build
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

pytestmark = pytest.mark.gpu


def assert_same(out, cnt, trunc, ohits, ocnt, otrunc, ctx=""):
    nq = len(cnt)
    assert np.array_equal(cnt, ocnt), f"{ctx}: counts differ {cnt[:8]} vs {ocnt[:8]}"
    assert np.array_equal(trunc, otrunc), f"{ctx}: truncated differs"
    for q in range(nq):
        c = int(cnt[q])
        ids = out[q, :c]["node"]
        oids = ohits[q, :c]["node"]
        if not np.array_equal(ids, oids):
            bad = int(np.argmax(ids != oids))
            raise AssertionError(f"{ctx}: query {q} id mismatch at {bad}: "
                                 f"{ids[max(0,bad-2):bad+3]} vs {oids[max(0,bad-2):bad+3]}")
        d = out[q, :c]["dist"]
        od = ohits[q, :c]["dist"]
        assert np.array_equal(d.view(np.uint32), od.view(np.uint32)), f"{ctx}: query {q} dist bits"


@pytest.fixture(scope="module")
def data100k():
    d = synth.gmm128(100_000, 96, seed=101)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    return d, cols, E.Postings(rows, vals)


@pytest.fixture(scope="module")
def index100k(data100k):
    d, cols, post = data100k
    g = E.GpuIndex(d["base"], cols, post.rows, precision=L.FV_GPU_PREC_BF16)
    yield g
    g.close()


def class_preds(cls, qcent):
    return [synth.filter_pred(cls, c) for c in qcent]


@pytest.mark.parametrize("prec", [L.FV_GPU_PREC_FP32, L.FV_GPU_PREC_BF16,
                                  L.FV_GPU_PREC_FP32_VALU])
@pytest.mark.parametrize("cls", synth.FILTER_CLASSES)
@pytest.mark.parametrize("topb", [10, 64])
def test_filter_classes(index100k, data100k, prec, cls, topb):
    d, cols, post = data100k
    qs = d["queries"][:64]
    preds = class_preds(cls, d["qcent"][:64])
    b = E.concat([E.serialize(p, len(d["base"]), post) for p in preds], len(d["base"]))
    print(f"case prec={prec} cls={cls} topb={topb}", flush=True)
    opts = L.BruteOpts(precision=prec, route=1)
    out, cnt, trunc = index100k.brute(qs, b, topb, opts)
    oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children, b.invals, topb)
    assert_same(out, cnt, trunc, oh, oc, ot, f"{cls}/{prec}/{topb}")


@pytest.mark.parametrize("cls", ["price<100", "cat_in_3_7", "price<1", "cat_eq_centre"])
def test_gather_route(index100k, data100k, cls):
    d, cols, post = data100k
    qs = d["queries"][:40]
    preds = class_preds(cls, d["qcent"][:40])
    b = E.concat([E.serialize(p, len(d["base"]), post) for p in preds], len(d["base"]))
    print(f"case gather cls={cls}", flush=True)
    out, cnt, trunc = index100k.brute(qs, b, 32, L.BruteOpts(route=2))
    assert index100k.stats().route == 2
    oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children, b.invals, 32)
    assert_same(out, cnt, trunc, oh, oc, ot, f"gather/{cls}")


def test_auto_route_and_default_abi(index100k, data100k):
    d, cols, post = data100k
    qs = d["queries"][:64]
    for cls in ["price<1", "all"]:
        preds = class_preds(cls, d["qcent"][:64])
        b = E.concat([E.serialize(p, len(d["base"]), post) for p in preds], len(d["base"]))
        out, cnt, trunc = index100k.brute(qs, b, 10)   # plain fv_gpu_brute_batch
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children, b.invals, 10)
        assert_same(out, cnt, trunc, oh, oc, ot, f"auto/{cls}")
    # price<1 on 100k rows enumerates ~100 rows per query: gather route
    preds = class_preds("price<1", d["qcent"][:64])
    b = E.concat([E.serialize(p, len(d["base"]), post) for p in preds], len(d["base"]))
    index100k.brute(qs, b, 10)
    assert index100k.stats().route == 2


@pytest.mark.parametrize("nq", [1, 63, 65, 96])
def test_uneven_batch_sizes(index100k, data100k, nq):
    d, cols, post = data100k
    r = np.random.default_rng(nq)
    qs = d["queries"][:nq]
    classes = r.choice(synth.FILTER_CLASSES, size=nq)
    preds = [synth.filter_pred(c, d["qcent"][i]) for i, c in enumerate(classes)]
    b = E.concat([E.serialize(p, len(d["base"]), post) for p in preds], len(d["base"]))
    out, cnt, trunc = index100k.brute(qs, b, 16, L.BruteOpts(route=1))
    oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children, b.invals, 16)
    assert_same(out, cnt, trunc, oh, oc, ot, f"nq={nq}")


def test_max_batch_and_large_topb():
    d = synth.gmm128(20_000, 1024, seed=5)
    cols = synth.columns(d)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_BF16)
    try:
        b = E.concat([E.serialize(None, 20_000)] * 1024, 20_000)
        out, cnt, trunc = g.brute(d["queries"], b, 8, L.BruteOpts(route=1))
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes, b.children,
                                      b.invals, 8)
        assert_same(out, cnt, trunc, oh, oc, ot, "nq=1024")
        qs = d["queries"][:4]
        b4 = E.concat([E.serialize(None, 20_000)] * 4, 20_000)
        for topb in (1, 300, 4096):
            out, cnt, trunc = g.brute(qs, b4, topb, L.BruteOpts(route=1))
            oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b4.qd, b4.nodes, b4.children,
                                          b4.invals, topb)
            assert_same(out, cnt, trunc, oh, oc, ot, f"topb={topb}")
            # 8-wave WGs forced: at topb 4096 a WG holds more of a query's
            # candidates than its LDS slots and spill area; the rest go
            # straight to its global pool segment
            out, cnt, trunc = g.brute(qs, b4, topb, L.BruteOpts(route=1, waves=8))
            assert_same(out, cnt, trunc, oh, oc, ot, f"topb={topb} waves=8")
    finally:
        g.close()


def test_edge_cases():
    # row count not a multiple of the 32/64-row tiles; few passing rows; none
    n = 1000 + 37
    d = synth.gmm128(n, 8, seed=3)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16)
    try:
        preds = [None, E.cmp(0, E.CMP_LT, -1.0), E.cmp(0, E.CMP_LT, 3.0),
                 E.cmp(0, E.CMP_GE, 999.0), E.not_(E.cmp(1, E.CMP_NE, 5.0)),
                 E.or_(E.cmp(0, E.CMP_LT, 2.0), E.cmp(0, E.CMP_GT, 998.0)),
                 E.isin(1, []), E.and_()]
        b = E.concat([E.serialize(p, n, post) for p in preds], n)
        for prec in (L.FV_GPU_PREC_FP32, L.FV_GPU_PREC_BF16, L.FV_GPU_PREC_FP32_VALU):
            for route in (1, 0):
                out, cnt, trunc = g.brute(d["queries"], b, 20, L.BruteOpts(precision=prec, route=route))
                oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                              b.children, b.invals, 20)
                assert_same(out, cnt, trunc, oh, oc, ot, f"edge/{prec}/{route}")
        assert cnt[1] == 0 and trunc[1] == 0
    finally:
        g.close()
    # the INT8 scan (an SQ8 index over these integer rows): seeds of filtered
    # rows must stay out also where a query's threshold admits every row
    g8 = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_SQ8)
    try:
        out, cnt, trunc = g8.brute(d["queries"], b, 20, L.BruteOpts(route=1))
        assert g8.stats().precision == L.FV_GPU_PREC_SQ8
        assert_same(out, cnt, trunc, oh, oc, ot, "edge/sq8")
    finally:
        g8.close()


def test_exact_ties_broken_by_id():
    # every row at the same distance: ids must come out ascending
    n = 5000
    base = np.full((n, 128), 7.0, dtype=np.float32)
    base[::3, 0] = 9.0
    q = np.zeros((3, 128), dtype=np.float32)
    g = E.GpuIndex(base, None, None, precision=L.FV_GPU_PREC_BF16)
    try:
        b = E.concat([E.serialize(None, n)] * 3, n)
        for prec in (L.FV_GPU_PREC_FP32, L.FV_GPU_PREC_BF16, L.FV_GPU_PREC_FP32_VALU):
            out, cnt, trunc = g.brute(q, b, 100, L.BruteOpts(precision=prec, route=1))
            oh, oc, ot, _ = O.brute_batch(base, None, q, b.qd, b.nodes, b.children, b.invals, 100)
            assert_same(out, cnt, trunc, oh, oc, ot, f"ties/{prec}")
            want = [i for i in range(n) if i % 3 != 0][:100]
            assert list(out[0, :100]["node"]) == want
    finally:
        g.close()


@pytest.mark.parametrize("dim,lim", [(128, 127), (256, 90)])
def test_mixed_sign_integers_exact_without_rerank(dim, lim):
    # integer rows and queries of both signs at the edge of the exactness
    # bounds (|v| <= lim; D delta^2 and D (V^2 + Q^2) below 2^24, so eps = 0):
    # the seeded MFMA sums pass through large negative values (a third of
    # the queries are negated rows) and must still equal the reference's
    # FP32 distances, since the select stage skips its re-rank at eps = 0
    rng = np.random.default_rng(7 + dim)
    n, nq = 4000 + 13, 12
    base = rng.integers(-lim, lim + 1, (n, dim)).astype(np.float32)
    qs = rng.integers(-lim, lim + 1, (nq, dim)).astype(np.float32)
    qs[:4] = -base[rng.integers(0, n, 4)]
    cols = np.stack([rng.uniform(0, 1000, n).astype(np.float32).astype(np.float64),
                     rng.integers(0, 32, n).astype(np.float64)])
    preds = [None, E.cmp(0, E.CMP_LT, 500.0), E.isin(1, [3.0, 7.0])] * 4
    b = E.concat([E.serialize(p, n) for p in preds], n)
    oh, oc, ot, _ = O.brute_batch(base, cols, qs, b.qd, b.nodes, b.children, b.invals, 20)
    for prec in (L.FV_GPU_PREC_BF16, L.FV_GPU_PREC_FP32, L.FV_GPU_PREC_SQ8):
        g = E.GpuIndex(base, cols, None, precision=prec)
        try:
            out, cnt, trunc = g.brute(qs, b, 20, L.BruteOpts(route=1))
            if prec == L.FV_GPU_PREC_SQ8:
                assert g.stats().precision == L.FV_GPU_PREC_SQ8
            assert_same(out, cnt, trunc, oh, oc, ot, f"mixed-sign {dim}/{prec}")
            if prec == L.FV_GPU_PREC_FP32:
                # an FP32 index is scanned in BF16 unless FP32 is asked for
                assert g.stats().precision == L.FV_GPU_PREC_BF16
                for p2 in (L.FV_GPU_PREC_FP32, L.FV_GPU_PREC_FP32_VALU):
                    out, cnt, trunc = g.brute(qs, b, 20, L.BruteOpts(route=1, precision=p2))
                    assert g.stats().precision == p2
                    assert_same(out, cnt, trunc, oh, oc, ot, f"mixed-sign {dim}/{p2}")
        finally:
            g.close()


def test_non_integer_768d_certified_rerank():
    d = synth.gmm768(20_000, 16)
    cols = synth.columns(d)
    g = E.GpuIndex(d["base"], cols, None, precision=L.FV_GPU_PREC_BF16)
    try:
        preds = [synth.filter_pred(c, d["qcent"][i]) for i, c in
                 enumerate(["all", "price<500", "price<100", "cat_in_3_7"] * 4)]
        b = E.concat([E.serialize(p, 20_000) for p in preds], 20_000)
        for prec in (L.FV_GPU_PREC_BF16, L.FV_GPU_PREC_FP32):
            out, cnt, trunc = g.brute(d["queries"], b, 10, L.BruteOpts(precision=prec, route=1))
            oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"], b.qd, b.nodes,
                                          b.children, b.invals, 10)
            assert_same(out, cnt, trunc, oh, oc, ot, f"768/{prec}")
            # distances within 1e-5 relative of the FP64 recomputation
            for qi in range(16):
                for k in range(int(cnt[qi])):
                    f64 = O.l2sq_f64(d["queries"][qi], d["base"][int(out[qi, k]["node"])])
                    assert abs(float(out[qi, k]["dist"]) - f64) <= 1e-5 * max(f64, 1e-30)
    finally:
        g.close()


def test_device_resident_batch(index100k, data100k):
    d, cols, post = data100k
    qs = d["queries"][:64]
    preds = class_preds("price<100", d["qcent"][:64])
    b = E.concat([E.serialize(p, len(d["base"]), post) for p in preds], len(d["base"]))
    bh = index100k.stage(qs, b, 10)
    try:
        for _ in range(3):
            index100k.run_brute(bh, L.BruteOpts(route=1))
        index100k.sync()
        out, cnt, flag, rc = index100k.fetch(bh, 64, 10)
        assert rc == 0
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children, b.invals, 10)
        assert_same(out, cnt, flag, oh, oc, ot, "staged")
    finally:
        index100k.free_batch(bh)


def test_abi_error_returns(index100k, data100k):
    d, cols, post = data100k
    b = E.concat([E.serialize(None, len(d["base"]))], len(d["base"]))
    with pytest.raises(E.GpuError, match="bad brute batch"):
        index100k.brute(d["queries"][:1], b, 0)
    with pytest.raises(E.GpuError, match="bad brute batch"):
        index100k.brute(d["queries"][:1], b, 5000)
    bad = E.concat([E.serialize(E.cmp(7, E.CMP_LT, 1.0), len(d["base"]))], len(d["base"]))
    with pytest.raises(E.GpuError, match="column out of range"):
        index100k.brute(d["queries"][:1], bad, 10)
    lib = L.load()
    eb = L.errbuf()
    rc = lib.fv_gpu_brute_batch(index100k.h, 0, None, None, None, None, 0, None, 0, None, 0,
                                None, 0, None, 0, 10, None, None, None, eb, len(eb))
    assert rc == -1 and b"bad brute batch" in eb.value


@pytest.mark.skipif(not have_reference_build(), reason="oracle/_ref not built")
def test_reference_host_glue_drives_engine(tmp_path):
    """engine.cpp (-DFV_GPU) linked against libpaves_amd.so: its GPU cursor
    and batch API must return what its CPU BruteCursor returns."""
    d = synth.gmm128(20_000, 12, seed=21)
    cols = synth.columns(d)
    cpu = O.RefEngine()
    gpu = O.RefEngine(gpu=True)
    path = str(tmp_path / "r.idx")
    cpu.build(path, d["base"], cols, M=8, efc=40, threads=8)
    gpu.L.fv_engine_gpu_precision(L.FV_GPU_PREC_BF16)
    ic, ig = cpu.open(path), gpu.open(path)
    try:
        assert gpu.L.fv_index_gpu_ensure(ig, -1) == 0
        for cls in synth.FILTER_CLASSES:
            for qi in range(len(d["queries"])):
                p = synth.filter_pred(cls, d["qcent"][qi])
                pc, k1 = cpu.compile(ic, p)
                pg, k2 = gpu.compile(ig, p)
                q = d["queries"][qi]
                want = cpu.brute(ic, q, pc, 10)
                # the batch entry the worker uses
                nodes = np.zeros(16, dtype=O.PRED_DT)
                children = np.zeros(16, np.int32)
                invals = np.zeros(16, np.float64)
                spans = np.zeros(64, dtype=L.SPAN_DT)
                bp = O.FvGpuBatchPred()
                tot = C.c_uint64()
                assert gpu.L.fv_pred_serialize(ig, pg, nodes.ctypes.data, 16, children.ctypes.data,
                                               16, invals.ctypes.data, 16, spans.ctypes.data, 64,
                                               C.byref(bp), C.byref(tot)) == 0
                out = np.zeros(64, dtype=L.HIT_DT)
                cnt = np.zeros(1, np.uint32)
                tr = np.zeros(1, np.uint8)
                qq = np.ascontiguousarray(q, np.float32)
                rc = gpu.L.fv_gpu_exec_brute_batch(ig, 1, qq.ctypes.data, C.byref(bp), 64,
                                                   out.ctypes.data, cnt.ctypes.data, tr.ctypes.data)
                assert rc == 0
                got = [(int(h["node"]), float(h["dist"])) for h in out[: min(10, int(cnt[0]))]]
                assert got == want[: len(got)] and len(got) == len(want), (cls, qi)
                if pc:
                    cpu.L.fv_pred_free(pc)
                if pg:
                    gpu.L.fv_pred_free(pg)
    finally:
        cpu.close(ic)
        gpu.close(ig)


def test_feedback_table_fills(data100k, index100k):
    """fv_gpu_brute_batch and fv_gpu_graph_batch feed the index's service-
    time table: device time per batch, queries bucketed by the selectivity
    their bitmaps measure (six classes: buckets 0, 1, 2 and 4)."""
    d, cols, post = data100k
    n = d["base"].shape[0]
    qs = d["queries"][:48]
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(48)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    before = index100k.feedback()
    index100k.brute(qs, b, 10, L.BruteOpts(route=1))
    fb = index100k.feedback()
    assert fb.batches[0] == before.batches[0] + 1
    assert fb.batch_us[0] > 0.0 and fb.batch_n[0] > 0.0
    # all-pass / price<500 (>= 20%), price<100 (10%), category IN / = (3-6%),
    # price<1 (0.1%)
    for bk in (0, 1, 2, 4):
        assert fb.q_us[0][bk] > 0.0, bk
