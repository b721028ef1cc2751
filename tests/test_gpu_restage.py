"""A stream of distinct batches through one captured run
(fv_gpu_batch_restage).

* Uneven batches of 4096 queries (k in {1, 10, 100}, selectivity
  10^U(-4, 0), one distinct price predicate per query, evaluated inline):
  the first batch is staged and its graph-mode run captured; three more
  batches with new queries, predicates and k are restaged into it. Every
  restage keeps the capture (return value 1), the replay for each batch
  equals the oracle's exact top-k for its own k on a sample of queries, and
  equals the eager run of the same batch staged fresh.
* Batches of 64 queries over the six filter classes through predicate
  bitmaps: program, op and IN-value counts differ from batch to batch (the
  kernels read them from the staging area), the capture is kept and every
  replay equals the oracle bit-exact.
* A restage whose predicate table outgrows the reserved capacity reports
  that the capture will be redone (0), and the next graph runs (eager,
  capture, replay) still equal the oracle.
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import synth

pytestmark = pytest.mark.gpu


def _check_sample(out, cnt, trunc, d, cols, qs, b, kq, sel):
    for k in np.unique(kq[sel]):
        s = sel[kq[sel] == k]
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs[s], b.qd[s], b.nodes, b.children,
                                      b.invals, int(k))
        for j, q in enumerate(s):
            c = int(cnt[q])
            assert c == int(oc[j]) and int(trunc[q]) == int(ot[j]), (q, k, c, int(oc[j]))
            assert np.array_equal(out[q, :c]["node"], oh[j, :c]["node"]), (q, k)
            assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                                  oh[j, :c]["dist"].view(np.uint32)), (q, k)


def _uneven_batch(n, nq, rng, post):
    kq = rng.choice(np.array([1, 10, 100], np.uint32), nq)
    kq[0] = 100                      # every batch has the same largest k
    sel = 10.0 ** rng.uniform(-4, 0, nq)
    preds = [E.cmp(0, E.CMP_LT, float(1000.0 * s)) for s in sel]
    return kq, E.concat([E.serialize(p, n, post) for p in preds], n)


def test_uneven_stream_replays_one_capture():
    n, nq, nb = 100_000, 4096, 4
    d = synth.gmm128(n, nq * nb, seed=61)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    rng = np.random.default_rng(62)
    batches = [_uneven_batch(n, nq, rng, post) for _ in range(nb)]
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        kq0, b0 = batches[0]
        q0 = d["queries"][:nq]
        bh = g.stage_ex(q0, b0, kq0)
        try:
            opts = L.BruteOpts(route=1, inline_pred=1, graph=1)
            for _ in range(2):          # eager run, then capture + launch
                g.run_brute(bh, opts)
                g.sync()
            for i, (kq, b) in enumerate(batches):
                qs = d["queries"][i * nq:(i + 1) * nq]
                if i > 0:
                    print(f"case restage uneven batch {i}", flush=True)
                    assert g.restage(bh, qs, b, 100, kq), i
                    g.run_brute(bh, opts)
                    g.sync()
                out, cnt, flag, rc = g.fetch(bh, nq, 100)
                assert rc == 0, i
                sel = np.random.default_rng(70 + i).choice(nq, 192, replace=False)
                _check_sample(out, cnt, flag, d, cols, qs, b, kq, sel)
                # the replay equals an eager run of the same batch staged fresh
                fresh = g.stage_ex(qs, b, kq)
                try:
                    g.run_brute(fresh, L.BruteOpts(route=1, inline_pred=1))
                    g.sync()
                    f = g.fetch(fresh, nq, 100)
                    assert np.array_equal(f[1], cnt) and np.array_equal(f[2], flag), i
                    for q in range(0, nq, 7):
                        c = int(cnt[q])
                        assert np.array_equal(f[0][q, :c], out[q, :c]), (i, q)
                finally:
                    g.free_batch(fresh)
        finally:
            g.free_batch(bh)
    finally:
        g.close()


def _class_batch(d, n, post, i0, nq, shift):
    cls = [synth.FILTER_CLASSES[(i + shift) % 6] for i in range(nq)]
    preds = [synth.filter_pred(c, d["qcent"][i0 + i]) for i, c in enumerate(cls)]
    return E.concat([E.serialize(p, n, post) for p in preds], n)


def test_bitmap_batches_replay_one_capture():
    n, nq, nb = 100_000, 64, 5
    d = synth.gmm128(n, nq * nb, seed=81)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        b0 = _class_batch(d, n, post, 0, nq, 0)
        bh = g.stage(d["queries"][:nq], b0, 10)
        try:
            opts = L.BruteOpts(route=1, graph=1)
            for _ in range(2):
                g.run_brute(bh, opts)
                g.sync()
            nprogs = set()
            for i in range(nb):
                qs = d["queries"][i * nq:(i + 1) * nq]
                b = b0 if i == 0 else _class_batch(d, n, post, i * nq, nq, i)
                if i > 0:
                    print(f"case restage bitmap batch {i}", flush=True)
                    assert g.restage(bh, qs, b, 10), i
                    g.run_brute(bh, opts)
                    g.sync()
                out, cnt, flag, rc = g.fetch(bh, nq, 10)
                assert rc == 0, i
                nprogs.add(len(b.nodes))
                oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b.qd, b.nodes, b.children,
                                              b.invals, 10)
                assert np.array_equal(cnt, oc) and np.array_equal(flag, ot), i
                for q in range(nq):
                    c = int(cnt[q])
                    assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), (i, q)
                    assert np.array_equal(out[q, :c]["dist"].view(np.uint32),
                                          oh[q, :c]["dist"].view(np.uint32)), (i, q)
            assert len(nprogs) > 1      # the predicate tables really differed
        finally:
            g.free_batch(bh)
    finally:
        g.close()


def test_restage_beyond_capacity_recaptures():
    n, nq = 50_000, 64
    d = synth.gmm128(n, nq * 2, seed=91)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        # one shared predicate: tables of one program, one op
        b0 = E.concat([E.serialize(E.cmp(0, E.CMP_LT, 500.0), n, post)] * nq, n)
        bh = g.stage(d["queries"][:nq], b0, 10)
        try:
            opts = L.BruteOpts(route=1, graph=1)
            for _ in range(3):
                g.run_brute(bh, opts)
                g.sync()
            # 64 distinct OR-of-IN predicates: far more ops and IN values
            preds = [E.or_(E.isin(1, [float(i % 32), float((i + 5) % 32)]),
                           E.cmp(0, E.CMP_GT, 10.0 * i)) for i in range(nq)]
            b1 = E.concat([E.serialize(p, n, post) for p in preds], n)
            qs = d["queries"][nq:]
            print("case restage beyond capacity", flush=True)
            assert not g.restage(bh, qs, b1, 10)
            oh, oc, ot, _ = O.brute_batch(d["base"], cols, qs, b1.qd, b1.nodes, b1.children,
                                          b1.invals, 10)
            for it in range(3):     # eager, capture, replay
                g.run_brute(bh, opts)
                g.sync()
                out, cnt, flag, rc = g.fetch(bh, nq, 10)
                assert rc == 0 and np.array_equal(cnt, oc) and np.array_equal(flag, ot), it
                for q in range(nq):
                    c = int(cnt[q])
                    assert np.array_equal(out[q, :c]["node"], oh[q, :c]["node"]), (it, q)
        finally:
            g.free_batch(bh)
    finally:
        g.close()
