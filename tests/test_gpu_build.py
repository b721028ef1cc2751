"""GPU parity of the device HNSW builder (fv_hnsw_build_gpu, pv_build.hip).

* With one frontier head per search iteration and the unbounded frontier,
  the device phase equals the oracle's restatement of the batch semantics
  (or_hnsw_batch_insert, pinned on the CPU by test_oracle_build.py) bit for
  bit: every level-0 list, for the default schedule and for batches of one
  node — which is also the host builder's sequential insertion
  (fv_hnsw_build_seeded with rest_on_host). Dims 24, 37 (tail paths of the
  distance order) and 128, M 4, 8 and 16 (hub lists re-pruned in chunks).
* The default shape (eight heads, bounded frontier) at 60k x 128 builds a
  deterministic graph that keeps the invariants (ids in range, no self
  links or duplicates, counts <= M0, every node linked) and searches with
  the recall of the host-built graph (within 0.005 at ef 64, recall@10
  against exact truth).
* keep = 1 leaves the built graph on the device: searches on it equal
  searches on an index uploaded from the exported lists.
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import hnsw
from paves_amd import synth

pytestmark = pytest.mark.gpu


def _lists(l0):
    return [tuple(r[1:1 + r[0]].tolist()) for r in l0]


def _rest_entries(g, x):
    lev = g["levels"]
    rest = np.array([i for i in range(1, len(lev)) if lev[i] == 0], np.uint32)
    ents = np.array([O.graph_entry(x, g["upper_offs"], g["upper"], g["M"], g["entry_node"],
                                   g["max_level"], x[i]) for i in rest], np.uint32)
    return rest, ents


@pytest.mark.parametrize("n,dim,M,efc,div", [(3000, 24, 8, 48, 8), (2500, 37, 4, 20, 4),
                                             (4000, 128, 16, 100, 8)])
def test_device_phase_equals_oracle(n, dim, M, efc, div):
    d = synth.gmm768(n, 8, seed=40 + dim, dim=dim, ncent=32)
    x = d["base"]
    g = E.GpuIndex(x, None, None, precision=L.FV_GPU_PREC_FP32, device=0)
    try:
        seed = hnsw.build_seeded(x, M, efc, threads=1)
        rest, ents = _rest_entries(seed, x)
        sizes = O.build_schedule(n, len(rest), div, 32768)
        print(f"case build n={n} dim={dim} M={M} efc={efc} batches={len(sizes)}", flush=True)
        dev = hnsw.build_gpu(g, x, M, efc, threads=1, heads=1, exact=True, batch_div=div,
                             keep=False)
        assert dev["stats"]["batches"] == len(sizes)
        assert np.array_equal(dev["upper"], seed["upper"])
        assert np.array_equal(dev["levels"], seed["levels"])
        want = O.hnsw_batch_insert(x, seed["l0"], rest, ents, sizes, M, efc)
        got, exp = _lists(dev["l0"]), _lists(want)
        bad = [i for i in range(n) if got[i] != exp[i]]
        assert not bad, (len(bad), bad[:3], got[bad[0]], exp[bad[0]])
        # batches of one node: the sequential host insertion
        print(f"case build batches of one n={n} dim={dim}", flush=True)
        one = hnsw.build_gpu(g, x, M, efc, threads=1, heads=1, exact=True, max_batch=1,
                             keep=False)
        host = hnsw.build_seeded(x, M, efc, threads=1, rest_on_host=True)
        assert _lists(one["l0"]) == _lists(host["l0"])
    finally:
        g.close()


def _recall(g, x, graph, qs, truth, ef):
    ents = hnsw.entries(graph, x, qs)
    hit = 0
    for s in range(0, len(qs), 1024):
        qq = qs[s:s + 1024]
        b = E.concat([E.serialize(None, len(x)) for _ in range(len(qq))], len(x),
                     entries=ents[s:s + 1024])
        out, cnt, _ = g.graph(qq, b, ef, len(x) + 1)
        for i in range(len(qq)):
            hit += len(set(out[i, :min(10, cnt[i])]["node"].tolist()) & set(truth[s + i].tolist()))
    return hit / truth.size


def test_default_build_quality_and_determinism():
    n, nq, M, efc = 60_000, 2048, 16, 200
    d = synth.gmm128(n, nq, seed=77)
    x, qs = d["base"], d["queries"]
    g = E.GpuIndex(x, None, None, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        print("case build default 60k x 128", flush=True)
        a = hnsw.build_gpu(g, x, M, efc, threads=1, keep=False)
        b = hnsw.build_gpu(g, x, M, efc, threads=1, keep=True)
        print("stats", a["stats"], b["stats"], flush=True)
        assert _lists(a["l0"]) == _lists(b["l0"]), "device build not deterministic"
        l0 = a["l0"]
        c = l0[:, 0]
        assert (c >= 1).all() and (c <= 2 * M).all()
        for v in range(0, n, 7):
            ids = l0[v, 1:1 + c[v]]
            assert (ids < n).all() and v not in ids and len(set(ids.tolist())) == c[v]
        host = hnsw.build(x, M, efc, threads=16)
        # exact truth from the GPU brute force (bit-exact with the oracle)
        truth = np.zeros((nq, 10), np.int64)
        for s in range(0, nq, 1024):
            sub = E.concat([E.serialize(None, n) for _ in range(min(1024, nq - s))], n)
            out, cnt, _ = g.brute(qs[s:s + 1024], sub, 10, L.BruteOpts(route=1))
            truth[s:s + 1024] = out["node"]
        hg = E.GpuIndex(x, None, None, l0=host["l0"], precision=L.FV_GPU_PREC_BF16, device=0)
        try:
            rh = _recall(hg, x, host, qs, truth, 64)
        finally:
            hg.close()
        rd = _recall(g, x, a, qs, truth, 64)   # g holds the device-built graph (keep)
        print(f"recall@10 ef64: device-built {rd:.4f} host-built {rh:.4f}", flush=True)
        assert rd >= rh - 0.005, (rd, rh)
    finally:
        g.close()


def test_keep_installs_graph():
    n, nq, M, efc = 8000, 256, 8, 64
    d = synth.gmm128(n, nq, seed=78)
    x, qs = d["base"], d["queries"]
    g = E.GpuIndex(x, None, None, precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        built = hnsw.build_gpu(g, x, M, efc, threads=4, keep=True)
        ref = E.GpuIndex(x, None, None, l0=built["l0"], precision=L.FV_GPU_PREC_BF16, device=0)
        try:
            ents = hnsw.entries(built, x, qs)
            b = E.concat([E.serialize(None, n) for _ in range(nq)], n, entries=ents)
            o1 = g.graph(qs, b, 48, n + 1)
            o2 = ref.graph(qs, b, 48, n + 1)
            for u, v in zip(o1, o2):
                assert np.array_equal(u, v)
        finally:
            ref.close()
    finally:
        g.close()
