"""The committed reference vectors (tests/golden/reference_vectors.npz,
outputs of the reference engine compiled from source) run through the GPU
path: brute top-20 under five predicate shapes with 5% NULL categories
(every scan precision the index offers, full-scan and automatic routes), and
the level-0 beam search with one frontier head per iteration against the
reference HnswCursor hits on the reference-built graph. Bit-exact ids and
FP32 distances; no reference sources are needed."""
import os

import numpy as np
import pytest

import oracle as O

from paves_amd import _lib as L
from paves_amd import engine as E

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "reference_vectors.npz")

import importlib.util  # noqa: E402

_spec = importlib.util.spec_from_file_location("make_golden",
                                               os.path.join(HERE, "golden", "make_golden.py"))
_mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_mg)


@pytest.fixture(scope="module")
def gv():
    with np.load(PATH) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("prec", [L.FV_GPU_PREC_FP32, L.FV_GPU_PREC_BF16, L.FV_GPU_PREC_SQ8])
@pytest.mark.parametrize("route", [0, 1])
def test_brute_golden(gv, prec, route):
    base, cols, queries = gv["b_base"], gv["b_cols"], gv["b_queries"]
    n, k = len(base), _mg.BRUTE_K
    rows, vals = _postings(cols)
    g = E.GpuIndex(base, cols, rows, precision=prec, device=0)
    try:
        for name, p in _mg.predicates().items():
            print(f"case golden brute prec={prec} route={route} pred={name}", flush=True)
            b = E.concat([E.serialize(p, n, E.Postings(rows, vals))] * len(queries), n)
            out, cnt, trunc = g.brute(queries, b, k, L.BruteOpts(route=route))
            want_n, want_d, want_c = gv[f"b_{name}_node"], gv[f"b_{name}_dist"], gv[f"b_{name}_cnt"]
            st = g.stats()
            if prec == L.FV_GPU_PREC_SQ8 and st.precision == L.FV_GPU_PREC_SQ8 and st.route == 1:
                # these rows are not integers: on the scan route the
                # quantized codes rank and the top 4 k are re-ranked exactly
                # (the oracle's restatement); the gather route is exact
                oh, oc, ot = O.sq8_brute_batch(base, cols, queries, b.qd, b.nodes, b.children,
                                               b.invals, k, 4)
                want_n, want_d, want_c = oh["node"], oh["dist"], oc
            for qi in range(len(queries)):
                c = int(want_c[qi])
                assert int(cnt[qi]) == c, (name, qi)
                assert int(trunc[qi]) == int(c == k), (name, qi)
                assert np.array_equal(out[qi, :c]["node"], want_n[qi, :c]), (name, qi)
                assert np.array_equal(out[qi, :c]["dist"].view(np.uint32),
                                      want_d[qi, :c].view(np.uint32)), (name, qi)
    finally:
        g.close()


def test_graph_golden(gv):
    base, cols, queries = gv["h_base"], gv["h_cols"], gv["h_queries"]
    n = len(base)
    ef = _mg.HNSW_EF
    entries = gv["h_entry"].astype(np.uint32)
    g = E.GpuIndex(base, cols, None, l0=gv["h_l0"], precision=L.FV_GPU_PREC_BF16, device=0)
    try:
        for name, p in _mg.predicates().items():
            print(f"case golden graph pred={name}", flush=True)
            b = E.concat([E.serialize(p, n)] * len(queries), n, entries=entries)
            out, cnt, exh = g.graph(queries, b, ef, max(200000, 400 * ef), wmax=1, fcap=1024)
            for qi in range(len(queries)):
                c = int(gv[f"h_{name}_cnt"][qi])
                assert int(cnt[qi]) >= c, (name, qi)
                assert np.array_equal(out[qi, :c]["node"], gv[f"h_{name}_node"][qi, :c]), (name, qi)
                assert np.array_equal(out[qi, :c]["dist"].view(np.uint32),
                                      gv[f"h_{name}_dist"][qi, :c].view(np.uint32)), (name, qi)
    finally:
        g.close()


def _postings(cols):
    from paves_amd import synth
    return synth.postings(cols)
