"""The oracle and the host HNSW builder against vectors the compiled
reference engine produced (tests/golden/reference_vectors.npz, written by
tests/golden/make_golden.py from oracle/_ref): distances bit for bit, brute
top-20 with five predicate shapes (NULLs included), the single-threaded
HNSW build section by section, HnswCursor hits and graph entry points.
Runs without the reference sources."""
import os

import numpy as np
import pytest

import oracle as O
from paves_amd import engine as E
from paves_amd import hnsw

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


@pytest.mark.parametrize("dim", _mg.L2_DIMS)
def test_l2sq_bits(gv, dim):
    x, q, dist = gv[f"l2_{dim}_x"], gv[f"l2_{dim}_q"], gv[f"l2_{dim}_dist"]
    for qi in range(len(q)):
        for i in range(len(x)):
            assert np.float32(O.l2sq(q[qi], x[i])).view(np.uint32) == dist[qi, i].view(np.uint32)


@pytest.mark.parametrize("name", list(_mg.predicates()))
def test_brute_topk(gv, name):
    base, cols, queries = gv["b_base"], gv["b_cols"], gv["b_queries"]
    p = _mg.predicates()[name]
    n, k = len(base), _mg.BRUTE_K
    b = E.concat([E.serialize(p, n)] * len(queries), n)
    hits, cnt, _, _ = O.brute_batch(base, cols, queries, b.qd, b.nodes, b.children, b.invals, k)
    want_n, want_d, want_c = gv[f"b_{name}_node"], gv[f"b_{name}_dist"], gv[f"b_{name}_cnt"]
    for qi in range(len(queries)):
        c = int(want_c[qi])
        assert int(cnt[qi]) == c, (name, qi)
        assert np.array_equal(hits[qi, :c]["node"], want_n[qi, :c]), (name, qi)
        assert np.array_equal(hits[qi, :c]["dist"].view(np.uint32),
                              want_d[qi, :c].view(np.uint32)), (name, qi)


def test_hnsw_build_single_thread(gv):
    M, efc, entry, top = (int(v) for v in gv["h_meta"])
    g = hnsw.build(gv["h_base"], M=M, ef_construction=efc, threads=1)
    assert g["entry_node"] == entry and g["max_level"] == top
    assert np.array_equal(g["levels"], gv["h_levels"])
    assert np.array_equal(g["upper_offs"], gv["h_upper_offs"])
    assert np.array_equal(g["upper"], gv["h_upper"])
    assert np.array_equal(g["l0"], gv["h_l0"])


@pytest.mark.parametrize("name", list(_mg.predicates()))
def test_hnsw_cursor_and_entry(gv, name):
    base, cols, queries = gv["h_base"], gv["h_cols"], gv["h_queries"]
    M, efc, entry_node, top = (int(v) for v in gv["h_meta"])
    p = _mg.predicates()[name]
    s = E.serialize(p, len(base))
    proot = 0 if len(s.nodes) else -1
    ef = _mg.HNSW_EF
    for qi, q in enumerate(queries):
        e = O.graph_entry(base, gv["h_upper_offs"], gv["h_upper"], M, entry_node, top, q)
        assert e == int(gv["h_entry"][qi])
        hits, exh, vis = O.hnsw_one(base, gv["h_l0"], cols, q, s.nodes, s.children, s.invals,
                                    proot, e, ef, max(200000, ef * 400))
        c = int(gv[f"h_{name}_cnt"][qi])
        assert len(hits) >= c
        assert np.array_equal(hits[:c]["node"], gv[f"h_{name}_node"][qi, :c]), (name, qi)
        assert np.array_equal(hits[:c]["dist"].view(np.uint32),
                              gv[f"h_{name}_dist"][qi, :c].view(np.uint32)), (name, qi)
