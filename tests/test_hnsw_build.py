"""The host HNSW builder and entry descent (paves_amd/hnsw.py) against the
reference builder compiled under oracle/_ref and the CPU oracle."""
import numpy as np
import pytest

from paves_amd import engine as E
from paves_amd import hnsw, synth

import oracle as O


def _ref():
    try:
        return O.RefEngine()
    except RuntimeError:
        pytest.skip("reference build not present (oracle/_ref)")


@pytest.mark.parametrize("dim", [8, 17, 31, 48, 128, 130])
def test_l2sq_equals_oracle(dim):
    rng = np.random.default_rng(dim)
    a = rng.standard_normal((64, dim)).astype(np.float32) * 3
    b = rng.standard_normal((64, dim)).astype(np.float32) * 3
    for i in range(64):
        assert hnsw.l2sq(a[i], b[i]) == O.l2sq(a[i], b[i])


@pytest.mark.parametrize("n,M,efc", [(3000, 8, 48), (2500, 16, 100)])
def test_single_thread_build_equals_reference(tmp_path, n, M, efc):
    """One worker: every section equals the reference builder's file."""
    ref = _ref()
    d = synth.gmm128(n, 4, seed=n + M)
    cols = synth.columns(d)
    path = str(tmp_path / "r.idx")
    ref.build(path, d["base"], cols, M=M, efc=efc, threads=1)
    f = O.read_index_file(path)
    g = hnsw.build(d["base"], M=M, ef_construction=efc, threads=1)
    assert g["entry_node"] == f["entry_node"] and g["max_level"] == f["max_level"]
    assert np.array_equal(g["levels"], f["levels"])
    assert np.array_equal(g["upper_offs"], f["upper_offs"])
    assert np.array_equal(g["upper"], f["upper"])
    assert np.array_equal(g["l0"], f["l0"])


def test_entries_equal_oracle():
    d = synth.gmm128(5000, 32, seed=5)
    g = hnsw.build(d["base"], M=8, ef_construction=64, threads=4)
    got = hnsw.entries(g, d["base"], d["queries"], threads=4)
    for i, q in enumerate(d["queries"]):
        assert got[i] == O.graph_entry(d["base"], g["upper_offs"], g["upper"], g["M"],
                                       g["entry_node"], g["max_level"], q)


def test_parallel_build_recall():
    """Several workers (scheduling-dependent graph): structure is valid and
    the oracle's cursor reaches recall@10 >= 0.95 unfiltered at ef = 64."""
    n = 8000
    d = synth.gmm128(n, 40, seed=8)
    g = hnsw.build(d["base"], M=16, ef_construction=100, threads=8)
    l0 = g["l0"]
    assert l0.shape == (n, 33)
    assert (l0[:, 0] <= 32).all() and (l0[:, 0] >= 1).all()
    for i in range(0, n, 97):
        nb = l0[i, 1: 1 + l0[i, 0]]
        assert (nb < n).all() and i not in nb and len(set(nb.tolist())) == len(nb)
    ent = hnsw.entries(g, d["base"], d["queries"])
    b = E.concat([E.serialize(None, n)] * 40, n, entries=ent)
    oh, oc, _ = O.hnsw_batch(d["base"], l0, None, d["queries"], b.qd, b.nodes, b.children,
                             b.invals, ent, 64, 200000)
    th, tc, _, _ = O.brute_batch(d["base"], None, d["queries"], b.qd, b.nodes, b.children,
                                 b.invals, 10)
    rec = [O.recall(oh[q, :10]["node"], th[q, :10]["node"]) for q in range(40)]
    assert np.mean(rec) >= 0.95, np.mean(rec)
