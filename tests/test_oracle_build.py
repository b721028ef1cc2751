"""Pins the oracle's restatement of the batched HNSW builder
(or_hnsw_batch_insert) on the CPU, before the device builder is checked
against it.

* Batches of one node reduce the batch semantics to the sequential
  builder: from the host phase's graph (node 0 and every node above level
  0), or_hnsw_batch_insert with every batch of size 1 must give, bit for
  bit, the library's host insertion of the remaining nodes in index order
  (fv_hnsw_build_seeded with rest_on_host, one thread) — the host builder
  whose single-threaded graph equals the reference builder's
  (test_golden.py). This pins the restated beam, the heuristic (new-node
  lists and full-list re-pruning) and the append rule at once.
* The batch schedule: sizes cover the remaining nodes, start at
  inserted / batch_div and never exceed max_batch.
* With real batches the result still satisfies the graph's invariants (no
  self links or duplicates, ids in range, counts <= M0) and its searches
  keep their recall close to the sequential graph's.
"""
import numpy as np
import pytest

import oracle as O
from paves_amd import hnsw
from paves_amd import synth


def _data(n, dim, seed):
    d = synth.gmm768(n, 64, seed=seed, dim=dim, ncent=24)
    return d["base"], d["queries"]


def _rest_entries(g, x):
    lev = g["levels"]
    rest = np.array([i for i in range(1, len(lev)) if lev[i] == 0], np.uint32)
    ents = np.array([O.graph_entry(x, g["upper_offs"], g["upper"], g["M"], g["entry_node"],
                                   g["max_level"], x[i]) for i in rest], np.uint32)
    return rest, ents


def _lists(l0):
    return [tuple(r[1:1 + r[0]].tolist()) for r in l0]


@pytest.mark.parametrize("n,dim,M,efc", [(1500, 24, 8, 40), (1200, 37, 4, 16), (900, 64, 16, 60)])
def test_batches_of_one_equal_sequential_host_insertion(n, dim, M, efc):
    x, _ = _data(n, dim, 7 + dim)
    seed = hnsw.build_seeded(x, M, efc, threads=1)
    host = hnsw.build_seeded(x, M, efc, threads=1, rest_on_host=True)
    # the upper levels and the seed nodes' state come from the same host phase
    assert np.array_equal(seed["levels"], host["levels"])
    assert np.array_equal(seed["upper_offs"], host["upper_offs"])
    assert seed["entry_node"] == host["entry_node"] and seed["max_level"] == host["max_level"]
    rest, ents = _rest_entries(seed, x)
    assert len(rest) > n // 2
    # before the remaining nodes go in, they have no lists
    assert all(seed["l0"][i, 0] == 0 for i in rest)
    got = O.hnsw_batch_insert(x, seed["l0"], rest, ents, [1] * len(rest), M, efc)
    assert _lists(got) == _lists(host["l0"])
    # the upper levels are final after the host phase
    assert np.array_equal(seed["upper"], host["upper"])


def test_schedule():
    s = O.build_schedule(100_000, 93_000, 8, 4096)
    assert sum(s) == 93_000
    assert s[0] == 7_000 // 8
    assert max(s) == 4096 and min(s) >= 1
    assert O.build_schedule(10, 0, 8, 4096) == []
    assert O.build_schedule(10, 9, 8, 4096) == [1] * 9


def test_batched_invariants_and_recall():
    n, dim, M, efc = 4000, 32, 8, 64
    x, qs = _data(n, dim, 3)
    seed = hnsw.build_seeded(x, M, efc, threads=1)
    host = hnsw.build_seeded(x, M, efc, threads=1, rest_on_host=True)
    rest, ents = _rest_entries(seed, x)
    sizes = O.build_schedule(n, len(rest), 4, 512)
    assert len(sizes) > 5 and max(sizes) > 100
    l0 = O.hnsw_batch_insert(x, seed["l0"], rest, ents, sizes, M, efc)
    for v, r in enumerate(l0):
        c = int(r[0])
        ids = r[1:1 + c]
        assert c <= 2 * M
        assert len(set(ids.tolist())) == c and v not in ids and (ids < n).all()
    assert (l0[:, 0] > 0).all()
    # recall@10 of the oracle cursor at ef 40 on both graphs
    truth = np.argsort(((qs[:, None, :] - x[None, :, :]) ** 2).sum(-1), axis=1)[:, :10]

    def rec(g):
        ents_q = [O.graph_entry(x, g["upper_offs"], g["upper"], g["M"], g["entry_node"],
                                g["max_level"], q) for q in qs]
        hit = 0
        for qi, q in enumerate(qs):
            h, _, _ = O.hnsw_one(x, g["l0"], None, q, None, None, None, -1, ents_q[qi], 40, n + 1)
            hit += len(set(h[:10]["node"].tolist()) & set(truth[qi].tolist()))
        return hit / truth.size

    rb = rec(dict(seed, l0=l0))
    rh = rec(host)
    assert rb >= rh - 0.02, (rb, rh)
