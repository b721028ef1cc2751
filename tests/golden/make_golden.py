"""Writes tests/golden/reference_vectors.npz: outputs of the reference engine
(engine.cpp compiled from source under oracle/_ref, see oracle/Makefile) on
small seeded inputs, so the oracle and the host HNSW builder stay pinned
where the reference sources are not available. Inputs are stored with the
outputs (no regeneration needed).

Contents (all plain arrays):
  l2_<dim>_x, l2_<dim>_q, l2_<dim>_dist   BruteCursor distances of every row
                                          (float32, the reference l2sq order)
  b_base, b_cols, b_queries               brute-force data (float, 24 dims)
  b_<p>_node, b_<p>_dist, b_<p>_cnt       top-20 of BruteCursor per query for
                                          predicate p of PREDICATES
  h_base, h_l0, h_levels, h_upper_offs,   reference HNSW build, one thread
  h_upper, h_meta (M, efc, entry, top)
  h_queries, h_<p>_node, h_<p>_dist,      HnswCursor first ef hits per query
  h_<p>_cnt, h_entry                      and fv_index_graph_entry

Run: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from paves_amd import engine as E  # noqa: E402
from paves_amd import synth  # noqa: E402

L2_DIMS = (8, 17, 31, 77, 128, 130)
BRUTE_K = 20
HNSW_EF = 16


def predicates():
    """Predicate set of the brute and HNSW vectors (rebuilt by the test)."""
    return {
        "none": None,
        "lt": E.cmp(0, E.CMP_LT, 300.0),
        "in": E.isin(1, [1.0, 4.0, 9.0]),
        "ne_or": E.or_(E.cmp(1, E.CMP_NE, 3.0), E.cmp(0, E.CMP_GE, 950.0)),
        "not_and": E.not_(E.and_(E.cmp(0, E.CMP_GE, 100.0), E.cmp(0, E.CMP_LT, 800.0))),
    }


def _cols(n, seed):
    r = np.random.default_rng(seed)
    price = (r.random(n) * 1000.0).astype(np.float32).astype(np.float64)
    cat = r.integers(0, 16, n).astype(np.float64)
    cat[r.random(n) < 0.05] = np.nan          # NULLs
    return np.ascontiguousarray(np.stack([price, cat]))


def main():
    ref = O.RefEngine()
    out = {}
    tmp = tempfile.mkdtemp(prefix="paves_golden_")
    try:
        # distances: every row of a small index, non-integer data
        for dim in L2_DIMS:
            r = np.random.default_rng(1000 + dim)
            x = (r.standard_normal((40, dim)) * 3).astype(np.float32)
            q = (r.standard_normal((2, dim)) * 3).astype(np.float32)
            path = os.path.join(tmp, f"l2_{dim}.idx")
            ref.build(path, x, None, M=8, efc=16, threads=1)
            idx = ref.open(path)
            dist = np.zeros((2, 40), np.float32)
            for qi in range(2):
                for node, dd in ref.brute(idx, q[qi], None, 40, threads=1):
                    dist[qi, node] = np.float32(dd)
            ref.close(idx)
            out[f"l2_{dim}_x"], out[f"l2_{dim}_q"], out[f"l2_{dim}_dist"] = x, q, dist

        # filtered brute force
        n, dim, nq = 1500, 24, 8
        r = np.random.default_rng(77)
        base = (r.standard_normal((n, dim)) * 5).astype(np.float32)
        queries = (r.standard_normal((nq, dim)) * 5).astype(np.float32)
        cols = _cols(n, 78)
        path = os.path.join(tmp, "b.idx")
        ref.build(path, base, cols, M=8, efc=32, threads=1)
        idx = ref.open(path)
        out["b_base"], out["b_cols"], out["b_queries"] = base, cols, queries
        for name, p in predicates().items():
            ph, keep = ref.compile(idx, p)
            node = np.zeros((nq, BRUTE_K), np.uint64)
            dist = np.zeros((nq, BRUTE_K), np.float32)
            cnt = np.zeros(nq, np.int32)
            for qi in range(nq):
                hits = ref.brute(idx, queries[qi], ph, BRUTE_K, threads=1)
                cnt[qi] = len(hits)
                for k, (nd, dd) in enumerate(hits):
                    node[qi, k], dist[qi, k] = nd, np.float32(dd)
            if ph:
                ref.L.fv_pred_free(ph)
            out[f"b_{name}_node"], out[f"b_{name}_dist"], out[f"b_{name}_cnt"] = node, dist, cnt
        ref.close(idx)

        # HNSW: single-threaded build, cursor and entry point
        d = synth.gmm128(600, 6, seed=61)
        hb = np.ascontiguousarray(d["base"][:, :32])
        hq = np.ascontiguousarray(d["queries"][:, :32])
        hcols = _cols(600, 62)
        M, efc = 8, 32
        path = os.path.join(tmp, "h.idx")
        ref.build(path, hb, hcols, M=M, efc=efc, threads=1)
        f = O.read_index_file(path)
        idx = ref.open(path)
        out["h_base"], out["h_queries"], out["h_cols"] = hb, hq, hcols
        out["h_l0"], out["h_levels"] = f["l0"], f["levels"]
        out["h_upper_offs"], out["h_upper"] = f["upper_offs"], f["upper"]
        out["h_meta"] = np.array([M, efc, f["entry_node"], f["max_level"]], np.int64)
        out["h_entry"] = np.array([ref.graph_entry(idx, q) for q in hq], np.int64)
        for name, p in predicates().items():
            ph, keep = ref.compile(idx, p)
            node = np.zeros((len(hq), HNSW_EF), np.uint64)
            dist = np.zeros((len(hq), HNSW_EF), np.float32)
            cnt = np.zeros(len(hq), np.int32)
            for qi, q in enumerate(hq):
                hits = ref.hnsw(idx, q, ph, HNSW_EF, HNSW_EF)
                cnt[qi] = len(hits)
                for k, (nd, dd) in enumerate(hits):
                    node[qi, k], dist[qi, k] = nd, np.float32(dd)
            if ph:
                ref.L.fv_pred_free(ph)
            out[f"h_{name}_node"], out[f"h_{name}_dist"], out[f"h_{name}_cnt"] = node, dist, cnt
        ref.close(idx)
    finally:
        for fn in os.listdir(tmp):
            os.unlink(os.path.join(tmp, fn))
        os.rmdir(tmp)
    dst = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(dst, **out)
    print(f"wrote {dst} ({os.path.getsize(dst)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
