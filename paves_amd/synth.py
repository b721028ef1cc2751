"""Seeded synthetic workloads for the filtered-search query path.

Base 128-d data (used at 100k and 1M rows): a 256-centre Gaussian mixture,
centres ~ U[0,120]^128 with equal weights, x = c + N(0, 20^2 I), clipped to
[0, 218] and rounded to integers (stored fp32). Attributes: price ~
U[0,1000) (fp32 values), category ~ U{0..31}. Queries come from the same
mixture and keep their centre index (the "correlated" filter class uses
category == centre mod 32).

Because every coordinate is an integer in [0, 218], every squared distance,
dot product and norm of 128-d rows is an integer below 2^24 and therefore
exact in FP32 and in BF16 x BF16 -> FP32 arithmetic.

768-d data: a 1024-centre mixture with centres ~ N(0, I) normalised to unit
length, x = c + N(0, 0.15^2 I) per coordinate, then normalised and scaled
by 10 (values are not integers; FP32 sums depend on summation order).

Seeds: base seed 20240617 split into independent streams per array.
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 20240617
NCOLS = 2          # column slots: 0 = price, 1 = category
COL_PRICE = 0
COL_CATEGORY = 1


def _rngs(seed: int, n: int):
    ss = np.random.SeedSequence(seed)
    return [np.random.default_rng(s) for s in ss.spawn(n)]


def gmm128(n: int, nq: int, seed: int = BASE_SEED, ncent: int = 256,
           dim: int = 128, chunk: int = 65536):
    """Returns dict: base (n,dim) f32, queries (nq,dim) f32, price (n,) f32,
    category (n,) i32, qcent (nq,) i32, bcent (n,) i32, centers."""
    r_c, r_b, r_bn, r_q, r_qn, r_p, r_cat = _rngs(seed, 7)
    centers = r_c.uniform(0.0, 120.0, size=(ncent, dim)).astype(np.float32)

    def draw(rng_assign, rng_noise, m):
        out = np.empty((m, dim), dtype=np.float32)
        cid = rng_assign.integers(0, ncent, size=m, dtype=np.int32)
        for s in range(0, m, chunk):
            e = min(m, s + chunk)
            x = centers[cid[s:e]] + rng_noise.standard_normal((e - s, dim), dtype=np.float32) * 20.0
            np.clip(x, 0.0, 218.0, out=x)
            np.rint(x, out=x)
            out[s:e] = x
        return out, cid

    base, bcent = draw(r_b, r_bn, n)
    queries, qcent = draw(r_q, r_qn, nq)
    price = (r_p.random(n, dtype=np.float32) * np.float32(1000.0)).astype(np.float32)
    price = np.minimum(price, np.nextafter(np.float32(1000.0), np.float32(0.0)))
    category = r_cat.integers(0, 32, size=n, dtype=np.int32)
    return {"base": base, "queries": queries, "price": price, "category": category,
            "qcent": qcent, "bcent": bcent, "centers": centers}


def more_queries(data, m: int, seed: int):
    """m further queries from the mixture of `data` (same centres, sigma 20,
    clip and rounding as gmm128), drawn from an independent seed: training
    sets for the learned router that never overlap the evaluated queries."""
    r_a, r_n = _rngs(seed, 2)
    centers = data["centers"]
    ncent, dim = centers.shape
    cid = r_a.integers(0, ncent, size=m, dtype=np.int32)
    out = np.empty((m, dim), dtype=np.float32)
    for s in range(0, m, 65536):
        e = min(m, s + 65536)
        x = centers[cid[s:e]] + r_n.standard_normal((e - s, dim), dtype=np.float32) * 20.0
        np.clip(x, 0.0, 218.0, out=x)
        np.rint(x, out=x)
        out[s:e] = x
    return out, cid


def gmm768(n: int, nq: int, seed: int = BASE_SEED + 768, ncent: int = 1024,
           dim: int = 768, sigma: float = 0.15, chunk: int = 32768):
    r_c, r_b, r_bn, r_q, r_qn, r_p, r_cat = _rngs(seed, 7)
    centers = r_c.standard_normal((ncent, dim), dtype=np.float32)
    centers /= np.linalg.norm(centers, axis=1, keepdims=True)

    def draw(rng_assign, rng_noise, m):
        out = np.empty((m, dim), dtype=np.float32)
        cid = rng_assign.integers(0, ncent, size=m, dtype=np.int32)
        for s in range(0, m, chunk):
            e = min(m, s + chunk)
            x = centers[cid[s:e]] + rng_noise.standard_normal((e - s, dim), dtype=np.float32) * np.float32(sigma)
            x /= np.linalg.norm(x, axis=1, keepdims=True)
            x *= np.float32(10.0)
            out[s:e] = x
        return out, cid

    base, bcent = draw(r_b, r_bn, n)
    queries, qcent = draw(r_q, r_qn, nq)
    price = (r_p.random(n, dtype=np.float32) * np.float32(1000.0)).astype(np.float32)
    price = np.minimum(price, np.nextafter(np.float32(1000.0), np.float32(0.0)))
    category = r_cat.integers(0, 32, size=n, dtype=np.int32)
    return {"base": base, "queries": queries, "price": price, "category": category,
            "qcent": qcent, "bcent": bcent, "centers": centers}


def columns(data) -> np.ndarray:
    """Column-major f64 scalar zone [ncols][n] (engine.cpp:207 layout)."""
    n = data["base"].shape[0]
    cols = np.empty((NCOLS, n), dtype=np.float64)
    cols[COL_PRICE] = data["price"].astype(np.float64)
    cols[COL_CATEGORY] = data["category"].astype(np.float64)
    return cols


def postings(cols: np.ndarray):
    """Per column: rows sorted by value (NaN last, ties by row) and the
    sorted values (engine.cpp:939-957)."""
    ncols, n = cols.shape
    rows = np.empty((ncols, n), dtype=np.uint32)
    vals = np.empty((ncols, n), dtype=np.float64)
    for c in range(ncols):
        v = cols[c]
        key = np.where(np.isnan(v), np.inf, v)
        nanflag = np.isnan(v).astype(np.int8)
        order = np.lexsort((np.arange(n), key, nanflag))
        rows[c] = order.astype(np.uint32)
        vals[c] = v[order]
    return rows, vals


# six filter classes of the 1M x 128 workload
FILTER_CLASSES = ["all", "price<500", "price<100", "cat_in_3_7", "price<1", "cat_eq_centre"]


def filter_pred(cls: str, qcent: int = 0):
    """Predicate tree of one filter class (None = no filter)."""
    from . import engine as E
    if cls == "all":
        return None
    if cls == "price<500":
        return E.cmp(COL_PRICE, E.CMP_LT, 500.0)
    if cls == "price<100":
        return E.cmp(COL_PRICE, E.CMP_LT, 100.0)
    if cls == "cat_in_3_7":
        return E.isin(COL_CATEGORY, [3.0, 7.0])
    if cls == "price<1":
        return E.cmp(COL_PRICE, E.CMP_LT, 1.0)
    if cls == "cat_eq_centre":
        return E.cmp(COL_CATEGORY, E.CMP_EQ, float(int(qcent) % 32))
    raise ValueError(cls)
