"""The learned per-query beam width / router: model handle, features,
training and prediction (bindings of fv_gpu_model_* in include/paves_gpu.h).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .engine import GpuError


def kmeans(x: np.ndarray, k: int = 256, iters: int = 10, sample: int = 50_000, seed: int = 7):
    """Centroids for the distance features: Lloyd iterations on a sample."""
    rng = np.random.default_rng(seed)
    xs = np.asarray(x[rng.choice(len(x), min(sample, len(x)), replace=False)], np.float64)
    c = xs[rng.choice(len(xs), k, replace=False)].copy()
    for _ in range(iters):
        d = (xs * xs).sum(1)[:, None] + (c * c).sum(1)[None, :] - 2 * xs @ c.T
        a = d.argmin(1)
        for j in range(k):
            m = a == j
            if m.any():
                c[j] = xs[m].mean(0)
    return c.astype(np.float32)


class Model:
    def __init__(self, dim: int, centroids: np.ndarray, device: int = 0):
        self.lib = L.load()
        self.dim = dim
        self.F = dim + 17
        self.np = self.lib.fv_gpu_model_nparams(dim)
        self.cent = np.ascontiguousarray(centroids, np.float32)
        eb = L.errbuf()
        h = self.lib.fv_gpu_model_create(device, dim, L.ptr(self.cent), self.cent.shape[0], eb,
                                         len(eb))
        if not h:
            raise GpuError(eb.value.decode())
        self.h = C.c_void_p(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fv_gpu_model_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, eb):
        if rc != 0:
            raise GpuError(eb.value.decode())

    def set_params(self, p, mean=None, std=None):
        eb = L.errbuf()
        p = np.ascontiguousarray(p, np.float32)
        m = None if mean is None else np.ascontiguousarray(mean, np.float32)
        s = None if std is None else np.ascontiguousarray(std, np.float32)
        self._chk(self.lib.fv_gpu_model_set_params(self.h, L.ptr(p), L.ptr(m), L.ptr(s), eb,
                                                   len(eb)), eb)

    def params(self):
        p = np.zeros(self.np, np.float32)
        m = np.zeros(self.F, np.float32)
        s = np.zeros(self.F, np.float32)
        eb = L.errbuf()
        self._chk(self.lib.fv_gpu_model_get_params(self.h, L.ptr(p), L.ptr(m), L.ptr(s), eb,
                                                   len(eb)), eb)
        return p, m, s

    def grad(self, feats, targets, lam=1.0):
        x = np.ascontiguousarray(feats, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        g = np.zeros(self.np, np.float32)
        loss = C.c_double()
        eb = L.errbuf()
        self._chk(self.lib.fv_gpu_model_grad(self.h, x.shape[0], L.ptr(x), L.ptr(t), lam,
                                             L.ptr(g), C.byref(loss), eb, len(eb)), eb)
        return g, loss.value

    def train(self, feats, targets, epochs=20, batch=1024, lr=2e-3, lam=1.0, seed=1):
        x = np.ascontiguousarray(feats, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        hist = np.zeros(epochs, np.float32)
        eb = L.errbuf()
        self._chk(self.lib.fv_gpu_model_train(self.h, x.shape[0], L.ptr(x), L.ptr(t), epochs,
                                              batch, lr, lam, seed, L.ptr(hist), eb, len(eb)), eb)
        return hist

    def predict(self, feats):
        x = np.ascontiguousarray(feats, np.float32)
        y = np.zeros((x.shape[0], 2), np.float32)
        eb = L.errbuf()
        self._chk(self.lib.fv_gpu_model_predict(self.h, x.shape[0], L.ptr(x), L.ptr(y), eb,
                                                len(eb)), eb)
        return y

    def batch_features(self, index, bh, nq):
        out = np.zeros((nq, self.F), np.float32)
        eb = L.errbuf()
        self._chk(self.lib.fv_gpu_batch_features(index.h, bh, self.h, L.ptr(out), eb, len(eb)),
                  eb)
        return out


# ------------------------------------------------------------------ pipeline

EF_GRID = (10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512)


def _recall(ids, truth) -> float:
    t = set(int(x) for x in truth)
    if not t:
        return 1.0
    return len(t & set(int(x) for x in ids)) / len(t)


def exact_truth(index, queries, sers, k, batch=1024, concat=None):
    """Exact top-k ids per query through the brute path (BF16 scan + exact
    FP32 re-rank; bit-identical to the reference cursor)."""
    from . import engine as E
    out = []
    for s in range(0, len(queries), batch):
        e = min(len(queries), s + batch)
        b = E.concat(sers[s:e], index.count)
        o, c, _ = index.brute(queries[s:e], b, k)
        out += [o[q, : c[q]]["node"].copy() for q in range(e - s)]
    return out


def ef_sweep(index, queries, sers, entries, truth, k=10, grid=EF_GRID, target=0.9,
             max_visited=50_000, batch=1024):
    """Smallest ef of `grid` whose graph result reaches recall@k >= target
    per query (-1: none within max_visited), and the nodes visited there."""
    from . import engine as E
    nq = len(queries)
    min_ef = np.full(nq, -1, np.int32)
    visited = np.zeros(nq, np.float64)
    for ef in grid:
        pend = np.flatnonzero(min_ef < 0)
        if len(pend) == 0:
            break
        for s in range(0, len(pend), batch):
            idx = pend[s: s + batch]
            b = E.concat([sers[i] for i in idx], index.count, entries=entries[idx])
            o, c, _, dbg = index.graph(queries[idx], b, ef, max_visited, debug=True)
            for j, qi in enumerate(idx):
                if _recall(o[j, : min(k, c[j])]["node"], truth[qi]) >= target:
                    min_ef[qi] = ef
                    visited[qi] = dbg[j, 1]
    return min_ef, visited


def service_costs(index, queries, sers, entries, k=10, ef=64, micro=64):
    """Measured per-query service time of the brute path (batches of
    `micro` queries, device-resident, overlapped as in production) and per
    visited node of the graph search (a batch of up to 1024 queries): the
    cost model that labels the route, for routed runs over batches of
    `micro` queries."""
    import time
    from . import engine as E
    from . import _lib as LL
    nb = max(1, min(len(queries), 4096) // micro)
    staged = []
    for i in range(nb):
        sl = slice(i * micro, (i + 1) * micro)
        staged.append(index.stage(queries[sl], E.concat(sers[sl], index.count), k))
    bo = LL.BruteOpts()
    for bh in staged:
        index.run_brute(bh, bo)
    index.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        for bh in staged:
            index.run_brute(bh, bo)
    index.sync()
    t_brute = (time.perf_counter() - t0) / (3 * nb * micro)
    for bh in staged:
        index.free_batch(bh)
    # graph throughput per visited node: the search kernel alone (HIP
    # events), a full batch of unfiltered queries, so the figure is the
    # kernel's rate with the chip occupied rather than the latency tail of a
    # few long filtered searches
    m = min(len(queries), 1024)
    b = E.concat([E.serialize(None, index.count)] * m, index.count, entries=entries[:m])
    index.graph(queries[:m], b, ef, 200_000)
    index.set_timing(True)
    _, _, _, dbg = index.graph(queries[:m], b, ef, 200_000, debug=True)
    t_graph = index.stats().t_graph_us * 1e-6
    index.set_timing(False)
    t_node = t_graph / max(1.0, float(dbg[:, 1].sum()))
    return t_brute, t_node


def targets_from_labels(min_ef, visited, t_brute, t_node):
    """[log ef, route, ef valid]: route 1 (graph) when the query has a
    graph label and its visited-node cost is below the brute cost."""
    valid = min_ef > 0
    log_ef = np.where(valid, np.log(np.maximum(min_ef, 1)), 0.0)
    route = valid & (visited * t_node < t_brute)
    return np.stack([log_ef, route.astype(np.float64), valid.astype(np.float64)], 1).astype(
        np.float32)


def features_of(model, index, queries, sers, batch=1024):
    from . import engine as E
    out = []
    for s in range(0, len(queries), batch):
        e = min(len(queries), s + batch)
        bh = index.stage(queries[s:e], E.concat(sers[s:e], index.count), 1)
        try:
            out.append(model.batch_features(index, bh, e - s))
        finally:
            index.free_batch(bh)
    return np.concatenate(out, 0)


def run_mode(index, model, queries, sers, entries, truth, opts, micro=64, reps=1,
             groups=None):
    """Routed runs over `queries` in micro-batches (device-resident,
    overlapped on the index's streams); returns (QPS, mean recall@k,
    graph-routed fraction, mean ef of graph-routed queries). With `groups`
    (a group id per query) the recall is the lowest of the groups' means."""
    import time
    from . import engine as E
    k = opts.k
    staged = []
    for s in range(0, len(queries), micro):
        e = min(len(queries), s + micro)
        b = E.concat(sers[s:e], index.count, entries=entries[s:e])
        staged.append((s, e, index.stage(queries[s:e], b, k)))
    try:
        for s, e, bh in staged:
            index.run_learned(bh, model, opts)
        index.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            for s, e, bh in staged:
                index.run_learned(bh, model, opts)
        index.sync()
        dt = (time.perf_counter() - t0) / reps
        rec, ng, efs = [], 0, []
        for s, e, bh in staged:
            o, c, f, rc = index.fetch(bh, e - s, k)
            r, ef = index.routes(bh, e - s)
            for j in range(e - s):
                rec.append(_recall(o[j, : c[j]]["node"], truth[s + j]))
                if r[j]:
                    ng += 1
                    efs.append(int(ef[j]))
        rec = np.asarray(rec)
        if groups is None:
            r = float(np.mean(rec))
        else:
            gid = np.asarray(groups)
            r = min(float(np.mean(rec[gid == u])) for u in np.unique(gid))
        return (len(queries) / dt, r, ng / len(queries), float(np.mean(efs)) if efs else 0.0)
    finally:
        for s, e, bh in staged:
            index.free_batch(bh)


def calibrate_scale(index, model, queries, sers, entries, truth, k=10, target=0.9,
                    scales=(1.0, 1.25, 1.5, 2.0, 3.0, 4.0), mode=1, micro=64):
    """Smallest safety factor on the predicted ef whose mean recall over
    (queries) reaches the target in the given mode."""
    from . import _lib as LL
    for s in scales:
        _, rec, _, _ = run_mode(index, model, queries, sers, entries, truth,
                                LL.LearnedOpts(k=k, mode=mode, ef_scale=s, max_visited=50_000),
                                micro=micro)
        if rec >= target:
            return s
    return scales[-1]


def train_router(index, graph, base, queries, sers, k=10, target=0.9, epochs=30, batch=1024,
                 lr=2e-3, centroids=None, log=None, micro=64):
    """End to end: exact truth, ef labels, cost-model routes, features,
    GPU training. Returns (model, info)."""
    import time
    from . import hnsw
    t0 = time.perf_counter()
    ent = hnsw.entries(graph, base, queries)
    truth = exact_truth(index, queries, sers, k)
    t1 = time.perf_counter()
    min_ef, visited = ef_sweep(index, queries, sers, ent, truth, k=k, target=target)
    t2 = time.perf_counter()
    t_b, t_v = service_costs(index, queries, sers, ent, k=k, micro=micro)
    tg = targets_from_labels(min_ef, visited, t_b, t_v)
    cent = centroids if centroids is not None else kmeans(base, 256, iters=6, sample=30_000)
    model = Model(base.shape[1], cent)
    feats = features_of(model, index, queries, sers)
    t3 = time.perf_counter()
    hist = model.train(feats, tg, epochs=epochs, batch=min(batch, len(queries)), lr=lr)
    t4 = time.perf_counter()
    info = {"n_train": len(queries), "truth_s": t1 - t0, "labels_s": t2 - t1,
            "features_s": t3 - t2, "train_s": t4 - t3, "loss_first": float(hist[0]),
            "loss_last": float(hist[-1]), "ef_valid": float(tg[:, 2].mean()),
            "route_graph": float(tg[:, 1].mean()), "t_brute_us": t_b * 1e6,
            "t_node_ns": t_v * 1e9,
            "median_label_ef": float(np.median(min_ef[min_ef > 0])) if (min_ef > 0).any() else 0}
    if log:
        log(f"router trained: {info}")
    return model, info


def calibrate_bias(index, model, queries, sers, entries, truth, k=10, target=0.9, scale=1.0,
                   micro=1024, fracs=(0.0, 0.05, 0.1, 0.2, 0.3, 0.5), reps=3, groups=None,
                   scales=None):
    """The route threshold (graph when the route logit exceeds it) and ef
    safety factor that give the highest measured throughput of routed runs
    over (queries) at recall@k >= target — the mean over all queries, or
    with `groups` (a filter-class id per query) every group's mean.
    Candidates send the given fractions of these queries to the graph (0:
    all brute, the fallback, which needs no factor) at each factor in
    `scales` (default: `scale` alone). Returns (bias, scale, table of
    (fraction, scale, bias, QPS, recall, graph-routed fraction))."""
    from . import _lib as LL
    y = model.predict(features_of(model, index, queries, sers))
    table, best = [], None
    for f in fracs:
        bias = float("inf") if f == 0.0 else float(np.quantile(y[:, 1], 1.0 - f))
        for sc in ((scale,) if f == 0.0 or not scales else scales):
            qps, rec, gfrac, _ = run_mode(index, model, queries, sers, entries, truth,
                                          LL.LearnedOpts(k=k, mode=0, ef_scale=sc,
                                                         route_bias=bias, max_visited=50_000),
                                          micro=micro, reps=reps, groups=groups)
            table.append((f, sc, bias, qps, rec, gfrac))
            if rec >= target and (best is None or qps > best[3]):
                best = table[-1]
    if best is None:
        return float("inf"), scale, table
    return best[2], best[1], table
