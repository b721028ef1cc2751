"""Host-side glue around the C ABI: predicates, driver spans, batch concat.

Mirrors the reference engine's host half of the GPU path so that tests and
the benchmark can drive ``libpaves_amd.so`` exactly as ``engine.cpp`` does:

* ``Pred``/``cmp``/``isin``/``and_``/``or_``/``not_`` — the predicate tree of
  ``FvPredNode`` (engine_api.h:76-85);
* ``Driver`` — ``make_driver`` (engine.cpp:464-599): posting-list spans,
  ``covers`` and ``may_dup``;
* ``serialize`` — ``fv_pred_serialize`` (engine.cpp:1168-1232): node 0 is
  the root, children/values in side arrays, ``need_eval``/``dedup`` flags;
* ``concat`` — ``GpuConcat::build`` (engine.cpp:1241-1281): per-query slices
  rebased into one set of arrays plus ``FvGpuQueryDesc``;
* ``GpuIndex`` — the resident index and the batch entry points.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import (CMP_EQ, CMP_GE, CMP_GT, CMP_LE, CMP_LT, CMP_NE, NODE_AND, NODE_CMP,
                   NODE_IN, NODE_NOT, NODE_OR, HIT_DT, PRED_DT, QD_DT, SPAN_DT, ptr)


# ---------------------------------------------------------------- predicates

@dataclass
class Pred:
    kind: int
    col: int = 0
    op: int = 0
    value: float = 0.0
    values: tuple = ()
    children: tuple = ()


def cmp(col, op, value):
    return Pred(NODE_CMP, col=col, op=op, value=float(value))


def isin(col, values):
    return Pred(NODE_IN, col=col, values=tuple(float(v) for v in values))


def and_(*c):
    return Pred(NODE_AND, children=tuple(c))


def or_(*c):
    return Pred(NODE_OR, children=tuple(c))


def not_(c):
    return Pred(NODE_NOT, children=(c,))


def eval_pred(p: Pred, cols: np.ndarray) -> np.ndarray:
    """Vectorised FvPred::eval_node (engine.cpp:397-430) over all rows."""
    if p.kind == NODE_CMP:
        v = cols[p.col]
        with np.errstate(invalid="ignore"):
            if p.op == CMP_LT: return v < p.value
            if p.op == CMP_LE: return v <= p.value
            if p.op == CMP_EQ: return v == p.value
            if p.op == CMP_GE: return v >= p.value
            if p.op == CMP_GT: return v > p.value
            return (v == v) & (v != p.value)
    if p.kind == NODE_IN:
        v = cols[p.col]
        out = np.zeros(v.shape, dtype=bool)
        for x in p.values:
            out |= v == x
        return out
    if p.kind == NODE_AND:
        out = np.ones(cols.shape[1], dtype=bool)
        for c in p.children:
            out &= eval_pred(c, cols)
        return out
    if p.kind == NODE_OR:
        out = np.zeros(cols.shape[1], dtype=bool)
        for c in p.children:
            out |= eval_pred(c, cols)
        return out
    return ~eval_pred(p.children[0], cols)


# ---------------------------------------------------------------- driver spans

class Postings:
    """Per-column posting lists: rows sorted by value, NaN last."""

    def __init__(self, rows: np.ndarray, vals: np.ndarray):
        self.rows = np.ascontiguousarray(rows, dtype=np.uint32)
        self.vals = vals
        self.nan_start = [int(np.searchsorted(np.isnan(vals[c]), True)) for c in range(vals.shape[0])]

    def lower(self, col, v):
        return int(np.searchsorted(self.vals[col][: self.nan_start[col]], v, side="left"))

    def upper(self, col, v):
        return int(np.searchsorted(self.vals[col][: self.nan_start[col]], v, side="right"))


@dataclass
class Driver:
    ok: bool = False
    covers: bool = False
    may_dup: bool = False
    size: int = 0
    spans: list = field(default_factory=list)


def _cmp_range(post: Postings, col, op, value):
    fin = post.nan_start[col]
    if op == CMP_LT: return 0, post.lower(col, value)
    if op == CMP_LE: return 0, post.upper(col, value)
    if op == CMP_EQ: return post.lower(col, value), post.upper(col, value)
    if op == CMP_GE: return post.lower(col, value), fin
    if op == CMP_GT: return post.upper(col, value), fin
    return None


def make_driver(post: Postings, p: Pred) -> Driver:
    """make_driver (engine.cpp:464-599)."""
    d = Driver()
    if p.kind == NODE_CMP:
        if math.isnan(p.value):
            if p.op == CMP_EQ:
                d.ok, d.covers, d.size = True, True, 0
            return d
        r = _cmp_range(post, p.col, p.op, p.value)
        if r is None:
            return d
        lo, hi = r
        d.ok, d.covers = True, True
        if hi > lo:
            d.spans.append((p.col, lo, hi))
            d.size = hi - lo
        return d
    if p.kind == NODE_IN:
        d.ok, d.covers = True, True
        for v in p.values:
            if math.isnan(v):
                continue
            lo, hi = post.lower(p.col, v), post.upper(p.col, v)
            if hi > lo:
                d.spans.append((p.col, lo, hi))
                d.size += hi - lo
        return d
    if p.kind == NODE_AND:
        best = Driver()
        for c in p.children:
            cd = make_driver(post, c)
            if cd.ok and (not best.ok or cd.size < best.size):
                best = cd
        if best.ok and len(p.children) > 1:
            best = Driver(best.ok, False, best.may_dup, best.size, list(best.spans))
        percol = {}
        order = []
        for c in p.children:
            if c.kind != NODE_CMP or math.isnan(c.value):
                continue
            r = _cmp_range(post, c.col, c.op, c.value)
            if r is None:
                continue
            if c.col not in percol:
                percol[c.col] = [0, 2 ** 64 - 1, 0]
                order.append(c.col)
            e = percol[c.col]
            e[0] = max(e[0], r[0])
            e[1] = min(e[1], r[1])
            e[2] += 1
        for col in order:
            lo, hi, nconj = percol[col]
            if nconj < 2:
                continue
            cd = Driver(True, True, False, 0, [])
            if hi > lo:
                cd.spans.append((col, lo, hi))
                cd.size = hi - lo
            if not best.ok or cd.size < best.size:
                best = cd
                best.covers = nconj == len(p.children)
        return best
    if p.kind == NODE_OR:
        d.ok, d.covers = True, True
        for c in p.children:
            cd = make_driver(post, c)
            if not cd.ok:
                return Driver()
            d.spans.extend(cd.spans)
            d.size += cd.size
            d.covers = d.covers and cd.covers
            d.may_dup = d.may_dup or cd.may_dup
        if len(p.children) > 1:
            d.may_dup = True
        return d
    return d


# ---------------------------------------------------------------- serialization

@dataclass
class Serialized:
    nodes: np.ndarray
    children: np.ndarray
    invals: np.ndarray
    spans: np.ndarray
    need_eval: int
    dedup: int
    total: int


def _flatten(p: Pred, nodes: list, childlists: list):
    """copy_tree order (engine.cpp:610-630): node first, then children."""
    ni = len(nodes)
    nodes.append(p)
    childlists.append([])
    if p.kind in (NODE_AND, NODE_OR, NODE_NOT):
        kids = [_flatten(c, nodes, childlists) for c in p.children]
        childlists[ni] = kids
    return ni


def serialize(p: Pred | None, count: int, post: Postings | None = None,
              max_spans: int = L.FV_GPU_MAX_SPANS) -> Serialized:
    """fv_pred_serialize (engine.cpp:1168-1232)."""
    empty_nodes = np.zeros(0, dtype=PRED_DT)
    if p is None:
        return Serialized(empty_nodes, np.zeros(0, np.int32), np.zeros(0, np.float64),
                          np.zeros(0, dtype=SPAN_DT), 0, 0, count)
    nodes, kids = [], []
    _flatten(p, nodes, kids)
    arr = np.zeros(len(nodes), dtype=PRED_DT)
    children, invals = [], []
    for i, n in enumerate(nodes):
        arr[i]["kind"] = n.kind
        arr[i]["col"] = n.col
        arr[i]["op"] = n.op
        arr[i]["value"] = n.value
        arr[i]["nchildren"] = len(kids[i])
        arr[i]["child0"] = len(children)
        arr[i]["nvalues"] = len(n.values) if n.kind == NODE_IN else 0
        arr[i]["val0"] = len(invals)
        children.extend(kids[i])
        if n.kind == NODE_IN:
            invals.extend(n.values)
    drv = make_driver(post, p) if post is not None else Driver()
    need_eval = int((not drv.ok) or (not drv.covers) or drv.may_dup)
    dedup = int(drv.may_dup)
    spans = np.zeros(0, dtype=SPAN_DT)
    total = count
    if drv.ok and len(drv.spans) <= max_spans:
        spans = np.zeros(len(drv.spans), dtype=SPAN_DT)
        for i, (c, lo, hi) in enumerate(drv.spans):
            spans[i] = (c, lo, hi)
        total = int(sum(hi - lo for _, lo, hi in drv.spans))
    elif drv.ok:
        need_eval = 1
    return Serialized(arr, np.asarray(children, np.int32), np.asarray(invals, np.float64),
                      spans, need_eval, dedup, total)


@dataclass
class Batch:
    nq: int
    qd: np.ndarray
    dedup: np.ndarray
    nodes: np.ndarray
    children: np.ndarray
    invals: np.ndarray
    spans: np.ndarray
    spanoff: np.ndarray


def concat(sers: list, count: int, entries=None) -> Batch:
    """GpuConcat::build (engine.cpp:1241-1281)."""
    nq = len(sers)
    qd = np.zeros(nq, dtype=QD_DT)
    dedup = np.zeros(nq, dtype=np.uint8)
    nodes, children, invals, spans, soff = [], [], [], [], []
    nn = nc = nv = ns = 0
    for q, s in enumerate(sers):
        base_n, base_c, base_v = nn, nc, nv
        qd[q]["proot"] = base_n if len(s.nodes) > 0 else -1
        qd[q]["need_eval"] = s.need_eval
        qd[q]["entry"] = 0 if entries is None else int(entries[q])
        dedup[q] = s.dedup
        if len(s.nodes):
            a = s.nodes.copy()
            a["child0"] += base_c
            a["val0"] += base_v
            nodes.append(a)
            nn += len(a)
        if len(s.children):
            children.append(s.children + base_n)
            nc += len(s.children)
        if len(s.invals):
            invals.append(s.invals)
            nv += len(s.invals)
        qd[q]["span0"] = ns
        qd[q]["nspans"] = len(s.spans)
        qd[q]["soff0"] = len(soff)
        if len(s.spans) > 0:
            tot = 0
            for sp in s.spans:
                soff.append(tot)
                tot += int(sp["hi"]) - int(sp["lo"])
            soff.append(tot)
            spans.append(s.spans)
            ns += len(s.spans)
            qd[q]["total"] = tot
        else:
            qd[q]["soff0"] = 0
            qd[q]["total"] = count
    cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dtype=dt)
    return Batch(nq, qd, dedup, cat(nodes, PRED_DT), cat(children, np.int32),
                 cat(invals, np.float64), cat(spans, SPAN_DT), np.asarray(soff, np.uint64))


# ---------------------------------------------------------------- GPU index

class GpuError(RuntimeError):
    pass


class GpuIndex:
    """A GPU-resident index (fv_gpu_index_upload) and its batch calls."""

    def __init__(self, vectors: np.ndarray, cols: np.ndarray | None = None,
                 post_rows: np.ndarray | None = None, l0: np.ndarray | None = None,
                 precision: int = L.FV_GPU_PREC_BF16, device: int = -1):
        self.lib = L.load()
        self.vectors = np.ascontiguousarray(vectors, dtype=np.float32)
        self.count, self.dim = self.vectors.shape
        self.cols = None if cols is None else np.ascontiguousarray(cols, dtype=np.float64)
        self.ncols = 0 if cols is None else self.cols.shape[0]
        self.post_rows = None if post_rows is None else np.ascontiguousarray(post_rows, dtype=np.uint32)
        l0c = None if l0 is None else np.ascontiguousarray(l0, dtype=np.uint32)
        stride = 0 if l0c is None else l0c.shape[1]
        eb = L.errbuf()
        h = self.lib.fv_gpu_index_upload(ptr(self.vectors), ptr(self.cols), ptr(self.post_rows),
                                         ptr(l0c), stride, self.count, self.dim, self.ncols,
                                         device, precision, eb, len(eb))
        if not h:
            raise GpuError(eb.value.decode())
        self.h = C.c_void_p(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fv_gpu_index_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def precision(self):
        return self.lib.fv_gpu_index_precision(self.h)

    def set_timing(self, on=True):
        self.lib.fv_gpu_set_timing(self.h, 1 if on else 0)

    def stats(self) -> L.Stats:
        st = L.Stats()
        self.lib.fv_gpu_last_stats(self.h, C.byref(st))
        return st

    def feedback(self) -> L.Feedback:
        fb = L.Feedback()
        if self.lib.fv_gpu_feedback(self.h, C.byref(fb)) != 0:
            raise GpuError("gpu: feedback")
        return fb

    def brute(self, queries: np.ndarray, b: Batch, topb: int, opts: L.BruteOpts | None = None):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        out = np.zeros((b.nq, topb), dtype=HIT_DT)
        cnt = np.zeros(b.nq, dtype=np.uint32)
        trunc = np.zeros(b.nq, dtype=np.uint8)
        eb = L.errbuf()
        args = (self.h, b.nq, ptr(q), ptr(b.qd), ptr(b.dedup), ptr(b.nodes), len(b.nodes),
                ptr(b.children), len(b.children), ptr(b.invals), len(b.invals), ptr(b.spans),
                len(b.spans), ptr(b.spanoff), len(b.spanoff), topb)
        if opts is None:
            rc = self.lib.fv_gpu_brute_batch(*args, ptr(out), ptr(cnt), ptr(trunc), eb, len(eb))
        else:
            rc = self.lib.fv_gpu_brute_batch_ex(*args, C.byref(opts), ptr(out), ptr(cnt),
                                                ptr(trunc), eb, len(eb))
        if rc != 0:
            raise GpuError(eb.value.decode())
        return out, cnt, trunc

    def graph(self, queries: np.ndarray, b: Batch, ef: int, max_visited: int,
              wmax: int | None = None, fcap: int = 1024, debug: bool = False):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        out = np.zeros((b.nq, ef), dtype=HIT_DT)
        cnt = np.zeros(b.nq, dtype=np.uint32)
        exh = np.zeros(b.nq, dtype=np.uint8)
        dbg = np.zeros((b.nq, 4), dtype=np.uint32) if debug else None
        eb = L.errbuf()
        args = (self.h, b.nq, ptr(q), ptr(b.qd), ptr(b.nodes), len(b.nodes), ptr(b.children),
                len(b.children), ptr(b.invals), len(b.invals), ef, max_visited)
        if wmax is None and not debug:
            rc = self.lib.fv_gpu_graph_batch(*args, ptr(out), ptr(cnt), ptr(exh), eb, len(eb))
        else:
            rc = self.lib.fv_gpu_graph_batch_ex(*args, wmax or 8, fcap, ptr(out), ptr(cnt),
                                                ptr(exh), ptr(dbg), eb, len(eb))
        if rc != 0:
            raise GpuError(eb.value.decode())
        return (out, cnt, exh, dbg) if debug else (out, cnt, exh)

    # device-resident batches
    def stage(self, queries: np.ndarray, b: Batch, topb: int):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        eb = L.errbuf()
        h = self.lib.fv_gpu_batch_stage(self.h, b.nq, ptr(q), ptr(b.qd), ptr(b.dedup),
                                        ptr(b.nodes), len(b.nodes), ptr(b.children),
                                        len(b.children), ptr(b.invals), len(b.invals),
                                        ptr(b.spans), len(b.spans), ptr(b.spanoff),
                                        len(b.spanoff), topb, eb, len(eb))
        if not h:
            raise GpuError(eb.value.decode())
        return C.c_void_p(h)

    def stage_ex(self, queries: np.ndarray, b: Batch, k_per_query: np.ndarray):
        """Uneven batch: per-query k (outputs have stride max k)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        kq = np.ascontiguousarray(k_per_query, dtype=np.uint32)
        eb = L.errbuf()
        h = self.lib.fv_gpu_batch_stage_ex(self.h, b.nq, ptr(q), ptr(b.qd), ptr(b.dedup),
                                           ptr(b.nodes), len(b.nodes), ptr(b.children),
                                           len(b.children), ptr(b.invals), len(b.invals),
                                           ptr(b.spans), len(b.spans), ptr(b.spanoff),
                                           len(b.spanoff), ptr(kq), int(kq.max()), eb, len(eb))
        if not h:
            raise GpuError(eb.value.decode())
        return C.c_void_p(h)

    def restage(self, bh, queries: np.ndarray, b: Batch, topb: int,
                k_per_query: np.ndarray | None = None) -> bool:
        """Write a new batch into staged batch `bh` (its buffers, after its
        pending work). True when its captured run replays for the new
        batch as is."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        kq = None if k_per_query is None else np.ascontiguousarray(k_per_query, dtype=np.uint32)
        eb = L.errbuf()
        rc = self.lib.fv_gpu_batch_restage(self.h, bh, b.nq, ptr(q), ptr(b.qd), ptr(b.dedup),
                                           ptr(b.nodes), len(b.nodes), ptr(b.children),
                                           len(b.children), ptr(b.invals), len(b.invals),
                                           ptr(b.spans), len(b.spans), ptr(b.spanoff),
                                           len(b.spanoff), None if kq is None else ptr(kq),
                                           topb, eb, len(eb))
        if rc < 0:
            raise GpuError(eb.value.decode())
        return rc == 1

    def run_brute(self, bh, opts: L.BruteOpts | None = None):
        eb = L.errbuf()
        rc = self.lib.fv_gpu_batch_run_brute(self.h, bh, C.byref(opts) if opts else None, eb, len(eb))
        if rc != 0:
            raise GpuError(eb.value.decode())

    def tune(self, bh, base: L.BruteOpts | None = None, reps: int = 3):
        """Autotune the launch shape for a staged batch: (best opts, us per
        run, shapes timed)."""
        best = L.BruteOpts()
        us = C.c_double(0.0)
        n = C.c_int(0)
        eb = L.errbuf()
        rc = self.lib.fv_gpu_batch_tune(self.h, bh, C.byref(base) if base else None, reps,
                                        C.byref(best), C.byref(us), C.byref(n), eb, len(eb))
        if rc != 0:
            raise GpuError(eb.value.decode())
        return best, us.value, n.value

    def time_scan(self, bh, opts: L.BruteOpts | None = None, reps: int = 20) -> float:
        """Microseconds per main-scan launch, reps launches back to back."""
        us = C.c_double(0.0)
        eb = L.errbuf()
        rc = self.lib.fv_gpu_batch_time_scan(self.h, bh, C.byref(opts) if opts else None, reps,
                                             C.byref(us), eb, len(eb))
        if rc != 0:
            raise GpuError(eb.value.decode())
        return us.value

    def run_learned(self, bh, model, opts: L.LearnedOpts):
        eb = L.errbuf()
        rc = self.lib.fv_gpu_batch_run_learned(self.h, bh, model.h if model else None,
                                               C.byref(opts), eb, len(eb))
        if rc != 0:
            raise GpuError(eb.value.decode())

    def routes(self, bh, nq: int, with_y: bool = False):
        route = np.zeros(nq, np.uint8)
        ef = np.zeros(nq, np.uint32)
        y = np.zeros((nq, 2), np.float32) if with_y else None
        eb = L.errbuf()
        if self.lib.fv_gpu_batch_routes(self.h, bh, ptr(route), ptr(ef), ptr(y), eb,
                                        len(eb)) != 0:
            raise GpuError(eb.value.decode())
        return (route, ef, y) if with_y else (route, ef)

    def run_graph(self, bh, ef: int, max_visited: int, wmax: int | None = None,
                  fcap: int | None = None):
        """Graph search of a staged batch; wmax / fcap as in graph() (a
        negative fcap bounds the frontier without a spill store)."""
        eb = L.errbuf()
        if wmax is None and fcap is None:
            rc = self.lib.fv_gpu_batch_run_graph(self.h, bh, ef, max_visited, eb, len(eb))
        else:
            rc = self.lib.fv_gpu_batch_run_graph_ex(self.h, bh, ef, max_visited, wmax or 8,
                                                    1024 if fcap is None else fcap, eb, len(eb))
        if rc != 0:
            raise GpuError(eb.value.decode())

    def sync(self):
        eb = L.errbuf()
        if self.lib.fv_gpu_sync(self.h, eb, len(eb)) != 0:
            raise GpuError(eb.value.decode())

    def fetch(self, bh, nq: int, width: int):
        out = np.zeros((nq, width), dtype=HIT_DT)
        cnt = np.zeros(nq, dtype=np.uint32)
        flag = np.zeros(nq, dtype=np.uint8)
        eb = L.errbuf()
        rc = self.lib.fv_gpu_batch_fetch(self.h, bh, ptr(out), ptr(cnt), ptr(flag), eb, len(eb))
        return out, cnt, flag, rc

    def free_batch(self, bh):
        self.lib.fv_gpu_batch_free(bh)
