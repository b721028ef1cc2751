"""HNSW graphs for the GPU beam search (bindings of fv_hnsw_* in
include/paves_gpu.h).

`build` runs the library's host builder and returns the graph in the
reference index-file layout (the arrays `indexfile.read` also returns), so
built and loaded graphs are used the same way: `l0` goes to
fv_gpu_index_upload, `entries` gives each query its level-0 entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L


def _threads(threads):
    return max(1, int(threads or min(16, os.cpu_count() or 1)))


def build(vectors: np.ndarray, M: int = 16, ef_construction: int = 200, threads: int = 0):
    lib = L.load()
    x = np.ascontiguousarray(vectors, dtype=np.float32)
    n, dim = x.shape
    eb = L.errbuf()
    h = lib.fv_hnsw_build(L.ptr(x), n, dim, M, ef_construction, _threads(threads), eb, len(eb))
    if not h:
        raise RuntimeError(eb.value.decode())
    return _export(lib, h, n)


def build_gpu(index, vectors: np.ndarray, M: int = 16, ef_construction: int = 200,
              threads: int = 0, heads: int = 8, exact: bool = False, batch_div: int = 8,
              max_batch: int = 32768, keep: bool = True):
    """Level 0 built on the GPU of `index` (an engine.GpuIndex over the same
    vectors); upper levels and entry points on the host (fv_hnsw_build_gpu).
    Returns build()'s dict plus "stats" (host / device seconds, batches).
    With keep the built graph also becomes the index's graph."""
    lib = L.load()
    x = np.ascontiguousarray(vectors, dtype=np.float32)
    n, dim = x.shape
    o = L.HnswGpuOpts()
    o.heads, o.exact, o.batch_div, o.max_batch = heads, int(exact), batch_div, max_batch
    o.keep = int(keep)
    eb = L.errbuf()
    h = lib.fv_hnsw_build_gpu(index.h, L.ptr(x), M, ef_construction, _threads(threads),
                              C.byref(o), eb, len(eb))
    if not h:
        raise RuntimeError(eb.value.decode())
    out = _export(lib, h, n)
    out["stats"] = {"host_s": o.t_host_s, "entries_s": o.t_entries_s,
                    "device_s": o.t_device_s, "batches": o.batches,
                    "search_s": o.t_search_s, "prune_s": o.t_prune_s, "link_s": o.t_link_s,
                    "visited": o.visited, "heads_expanded": o.heads_expanded}
    return out


def build_seeded(vectors: np.ndarray, M: int = 16, ef_construction: int = 200,
                 threads: int = 0, rest_on_host: bool = False):
    """build_gpu's host phase alone (fv_hnsw_build_seeded): node 0 and the
    nodes above level 0 inserted, the rest left out — or, with rest_on_host,
    inserted on the host afterwards in index order."""
    lib = L.load()
    x = np.ascontiguousarray(vectors, dtype=np.float32)
    n, dim = x.shape
    eb = L.errbuf()
    h = lib.fv_hnsw_build_seeded(L.ptr(x), n, dim, M, ef_construction, _threads(threads),
                                 int(rest_on_host), eb, len(eb))
    if not h:
        raise RuntimeError(eb.value.decode())
    return _export(lib, h, n)


def _export(lib, h, n):
    h = C.c_void_p(h)
    try:
        m, stride, top = C.c_uint32(), C.c_uint32(), C.c_uint32()
        entry, slots = C.c_uint64(), C.c_uint64()
        lib.fv_hnsw_info(h, C.byref(m), C.byref(stride), C.byref(top), C.byref(entry),
                         C.byref(slots))
        l0 = np.zeros((n, stride.value), np.uint32)
        levels = np.zeros(n, np.uint8)
        uoffs = np.zeros(n + 1, np.uint64)
        upper = np.zeros(max(1, slots.value), np.uint32)
        lib.fv_hnsw_export(h, L.ptr(l0), L.ptr(levels), L.ptr(uoffs), L.ptr(upper))
    finally:
        lib.fv_hnsw_free(h)
    return {"M": m.value, "l0": l0, "levels": levels, "upper_offs": uoffs,
            "upper": upper[: slots.value], "entry_node": entry.value, "max_level": top.value}


def entries(graph: dict, vectors: np.ndarray, queries: np.ndarray, threads: int = 0):
    """Level-0 entry point per query (greedy descent through the upper levels)."""
    lib = L.load()
    x = np.ascontiguousarray(vectors, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    out = np.zeros(q.shape[0], np.uint32)
    uoffs = np.ascontiguousarray(graph["upper_offs"], np.uint64)
    upper = np.ascontiguousarray(graph["upper"], np.uint32)
    if upper.size == 0:
        upper = np.zeros(1, np.uint32)
    rc = lib.fv_hnsw_entries(L.ptr(x), x.shape[1], L.ptr(uoffs), L.ptr(upper), int(graph["M"]),
                             int(graph["entry_node"]), int(graph["max_level"]), q.shape[0],
                             L.ptr(q), L.ptr(out), _threads(threads))
    if rc != 0:
        raise RuntimeError("fv_hnsw_entries failed")
    return out


def l2sq(a: np.ndarray, b: np.ndarray) -> float:
    lib = L.load()
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(lib.fv_l2sq(L.ptr(a), L.ptr(b), a.shape[0]))
