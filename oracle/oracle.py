"""ctypes access to the oracle — TEST INFRASTRUCTURE ONLY.

Two checkers live here:

* ``liboracle.so`` — this repo's C restatement of the reference arithmetic
  (paves_oracle.c);
* ``_ref/libpaves_ref*.so`` — the reference engine.cpp compiled from its
  own sources by oracle/Makefile (CPU-only; plus a -DFV_GPU build linked
  against the HIP engine). Used to pin the restatement, to drive the
  reference host glue against the HIP engine, and as the CPU baseline.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PRED_DT = np.dtype([("kind", "<i4"), ("col", "<i4"), ("op", "<i4"), ("nchildren", "<i4"),
                    ("child0", "<i4"), ("nvalues", "<i4"), ("val0", "<i4"), ("pad", "<i4"),
                    ("value", "<f8")])
HIT_DT = np.dtype({"names": ["node", "dist"], "formats": ["<u8", "<f4"],
                   "offsets": [0, 8], "itemsize": 16})
_P = C.c_void_p


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data if a.size else None


_or = None


def lib():
    global _or
    if _or is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built")
        L = C.CDLL(path)
        L.or_l2sq.restype = C.c_float
        L.or_l2sq.argtypes = [_P, _P, C.c_int]
        L.or_l2sq_f64.restype = C.c_double
        L.or_l2sq_f64.argtypes = [_P, _P, C.c_int]
        L.or_bf16_bits.restype = C.c_uint16
        L.or_bf16_bits.argtypes = [C.c_float]
        L.or_pred_eval.restype = C.c_int
        L.or_pred_eval.argtypes = [_P, _P, _P, _P, C.c_uint64, C.c_int32, C.c_uint32]
        L.or_brute_batch.restype = C.c_int
        L.or_brute_batch.argtypes = [_P, C.c_uint64, C.c_int, _P, C.c_int, _P, _P, _P, _P, _P,
                                     C.c_uint32, _P, _P, _P, _P]
        L.or_sq8_params.restype = None
        L.or_sq8_params.argtypes = [_P, C.c_uint64, C.c_int, _P, _P]
        L.or_sq8_codes.restype = None
        L.or_sq8_codes.argtypes = [_P, C.c_uint64, C.c_int, _P, C.c_float, _P]
        L.or_fp8_encode.restype = C.c_uint8
        L.or_fp8_encode.argtypes = [C.c_float]
        L.or_fp8_decode.restype = C.c_float
        L.or_fp8_decode.argtypes = [C.c_uint8]
        L.or_fp8_params.restype = None
        L.or_fp8_params.argtypes = [_P, C.c_uint64, C.c_int, _P, _P]
        L.or_fp8_codes.restype = None
        L.or_fp8_codes.argtypes = [_P, C.c_uint64, C.c_int, _P, C.c_float, _P]
        L.or_fp8_brute_batch.restype = C.c_int
        L.or_fp8_brute_batch.argtypes = [_P, C.c_uint64, C.c_int, _P, C.c_int, _P, _P, _P, _P, _P,
                                         C.c_uint32, C.c_uint32, _P, _P, _P]
        L.or_sq8_brute_batch.restype = C.c_int
        L.or_sq8_brute_batch.argtypes = [_P, C.c_uint64, C.c_int, _P, C.c_int, _P, _P, _P, _P, _P,
                                         C.c_uint32, C.c_uint32, _P, _P, _P]
        L.or_hnsw_search.restype = C.c_uint32
        L.or_hnsw_search.argtypes = [_P, C.c_uint64, C.c_int, _P, C.c_uint32, _P, _P, _P, _P,
                                     C.c_int32, _P, C.c_uint32, C.c_uint32, C.c_uint64, _P, _P, _P]
        L.or_hnsw_batch.restype = C.c_int
        L.or_hnsw_batch.argtypes = [_P, C.c_uint64, C.c_int, _P, C.c_uint32, _P, _P, _P, _P,
                                    C.c_int, _P, _P, _P, C.c_uint32, C.c_uint64, _P, _P, _P]
        L.or_graph_entry.restype = C.c_uint32
        L.or_graph_entry.argtypes = [_P, C.c_int, _P, _P, C.c_uint32, C.c_uint64, C.c_int, _P]
        L.or_hnsw_batch_insert.restype = C.c_int
        L.or_hnsw_batch_insert.argtypes = [_P, C.c_uint64, C.c_int, C.c_uint32, C.c_uint32, _P,
                                           _P, _P, _P, C.c_int]
        _or = L
    return _or


def l2sq(a: np.ndarray, b: np.ndarray) -> np.float32:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(lib().or_l2sq(_ptr(a), _ptr(b), a.shape[0]))


def l2sq_f64(a, b) -> float:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().or_l2sq_f64(_ptr(a), _ptr(b), a.shape[0])


def bf16_bits(x: float) -> int:
    return int(lib().or_bf16_bits(C.c_float(x)))


def pred_eval(nodes, children, invals, cols, root, row) -> bool:
    cols = np.ascontiguousarray(cols, np.float64)
    return bool(lib().or_pred_eval(_ptr(nodes), _ptr(children), _ptr(invals), _ptr(cols),
                                   cols.shape[1], root, row))


def brute_batch(vectors, cols, queries, qd, nodes, children, invals, topb):
    """Exact filtered top-topb per query; returns (hits, cnt, trunc, passing)."""
    vectors = np.ascontiguousarray(vectors, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    n, dim = vectors.shape
    nq = queries.shape[0]
    cols = None if cols is None else np.ascontiguousarray(cols, np.float64)
    proots = np.ascontiguousarray(qd["proot"], np.int32)
    out = np.zeros((nq, topb), dtype=HIT_DT)
    cnt = np.zeros(nq, np.uint32)
    trunc = np.zeros(nq, np.uint8)
    passing = np.zeros(nq, np.uint64)
    lib().or_brute_batch(_ptr(vectors), n, dim, _ptr(cols), nq, _ptr(queries), _ptr(proots),
                         _ptr(nodes), _ptr(children), _ptr(invals), topb, _ptr(out), _ptr(cnt),
                         _ptr(trunc), _ptr(passing))
    return out, cnt, trunc, passing


def sq8_params(vectors):
    """SQ8 quantizer of a row set: (per-dim midrange centre, scale)."""
    vectors = np.ascontiguousarray(vectors, np.float32)
    n, dim = vectors.shape
    mid = np.zeros(dim, np.float32)
    inv = np.zeros(1, np.float32)
    lib().or_sq8_params(_ptr(vectors), n, dim, _ptr(mid), _ptr(inv))
    return mid, np.float32(inv[0])


def sq8_codes(x, mid, inv):
    """int8 codes clamp(rint((x - mid) * inv), -127, 127) of rows x."""
    x = np.ascontiguousarray(np.atleast_2d(x), np.float32)
    mid = np.ascontiguousarray(mid, np.float32)
    out = np.zeros(x.shape, np.int8)
    lib().or_sq8_codes(_ptr(x), x.shape[0], x.shape[1], _ptr(mid), float(inv), _ptr(out))
    return out


def sq8_brute_batch(vectors, cols, queries, qd, nodes, children, invals, topb, over):
    """The quantized SQ8 path: top over x topb by code distance, re-ranked
    exactly; returns (hits, cnt, trunc)."""
    vectors = np.ascontiguousarray(vectors, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    n, dim = vectors.shape
    nq = queries.shape[0]
    cols = None if cols is None else np.ascontiguousarray(cols, np.float64)
    proots = np.ascontiguousarray(qd["proot"], np.int32)
    out = np.zeros((nq, topb), dtype=HIT_DT)
    cnt = np.zeros(nq, np.uint32)
    trunc = np.zeros(nq, np.uint8)
    rc = lib().or_sq8_brute_batch(_ptr(vectors), n, dim, _ptr(cols), nq, _ptr(queries),
                                  _ptr(proots), _ptr(nodes), _ptr(children), _ptr(invals), topb,
                                  over, _ptr(out), _ptr(cnt), _ptr(trunc))
    if rc != 0:
        raise MemoryError("oracle sq8: allocation failed")
    return out, cnt, trunc


def fp8_encode(v):
    """E4M3 code of one float (round to nearest even, saturating)."""
    return int(lib().or_fp8_encode(float(v)))


def fp8_decode_table():
    """Value of every E4M3 code 0..255 (float32; 0x7F and 0xFF decode to
    the bit pattern's formula value, never produced by the encoder)."""
    return np.array([lib().or_fp8_decode(c) for c in range(256)], np.float32)


def fp8_params(vectors):
    """FP8 quantizer of a row set: (per-dim midrange centre, scale)."""
    vectors = np.ascontiguousarray(vectors, np.float32)
    n, dim = vectors.shape
    mid = np.zeros(dim, np.float32)
    sc = np.zeros(1, np.float32)
    lib().or_fp8_params(_ptr(vectors), n, dim, _ptr(mid), _ptr(sc))
    return mid, np.float32(sc[0])


def fp8_codes(x, mid, scale):
    """E4M3 codes of rows x: encode((x - mid) * scale)."""
    x = np.ascontiguousarray(np.atleast_2d(x), np.float32)
    mid = np.ascontiguousarray(mid, np.float32)
    out = np.zeros(x.shape, np.uint8)
    lib().or_fp8_codes(_ptr(x), x.shape[0], x.shape[1], _ptr(mid), float(scale), _ptr(out))
    return out


def fp8_brute_batch(vectors, cols, queries, qd, nodes, children, invals, topb, over):
    """The FP8 index path: top over x topb by exact code distance,
    re-ranked exactly; returns (hits, cnt, trunc)."""
    vectors = np.ascontiguousarray(vectors, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    n, dim = vectors.shape
    nq = queries.shape[0]
    cols = None if cols is None else np.ascontiguousarray(cols, np.float64)
    proots = np.ascontiguousarray(qd["proot"], np.int32)
    out = np.zeros((nq, topb), dtype=HIT_DT)
    cnt = np.zeros(nq, np.uint32)
    trunc = np.zeros(nq, np.uint8)
    rc = lib().or_fp8_brute_batch(_ptr(vectors), n, dim, _ptr(cols), nq, _ptr(queries),
                                  _ptr(proots), _ptr(nodes), _ptr(children), _ptr(invals), topb,
                                  over, _ptr(out), _ptr(cnt), _ptr(trunc))
    if rc != 0:
        raise MemoryError("oracle fp8: allocation failed")
    return out, cnt, trunc


def hnsw_batch(vectors, l0, cols, queries, qd, nodes, children, invals, entries, ef, max_visited):
    vectors = np.ascontiguousarray(vectors, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    l0 = np.ascontiguousarray(l0, np.uint32)
    n, dim = vectors.shape
    nq = queries.shape[0]
    cols = None if cols is None else np.ascontiguousarray(cols, np.float64)
    proots = np.ascontiguousarray(qd["proot"], np.int32)
    entries = np.ascontiguousarray(entries, np.uint32)
    out = np.zeros((nq, ef), dtype=HIT_DT)
    cnt = np.zeros(nq, np.uint32)
    exh = np.zeros(nq, np.uint8)
    lib().or_hnsw_batch(_ptr(vectors), n, dim, _ptr(l0), l0.shape[1], _ptr(cols), _ptr(nodes),
                        _ptr(children), _ptr(invals), nq, _ptr(proots), _ptr(queries),
                        _ptr(entries), ef, max_visited, _ptr(out), _ptr(cnt), _ptr(exh))
    return out, cnt, exh


def hnsw_one(vectors, l0, cols, query, nodes, children, invals, proot, entry, ef, max_visited):
    vectors = np.ascontiguousarray(vectors, np.float32)
    query = np.ascontiguousarray(query, np.float32)
    l0 = np.ascontiguousarray(l0, np.uint32)
    out = np.zeros(ef, dtype=HIT_DT)
    exh = np.zeros(1, np.uint8)
    vis = np.zeros(1, np.uint64)
    n = lib().or_hnsw_search(_ptr(vectors), vectors.shape[0], vectors.shape[1], _ptr(l0),
                             l0.shape[1], _ptr(cols), _ptr(nodes), _ptr(children), _ptr(invals),
                             proot, _ptr(query), entry, ef, max_visited, _ptr(out), _ptr(exh),
                             _ptr(vis))
    return out[:n], int(exh[0]), int(vis[0])


def hnsw_batch_insert(vectors, l0, rest, entries, sizes, M, efc):
    """Level-0 phase of the batched HNSW builder (or_hnsw_batch_insert) from
    the host phase's lists l0 [n][2M+1]; returns the new lists."""
    x = np.ascontiguousarray(vectors, dtype=np.float32)
    out = np.ascontiguousarray(l0, dtype=np.uint32).copy()
    r = np.ascontiguousarray(rest, dtype=np.uint32)
    e = np.ascontiguousarray(entries, dtype=np.uint32)
    sz = np.ascontiguousarray(sizes, dtype=np.int32)
    rc = lib().or_hnsw_batch_insert(_ptr(x), x.shape[0], x.shape[1], M, efc, _ptr(out), _ptr(r),
                                    _ptr(e), _ptr(sz), len(sz))
    assert rc == 0
    return out


def build_schedule(n, nrest, batch_div, max_batch):
    """Batch sizes of the device builder: min(max_batch, inserted // batch_div),
    at least 1, over the nodes left after the host phase."""
    sizes, done, have = [], 0, n - nrest
    while done < nrest:
        b = min(nrest - done, max(1, min(have // batch_div, max_batch)))
        sizes.append(b)
        done += b
        have += b
    return sizes


def graph_entry(vectors, upper_offs, upper, M, entry_node, max_level, query):
    vectors = np.ascontiguousarray(vectors, np.float32)
    query = np.ascontiguousarray(query, np.float32)
    return int(lib().or_graph_entry(_ptr(vectors), vectors.shape[1], _ptr(upper_offs),
                                    _ptr(upper), M, entry_node, max_level, _ptr(query)))


def recall(got_ids, truth_ids) -> float:
    truth = set(int(x) for x in truth_ids)
    if not truth:
        return 1.0
    return len(truth & set(int(x) for x in got_ids)) / len(truth)


# ------------------------------------------------------------------ reference

def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return " avx512f" in f.read()
    except OSError:
        return False


class FvPredNode(C.Structure):
    pass


FvPredNode._fields_ = [("kind", C.c_int32), ("col", C.c_int32), ("op", C.c_int32),
                       ("value", C.c_double), ("nvalues", C.c_int32),
                       ("values", C.POINTER(C.c_double)), ("nchildren", C.c_int32),
                       ("children", C.POINTER(C.POINTER(FvPredNode)))]


class FvGpuBatchPred(C.Structure):
    _fields_ = [("nodes", _P), ("nnodes", C.c_int), ("children", _P), ("nchildren", C.c_int),
                ("invals", _P), ("ninvals", C.c_int), ("spans", _P), ("nspans", C.c_int),
                ("need_eval", C.c_int), ("dedup", C.c_int)]


class RefEngine:
    """The compiled reference engine (oracle/_ref)."""

    def __init__(self, gpu: bool = False):
        if gpu:
            name = "libpaves_ref_gpu.so"
        else:
            name = "libpaves_ref.so" if _cpu_has_avx512() else "libpaves_ref_avx2.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise RuntimeError(f"reference build {path} missing")
        self.path = path
        L = C.CDLL(path)
        sz = C.c_size_t
        L.fv_builder_create.restype = _P
        L.fv_builder_create.argtypes = [C.c_int, C.c_int, _P, C.c_uint64]
        L.fv_builder_add_row.argtypes = [_P, _P, C.c_uint64, _P, _P]
        L.fv_builder_finish.restype = C.c_int
        L.fv_builder_finish.argtypes = [_P, C.c_char_p, C.c_int32, C.c_int, C.c_int, C.c_int,
                                        C.c_char_p, sz]
        L.fv_builder_destroy.argtypes = [_P]
        L.fv_index_open.restype = _P
        L.fv_index_open.argtypes = [C.c_char_p, C.c_char_p, sz]
        L.fv_index_close.argtypes = [_P]
        L.fv_pred_compile.restype = _P
        L.fv_pred_compile.argtypes = [_P, C.POINTER(FvPredNode)]
        L.fv_pred_free.argtypes = [_P]
        L.fv_brute_begin.restype = _P
        L.fv_brute_begin.argtypes = [_P, _P, _P, C.c_int]
        L.fv_hnsw_begin.restype = _P
        L.fv_hnsw_begin.argtypes = [_P, _P, _P, C.c_int, C.c_uint64]
        L.fv_cursor_next.restype = C.c_int
        L.fv_cursor_next.argtypes = [_P, _P, C.c_int]
        L.fv_cursor_end.argtypes = [_P]
        L.fv_index_graph_entry.restype = C.c_uint32
        L.fv_index_graph_entry.argtypes = [_P, _P]
        L.fv_pred_serialize.restype = C.c_int
        L.fv_pred_serialize.argtypes = [_P, _P, _P, C.c_int, _P, C.c_int, _P, C.c_int, _P,
                                        C.c_int, C.POINTER(FvGpuBatchPred), C.POINTER(C.c_uint64)]
        L.fv_gpu_exec_brute_batch.restype = C.c_int
        L.fv_gpu_exec_brute_batch.argtypes = [_P, C.c_int, _P, C.POINTER(FvGpuBatchPred),
                                              C.c_uint32, _P, _P, _P]
        L.fv_gpu_exec_graph_batch.restype = C.c_int
        L.fv_gpu_exec_graph_batch.argtypes = [_P, C.c_int, _P, _P, C.POINTER(FvGpuBatchPred),
                                              C.c_uint32, C.c_uint64, _P, _P, _P]
        L.fv_engine_gpu_precision.argtypes = [C.c_int]
        L.fv_index_gpu_ensure.restype = C.c_int
        L.fv_index_gpu_ensure.argtypes = [_P, C.c_int]
        self.L = L

    # -- build an index file with the reference builder
    def build(self, path, vectors, cols, M=16, efc=200, threads=8):
        n, dim = vectors.shape
        ncols = 0 if cols is None else cols.shape[0]
        att = np.arange(2, 2 + ncols, dtype=np.int32)
        b = self.L.fv_builder_create(dim, ncols, _ptr(att), n)
        vec = np.ascontiguousarray(vectors, np.float32)
        sc = np.zeros(max(ncols, 1), np.float64)
        nul = np.zeros(max(ncols, 1), np.uint8)
        for i in range(n):
            if ncols:
                sc[:ncols] = cols[:, i]
                nul[:ncols] = np.isnan(cols[:, i])
            self.L.fv_builder_add_row(b, vec[i].ctypes.data, i, _ptr(sc), _ptr(nul))
        eb = C.create_string_buffer(256)
        rc = self.L.fv_builder_finish(b, path.encode(), 1, M, efc, threads, eb, 256)
        self.L.fv_builder_destroy(b)
        if rc != 0:
            raise RuntimeError(eb.value.decode())

    def open(self, path):
        eb = C.create_string_buffer(256)
        h = self.L.fv_index_open(path.encode(), eb, 256)
        if not h:
            raise RuntimeError(eb.value.decode())
        return _P(h)

    def close(self, idx):
        self.L.fv_index_close(idx)

    # -- predicates as FvPredNode trees (engine_api.h:76-85)
    def make_pred(self, p, keep):
        node = FvPredNode()
        node.kind, node.col, node.op, node.value = p.kind, p.col, p.op, p.value
        if p.values:
            arr = (C.c_double * len(p.values))(*p.values)
            keep.append(arr)
            node.nvalues = len(p.values)
            node.values = C.cast(arr, C.POINTER(C.c_double))
        if p.children:
            kids = [self.make_pred(c, keep) for c in p.children]
            arr = (C.POINTER(FvPredNode) * len(kids))(*[C.pointer(k) for k in kids])
            keep.append(arr)
            node.nchildren = len(kids)
            node.children = C.cast(arr, C.POINTER(C.POINTER(FvPredNode)))
        keep.append(node)
        return node

    def compile(self, idx, p):
        if p is None:
            return None, []
        keep = []
        root = self.make_pred(p, keep)
        return _P(self.L.fv_pred_compile(idx, C.byref(root))), keep

    def drain(self, cur, want, step=256):
        out = []
        buf = np.zeros(step, dtype=HIT_DT)
        while len(out) < want:
            k = self.L.fv_cursor_next(cur, buf.ctypes.data, min(step, want - len(out)))
            if k <= 0:
                break
            out.extend((int(buf[i]["node"]), float(buf[i]["dist"])) for i in range(k))
        return out

    def brute(self, idx, query, pred_h, want, threads=8):
        q = np.ascontiguousarray(query, np.float32)
        cur = self.L.fv_brute_begin(idx, q.ctypes.data, pred_h, threads)
        try:
            return self.drain(cur, want)
        finally:
            self.L.fv_cursor_end(cur)

    def hnsw(self, idx, query, pred_h, ef, want, max_visited=0):
        q = np.ascontiguousarray(query, np.float32)
        cur = self.L.fv_hnsw_begin(idx, q.ctypes.data, pred_h, ef, max_visited)
        try:
            return self.drain(cur, want)
        finally:
            self.L.fv_cursor_end(cur)

    def graph_entry(self, idx, query):
        q = np.ascontiguousarray(query, np.float32)
        return int(self.L.fv_index_graph_entry(idx, q.ctypes.data))


def read_index_file(path):
    """Parse a PAVES index file (FileHeader, engine.cpp:193-215) with numpy
    memmaps; returns a dict of arrays."""
    with open(path, "rb") as f:
        hdr = f.read(8 + 4 + 4 + 8 + 4 + 4 + 128 + 4 + 4 + 4 + 4 + 8 * 11)
    magic = hdr[:8]
    if magic != b"PGFAVR01":
        raise ValueError("bad magic")
    version, dim, count, vec_att, ncols = struct.unpack_from("<IIqiI", hdr, 8)
    off = 8 + 4 + 4 + 8 + 4 + 4 + 128
    M, efc, max_level = struct.unpack_from("<III", hdr, off)
    off += 12 + 4  # padding to 8 before entry_node
    fields = struct.unpack_from("<11Q", hdr, off)
    entry_node, o_vec, o_tid, o_cols, o_pvals, o_prows, o_lvl, o_l0, o_uoffs, o_up, fsize = fields
    mm = np.memmap(path, dtype=np.uint8, mode="r")
    def arr(o, dt, n):
        return np.frombuffer(mm, dtype=dt, count=n, offset=o)
    stride = 2 * M + 1
    d = {"dim": dim, "count": count, "ncols": ncols, "M": M, "efc": efc,
         "max_level": max_level, "entry_node": entry_node,
         "vectors": arr(o_vec, np.float32, count * dim).reshape(count, dim),
         "tids": arr(o_tid, np.uint64, count),
         "cols": arr(o_cols, np.float64, ncols * count).reshape(ncols, count),
         "post_vals": arr(o_pvals, np.float64, ncols * count).reshape(ncols, count),
         "post_rows": arr(o_prows, np.uint32, ncols * count).reshape(ncols, count),
         "levels": arr(o_lvl, np.uint8, count),
         "l0": arr(o_l0, np.uint32, count * stride).reshape(count, stride),
         "upper_offs": arr(o_uoffs, np.uint64, count + 1)}
    nup = int(d["upper_offs"][-1])
    d["upper"] = arr(o_up, np.uint32, nup)
    return d
