"""PAVES on-disk index format (reader and writer).

Format of engine.cpp:190-215 (FileHeader, 264 bytes) followed by 64-byte
aligned sections: vectors f32[count*dim], tids u64[count], cols
f64[ncols*count] (column-major), post_vals f64[ncols*count], post_rows
u32[ncols*count], levels u8[count], l0 u32[count*(2M+1)] ([cnt, neigh...]),
upper_offs u64[count+1], upper u32[...] (per node, level slabs of M+1).
The layout must stay byte-identical so real PAVES index files can be read
and files written here open in the reference engine.
"""
from __future__ import annotations

import os
import struct

import numpy as np

MAGIC = b"PGFAVR01"
VERSION = 1
MAX_COLS = 32
HEADER_SIZE = 264


def _a64(x: int) -> int:
    return (x + 63) & ~63


def write(path: str, vectors: np.ndarray, cols: np.ndarray | None, *, tids=None,
          M: int = 16, efc: int = 200, levels=None, l0=None, upper_offs=None, upper=None,
          entry_node: int = 0, max_level: int = 0, post_rows=None, post_vals=None,
          col_attnums=None, vec_attnum: int = 1) -> None:
    """Write an index file. Without a graph, l0 lists are empty (count 0) and
    there are no upper levels — enough for the brute-force path."""
    vectors = np.ascontiguousarray(vectors, np.float32)
    n, dim = vectors.shape
    ncols = 0 if cols is None else cols.shape[0]
    if ncols > MAX_COLS:
        raise ValueError("too many columns")
    stride = 2 * M + 1
    if tids is None:
        tids = np.arange(n, dtype=np.uint64)
    if levels is None:
        levels = np.zeros(n, np.uint8)
    if l0 is None:
        l0 = np.zeros((n, stride), np.uint32)
    if upper_offs is None:
        upper_offs = np.zeros(n + 1, np.uint64)
    if upper is None:
        upper = np.zeros(0, np.uint32)
    if ncols and (post_rows is None or post_vals is None):
        from .synth import postings
        post_rows, post_vals = postings(cols)
    if col_attnums is None:
        col_attnums = list(range(2, 2 + ncols))
    off = _a64(HEADER_SIZE)
    o_vec = off; off = _a64(off + n * dim * 4)
    o_tid = off; off = _a64(off + n * 8)
    o_cols = off; off = _a64(off + ncols * n * 8)
    o_pv = off; off = _a64(off + ncols * n * 8)
    o_pr = off; off = _a64(off + ncols * n * 4)
    o_lvl = off; off = _a64(off + n)
    o_l0 = off; off = _a64(off + n * stride * 4)
    o_uo = off; off = _a64(off + (n + 1) * 8)
    o_up = off; off = _a64(off + int(upper_offs[-1]) * 4)
    fsize = off
    att = list(col_attnums) + [0] * (MAX_COLS - ncols)
    hdr = (MAGIC + struct.pack("<IIqiI", VERSION, dim, n, vec_attnum, ncols) +
           struct.pack("<32i", *att) + struct.pack("<IIII", M, efc, max_level, 0) +
           struct.pack("<11Q", entry_node, o_vec, o_tid, o_cols, o_pv, o_pr, o_lvl, o_l0,
                       o_uo, o_up, fsize))
    assert len(hdr) == HEADER_SIZE
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(hdr)
        def put(o, arr):
            f.seek(o)
            f.write(np.ascontiguousarray(arr).tobytes())
        put(o_vec, vectors)
        put(o_tid, np.asarray(tids, np.uint64))
        if ncols:
            put(o_cols, np.asarray(cols, np.float64))
            put(o_pv, np.asarray(post_vals, np.float64))
            put(o_pr, np.asarray(post_rows, np.uint32))
        put(o_lvl, np.asarray(levels, np.uint8))
        put(o_l0, np.asarray(l0, np.uint32))
        put(o_uo, np.asarray(upper_offs, np.uint64))
        if len(upper):
            put(o_up, np.asarray(upper, np.uint32))
        f.seek(fsize - 1)
        f.write(b"\0")
    os.replace(tmp, path)


def read(path: str) -> dict:
    """Memory-map an index file; returns its arrays and header fields."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        hdr = f.read(HEADER_SIZE)
    if len(hdr) < HEADER_SIZE or hdr[:8] != MAGIC:
        raise ValueError(f"{path}: bad magic")
    version, dim, count, vec_att, ncols = struct.unpack_from("<IIqiI", hdr, 8)
    M, efc, max_level = struct.unpack_from("<III", hdr, 160)
    (entry_node, o_vec, o_tid, o_cols, o_pv, o_pr, o_lvl, o_l0, o_uo, o_up,
     fsize) = struct.unpack_from("<11Q", hdr, 176)
    if version != VERSION or fsize != size:
        raise ValueError(f"{path}: bad version/size")
    mm = np.memmap(path, dtype=np.uint8, mode="r")

    def arr(o, dt, n):
        return np.frombuffer(mm, dtype=dt, count=n, offset=o)
    stride = 2 * M + 1
    d = {"dim": dim, "count": count, "ncols": ncols, "M": M, "efc": efc,
         "max_level": max_level, "entry_node": entry_node, "l0_stride": stride,
         "vectors": arr(o_vec, np.float32, count * dim).reshape(count, dim),
         "tids": arr(o_tid, np.uint64, count),
         "cols": arr(o_cols, np.float64, ncols * count).reshape(ncols, count),
         "post_vals": arr(o_pv, np.float64, ncols * count).reshape(ncols, count),
         "post_rows": arr(o_pr, np.uint32, ncols * count).reshape(ncols, count),
         "levels": arr(o_lvl, np.uint8, count),
         "l0": arr(o_l0, np.uint32, count * stride).reshape(count, stride),
         "upper_offs": arr(o_uo, np.uint64, count + 1)}
    d["upper"] = arr(o_up, np.uint32, int(d["upper_offs"][-1]))
    return d
