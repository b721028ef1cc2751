"""ctypes bindings of libpaves_amd.so (include/paves_gpu.h).

The shared library is built in-tree by ``paves_amd/Makefile``; importing
this module never falls back to anything else — a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpaves_amd.so")

# byte layouts of engine_api.h / gpu_engine.h
PRED_DT = np.dtype([("kind", "<i4"), ("col", "<i4"), ("op", "<i4"), ("nchildren", "<i4"),
                    ("child0", "<i4"), ("nvalues", "<i4"), ("val0", "<i4"), ("pad", "<i4"),
                    ("value", "<f8")])
SPAN_DT = np.dtype([("col", "<i4"), ("lo", "<u4"), ("hi", "<u4")])
QD_DT = np.dtype([("proot", "<i4"), ("need_eval", "<i4"), ("span0", "<i4"), ("nspans", "<i4"),
                  ("soff0", "<i4"), ("entry", "<u4"), ("total", "<u8")])
HIT_DT = np.dtype({"names": ["node", "dist"], "formats": ["<u8", "<f4"],
                   "offsets": [0, 8], "itemsize": 16})
assert PRED_DT.itemsize == 40 and SPAN_DT.itemsize == 12
assert QD_DT.itemsize == 32 and HIT_DT.itemsize == 16

FV_GPU_PREC_FP32 = 0
FV_GPU_PREC_SQ8 = 1
FV_GPU_PREC_BF16 = 2
FV_GPU_PREC_FP32_VALU = 3
FV_GPU_PREC_FP8 = 4
FV_GPU_MAX_TOPB = 4096
FV_GPU_MAX_SPANS = 4096
FV_GPU_MAX_EF = 512
FV_GPU_MAX_BATCH = 1024

# FvCmpOp / FvNodeKind (engine_api.h:69-70)
CMP_LT, CMP_LE, CMP_EQ, CMP_GE, CMP_GT, CMP_NE = range(6)
NODE_CMP, NODE_IN, NODE_AND, NODE_OR, NODE_NOT = range(5)

REF_SYMBOLS = ["fv_gpu_available", "fv_gpu_index_upload", "fv_gpu_index_free",
               "fv_gpu_index_device", "fv_gpu_index_precision", "fv_gpu_brute_batch",
               "fv_gpu_graph_batch", "fv_gpu_tile_variant_qt", "fv_gpu_mem_probe"]
EXT_SYMBOLS = ["fv_gpu_set_timing", "fv_gpu_last_stats", "fv_gpu_feedback",
               "fv_gpu_feedback_update", "fv_gpu_sel_bucket", "fv_gpu_brute_batch_ex",
               "fv_gpu_graph_batch_ex", "fv_gpu_batch_stage", "fv_gpu_batch_run_brute",
               "fv_gpu_batch_run_graph", "fv_gpu_sync", "fv_gpu_batch_fetch",
               "fv_gpu_batch_free", "fv_gpu_copy_bench", "fv_gpu_read_bench", "fv_hnsw_build", "fv_hnsw_info",
               "fv_hnsw_export", "fv_hnsw_free", "fv_hnsw_entries", "fv_l2sq", "fv_hnsw_build_gpu", "fv_hnsw_build_seeded",
               "fv_gpu_model_nparams", "fv_gpu_model_create", "fv_gpu_model_free",
               "fv_gpu_model_set_params", "fv_gpu_model_get_params", "fv_gpu_model_grad",
               "fv_gpu_model_train", "fv_gpu_model_predict", "fv_gpu_batch_features",
               "fv_gpu_batch_run_learned", "fv_gpu_batch_routes", "fv_gpu_batch_stage_ex",
               "fv_gpu_batch_tune", "fv_gpu_batch_time_scan", "fv_gpu_brute_batch_noalloc",
               "fv_gpu_batch_restage", "fv_gpu_batch_run_graph_ex"]


class BruteOpts(C.Structure):
    """FvGpuBruteOpts; precision -1 = the index's precision."""
    _fields_ = [("precision", C.c_int32), ("route", C.c_int32),
                ("overfetch", C.c_uint32), ("sample_stride", C.c_uint32),
                ("inline_pred", C.c_int32), ("wg_per_cu", C.c_int32), ("qtw", C.c_int32),
                ("graph", C.c_int32), ("order", C.c_int32), ("waves", C.c_int32)]

    def __init__(self, precision=-1, route=0, overfetch=0, sample_stride=0, inline_pred=0,
                 wg_per_cu=0, qtw=0, graph=0, order=0, waves=0):
        super().__init__(precision, route, overfetch, sample_stride, inline_pred, wg_per_cu,
                         qtw, graph, order, waves)


class LearnedOpts(C.Structure):
    """FvGpuLearnedOpts; modes: 0 learned route + ef, 1 graph with learned
    ef, 2 graph with fixed ef, 3 brute."""
    _fields_ = [("k", C.c_uint32), ("mode", C.c_int32), ("ef_fixed", C.c_uint32),
                ("ef_min", C.c_uint32), ("ef_max", C.c_uint32), ("ef_scale", C.c_float),
                ("route_bias", C.c_float), ("max_visited", C.c_uint64),
                ("precision", C.c_int32), ("graph_cus", C.c_int32)]

    def __init__(self, k=10, mode=0, ef_fixed=400, ef_min=10, ef_max=512, ef_scale=1.0,
                 route_bias=0.0, max_visited=0, precision=-1, graph_cus=0):
        super().__init__(k, mode, ef_fixed, ef_min, ef_max, ef_scale, route_bias, max_visited,
                         precision, graph_cus)


class Feedback(C.Structure):
    """FvGpuFeedback: service-time EWMAs (us) per strategy (0 brute, 1
    graph) and selectivity bucket, batch time and size, batches recorded."""
    _fields_ = [("q_us", (C.c_double * 5) * 2), ("batch_us", C.c_double * 2),
                ("batch_n", C.c_double * 2), ("batches", C.c_uint64 * 2)]


class HnswGpuOpts(C.Structure):
    """FvHnswGpuOpts: device graph build options (in) and timings (out)."""
    _fields_ = [("heads", C.c_int), ("exact", C.c_int), ("batch_div", C.c_int),
                ("max_batch", C.c_int), ("keep", C.c_int), ("pad0", C.c_int),
                ("t_host_s", C.c_double), ("t_device_s", C.c_double), ("t_entries_s", C.c_double),
                ("t_search_s", C.c_double), ("t_prune_s", C.c_double), ("t_link_s", C.c_double),
                ("visited", C.c_uint64), ("heads_expanded", C.c_uint64), ("batches", C.c_int),
                ("pad", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("t_bitmap_us", C.c_double), ("t_prepass_us", C.c_double),
                ("t_scan_us", C.c_double), ("t_merge_us", C.c_double),
                ("t_rerank_us", C.c_double), ("t_gather_us", C.c_double),
                ("t_graph_us", C.c_double), ("t_total_us", C.c_double),
                ("scan_rows", C.c_uint64), ("scan_bytes", C.c_uint64),
                ("candidates", C.c_uint64), ("route", C.c_int32),
                ("precision", C.c_int32), ("certified_fail", C.c_uint32),
                ("nprog", C.c_uint32)]


_P = C.c_void_p
_lib = None


def load() -> C.CDLL:
    """Load the engine library (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"paves_amd: {LIB_PATH} not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    sz = C.c_size_t
    lib.fv_gpu_available.restype = C.c_int
    lib.fv_gpu_index_upload.restype = _P
    lib.fv_gpu_index_upload.argtypes = [_P, _P, _P, _P, C.c_uint32, C.c_uint64, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_char_p, sz]
    lib.fv_gpu_index_free.argtypes = [_P]
    lib.fv_gpu_index_device.argtypes = [_P]
    lib.fv_gpu_index_precision.argtypes = [_P]
    lib.fv_gpu_brute_batch.restype = C.c_int
    lib.fv_gpu_brute_batch.argtypes = [_P, C.c_int, _P, _P, _P, _P, C.c_int, _P, C.c_int,
                                       _P, C.c_int, _P, C.c_int, _P, C.c_int, C.c_uint32,
                                       _P, _P, _P, C.c_char_p, sz]
    lib.fv_gpu_brute_batch_ex.restype = C.c_int
    lib.fv_gpu_brute_batch_ex.argtypes = [_P, C.c_int, _P, _P, _P, _P, C.c_int, _P, C.c_int,
                                          _P, C.c_int, _P, C.c_int, _P, C.c_int, C.c_uint32,
                                          C.POINTER(BruteOpts), _P, _P, _P, C.c_char_p, sz]
    lib.fv_gpu_brute_batch_noalloc.restype = C.c_int
    lib.fv_gpu_brute_batch_noalloc.argtypes = lib.fv_gpu_brute_batch.argtypes
    lib.fv_gpu_graph_batch.restype = C.c_int
    lib.fv_gpu_graph_batch.argtypes = [_P, C.c_int, _P, _P, _P, C.c_int, _P, C.c_int, _P,
                                       C.c_int, C.c_uint32, C.c_uint64, _P, _P, _P,
                                       C.c_char_p, sz]
    lib.fv_gpu_graph_batch_ex.restype = C.c_int
    lib.fv_gpu_graph_batch_ex.argtypes = [_P, C.c_int, _P, _P, _P, C.c_int, _P, C.c_int, _P,
                                          C.c_int, C.c_uint32, C.c_uint64, C.c_int, C.c_int,
                                          _P, _P, _P, _P, C.c_char_p, sz]
    lib.fv_gpu_tile_variant_qt.argtypes = [C.c_int]
    lib.fv_gpu_mem_probe.argtypes = [_P, _P]
    lib.fv_gpu_set_timing.argtypes = [_P, C.c_int]
    lib.fv_gpu_last_stats.argtypes = [_P, C.POINTER(Stats)]
    lib.fv_gpu_feedback.restype = C.c_int
    lib.fv_gpu_feedback.argtypes = [_P, C.POINTER(Feedback)]
    lib.fv_gpu_feedback_update.restype = None
    lib.fv_gpu_feedback_update.argtypes = [C.POINTER(Feedback), C.c_int, C.c_int, _P, _P,
                                           C.c_double]
    lib.fv_gpu_sel_bucket.restype = C.c_int
    lib.fv_gpu_sel_bucket.argtypes = [C.c_int32]
    lib.fv_gpu_batch_stage.restype = _P
    lib.fv_gpu_batch_stage.argtypes = [_P, C.c_int, _P, _P, _P, _P, C.c_int, _P, C.c_int,
                                       _P, C.c_int, _P, C.c_int, _P, C.c_int, C.c_uint32,
                                       C.c_char_p, sz]
    lib.fv_gpu_batch_stage_ex.restype = _P
    lib.fv_gpu_batch_stage_ex.argtypes = [_P, C.c_int, _P, _P, _P, _P, C.c_int, _P, C.c_int,
                                          _P, C.c_int, _P, C.c_int, _P, C.c_int, _P, C.c_uint32,
                                          C.c_char_p, sz]
    lib.fv_gpu_batch_restage.argtypes = [_P, _P, C.c_int, _P, _P, _P, _P, C.c_int, _P,
                                         C.c_int, _P, C.c_int, _P, C.c_int, _P, C.c_int, _P,
                                         C.c_uint32, C.c_char_p, sz]
    lib.fv_gpu_batch_run_brute.argtypes = [_P, _P, C.POINTER(BruteOpts), C.c_char_p, sz]
    lib.fv_gpu_batch_time_scan.argtypes = [_P, _P, C.POINTER(BruteOpts), C.c_int,
                                           C.POINTER(C.c_double), C.c_char_p, sz]
    lib.fv_gpu_batch_tune.argtypes = [_P, _P, C.POINTER(BruteOpts), C.c_int,
                                      C.POINTER(BruteOpts), C.POINTER(C.c_double),
                                      C.POINTER(C.c_int), C.c_char_p, sz]
    lib.fv_gpu_batch_run_graph.argtypes = [_P, _P, C.c_uint32, C.c_uint64, C.c_char_p, sz]
    lib.fv_gpu_batch_run_graph_ex.argtypes = [_P, _P, C.c_uint32, C.c_uint64, C.c_int, C.c_int,
                                              C.c_char_p, sz]
    lib.fv_gpu_sync.argtypes = [_P, C.c_char_p, sz]
    lib.fv_gpu_batch_fetch.argtypes = [_P, _P, _P, _P, _P, C.c_char_p, sz]
    lib.fv_gpu_batch_free.argtypes = [_P]
    lib.fv_gpu_copy_bench.argtypes = [C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_double),
                                      C.c_char_p, sz]
    lib.fv_gpu_read_bench.argtypes = [C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_double),
                                      C.c_char_p, sz]
    lib.fv_hnsw_build.restype = _P
    lib.fv_hnsw_build.argtypes = [_P, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_char_p, sz]
    lib.fv_hnsw_info.argtypes = [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64)]
    lib.fv_hnsw_export.argtypes = [_P, _P, _P, _P, _P]
    lib.fv_hnsw_free.argtypes = [_P]
    lib.fv_hnsw_build_seeded.restype = _P
    lib.fv_hnsw_build_seeded.argtypes = [_P, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_char_p, C.c_size_t]
    lib.fv_hnsw_build_gpu.restype = _P
    lib.fv_hnsw_build_gpu.argtypes = [_P, _P, C.c_int, C.c_int, C.c_int, C.POINTER(HnswGpuOpts),
                                      C.c_char_p, C.c_size_t]
    lib.fv_hnsw_entries.argtypes = [_P, C.c_int, _P, _P, C.c_uint32, C.c_uint64, C.c_uint32,
                                    C.c_int, _P, _P, C.c_int]
    lib.fv_l2sq.restype = C.c_float
    lib.fv_l2sq.argtypes = [_P, _P, C.c_int]
    lib.fv_gpu_model_nparams.argtypes = [C.c_int]
    lib.fv_gpu_model_create.restype = _P
    lib.fv_gpu_model_create.argtypes = [C.c_int, C.c_int, _P, C.c_int, C.c_char_p, sz]
    lib.fv_gpu_model_free.argtypes = [_P]
    lib.fv_gpu_model_set_params.argtypes = [_P, _P, _P, _P, C.c_char_p, sz]
    lib.fv_gpu_model_get_params.argtypes = [_P, _P, _P, _P, C.c_char_p, sz]
    lib.fv_gpu_model_grad.argtypes = [_P, C.c_int, _P, _P, C.c_float, _P,
                                      C.POINTER(C.c_double), C.c_char_p, sz]
    lib.fv_gpu_model_train.argtypes = [_P, C.c_int, _P, _P, C.c_int, C.c_int, C.c_float,
                                       C.c_float, C.c_uint64, _P, C.c_char_p, sz]
    lib.fv_gpu_model_predict.argtypes = [_P, C.c_int, _P, _P, C.c_char_p, sz]
    lib.fv_gpu_batch_features.argtypes = [_P, _P, _P, _P, C.c_char_p, sz]
    lib.fv_gpu_batch_run_learned.argtypes = [_P, _P, _P, C.POINTER(LearnedOpts), C.c_char_p, sz]
    lib.fv_gpu_batch_routes.argtypes = [_P, _P, _P, _P, _P, C.c_char_p, sz]
    _lib = lib
    return lib


def ptr(a) -> int | None:
    """Data pointer of a numpy array (None for None / empty)."""
    if a is None:
        return None
    return a.ctypes.data if a.size else None


def errbuf():
    return C.create_string_buffer(512)
