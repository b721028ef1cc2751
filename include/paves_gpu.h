/*
 * paves_gpu.h — C ABI of the MI355X query engine (libpaves_amd.so).
 *
 * This library is a drop-in for PAVES's CUDA object: it exports the same
 * nine entry points with the same argument meaning, return codes and byte
 * layouts, so engine.cpp (built with -DFV_GPU) links against it unchanged.
 *
 *   reference interface                                this header
 *   extension/paves/src/engine/gpu_engine.h:24      -> FvGpuIndex (opaque)
 *   extension/paves/src/engine/gpu_engine.h:52-60   -> FvGpuQueryDesc
 *   extension/paves/src/engine/engine_api.h:102-117 -> FvGpuPredNode, FvGpuSpan
 *   extension/paves/src/engine/engine_api.h:145-148 -> FvHit
 *   extension/paves/src/engine/gpu_engine.h:63-109  -> fv_gpu_* entry points
 *   extension/paves/src/engine/engine_api.h:194-195 -> fv_gpu_tile_variant_qt,
 *                                                     fv_gpu_mem_probe
 *
 * When compiled together with the reference headers (engine_api.h already
 * included) the shared types are taken from there; the static size checks
 * below hold either way.
 *
 * Contract (as in the reference): every entry point is synchronous; errors
 * return -1 (NULL for upload) with a message in errbuf; no C++ exception
 * crosses the ABI; device initialisation is lazy (first upload), so it
 * happens after fork().
 *
 * Entry points under "extensions" have no counterpart in the reference.
 */
#ifndef PAVES_AMD_GPU_H
#define PAVES_AMD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef PAVES_ENGINE_API_H
/* ---- shared types, byte-identical to engine_api.h ---- */
typedef enum { FV_CMP_LT, FV_CMP_LE, FV_CMP_EQ, FV_CMP_GE, FV_CMP_GT, FV_CMP_NE } FvCmpOp;
typedef enum { FV_NODE_CMP, FV_NODE_IN, FV_NODE_AND, FV_NODE_OR, FV_NODE_NOT } FvNodeKind;

/* engine_api.h:102-112 — one predicate node; children/IN values are the
 * ranges [child0, child0+nchildren) / [val0, val0+nvalues) of side arrays */
typedef struct FvGpuPredNode {
    int32_t kind;
    int32_t col;
    int32_t op;
    int32_t nchildren;
    int32_t child0;
    int32_t nvalues;
    int32_t val0;
    int32_t pad;
    double value;
} FvGpuPredNode;

/* engine_api.h:114-117 — [lo, hi) positions in post_rows(col) */
typedef struct FvGpuSpan {
    int32_t col;
    uint32_t lo, hi;
} FvGpuSpan;

/* engine_api.h:145-148 */
typedef struct FvHit {
    uint64_t node;
    float dist; /* squared L2 */
} FvHit;
#endif /* PAVES_ENGINE_API_H */

#ifndef PAVES_GPU_ENGINE_H
typedef struct FvGpuIndex FvGpuIndex;

/* gpu_engine.h:33-34. FP32 and BF16 return exact results: ids and FP32
 * distances bit-identical to the CPU cursors (a certified FP32 re-rank
 * makes every scan exact). An FP32 index is scanned in BF16
 * (FV_GPU_PREC_BF16) unless a call asks for FP32. SQ8 (the reference's
 * reduced-precision index, quantized scan + re-rank of the candidates):
 * over integer data spanning at most 255 values in at most 256 dims, an
 * exact INT8 copy (exact results, as FP32); over any other data up to 1024
 * dims, int8 codes of the midrange-centred rows under one scale — the scan
 * ranks by code distance and the top FvGpuBruteOpts.overfetch x B rows are
 * re-ranked with the exact FP32 distance (approximate, like the
 * reference's SQ8; returned distances are exact). */
#define FV_GPU_PREC_FP32 0
#define FV_GPU_PREC_SQ8 1

#define FV_GPU_MAX_TOPB 4096   /* gpu_engine.h:37 */
#define FV_GPU_MAX_SPANS 4096  /* gpu_engine.h:39 */
#define FV_GPU_MAX_EF 512      /* gpu_engine.h:42 */
#define FV_GPU_MAX_BATCH 1024  /* gpu_engine.h:44 */

/* gpu_engine.h:52-60 */
typedef struct FvGpuQueryDesc {
    int32_t proot;
    int32_t need_eval;
    int32_t span0;
    int32_t nspans;
    int32_t soff0;
    uint32_t entry;
    uint64_t total;
} FvGpuQueryDesc;
#endif /* PAVES_GPU_ENGINE_H */

/* BF16 MFMA distance scan + exact FP32 re-rank of an over-fetched,
 * error-bounded candidate set; ids and distances equal the FP32 path. */
#define FV_GPU_PREC_BF16 2
/* Scan precision of FvGpuBruteOpts only (not an index precision): the FP32
 * scan on the vector ALUs, sum (x - q)^2 (the reference tiled kernel's
 * arithmetic, gpu_engine.cu:534-566). FV_GPU_PREC_FP32 asked for in the
 * options scans x.q on FP32 MFMA (dim % 8 == 0, dim <= 512; other dims take
 * the VALU scan). Every scan returns the same exact results. */
#define FV_GPU_PREC_FP32_VALU 3
/* Index precision (extension; no reference counterpart): FP8 E4M3 codes of
 * the midrange-centred rows under one scale (the largest half-range maps to
 * 448), scanned with v_mfma_f32_32x32x16_fp8_fp8 (FP32 accumulation); the
 * code distance ranks the rows and the top FvGpuBruteOpts.overfetch x B are
 * re-ranked with the exact FP32 distance, as the quantized SQ8 index. Any
 * float data up to 1024 dims; returned distances are exact. Asking for
 * FV_GPU_PREC_SQ8 on an FP8 index scans the FP8 codes. */
#define FV_GPU_PREC_FP8 4

#ifdef __cplusplus
#define PV_STATIC_ASSERT static_assert
#else
#define PV_STATIC_ASSERT _Static_assert
#endif
PV_STATIC_ASSERT(sizeof(FvGpuPredNode) == 40, "FvGpuPredNode layout");
PV_STATIC_ASSERT(sizeof(FvGpuSpan) == 12, "FvGpuSpan layout");
PV_STATIC_ASSERT(sizeof(FvGpuQueryDesc) == 32, "FvGpuQueryDesc layout");
PV_STATIC_ASSERT(sizeof(FvHit) == 16, "FvHit layout");

/* ================= reference entry points ================= */

/* gpu_engine.h:63 — 1 if a usable HIP device exists (initialises HIP). */
int fv_gpu_available(void);

/* gpu_engine.h:69-73 — copy vectors (count x dim fp32, row-major), scalar
 * columns (ncols x count f64, column-major, NULL = NaN), posting rows
 * (ncols x count u32) and level-0 adjacency (count x l0_stride u32,
 * [cnt, neigh...]; may be NULL) to one GPU. device -1 picks the device
 * with the most free memory. NULL + errbuf on failure. */
FvGpuIndex *fv_gpu_index_upload(const float *vectors, const double *cols,
                                const uint32_t *post_rows,
                                const uint32_t *l0, uint32_t l0_stride,
                                uint64_t count, int dim, int ncols, int device,
                                int precision, char *errbuf, size_t errlen);
/* gpu_engine.h:74-76 */
void fv_gpu_index_free(FvGpuIndex *g);
int fv_gpu_index_device(const FvGpuIndex *g);
int fv_gpu_index_precision(const FvGpuIndex *g);

/* gpu_engine.h:86-94 — batched exact filtered top-topb. Hits of query q go
 * to out[q*topb ..] ascending by (dist, node); out_cnt[q] entries;
 * truncated[q] = 1 when out_cnt[q] == topb (more passing rows may exist,
 * gpu_engine.cu:1059-1064). dedup, spans and spanoff are accepted for ABI
 * compatibility and used by the span-gather route. 0 / -1 + errbuf. */
int fv_gpu_brute_batch(FvGpuIndex *g, int nq, const float *queries,
                       const FvGpuQueryDesc *qd, const uint8_t *dedup,
                       const FvGpuPredNode *nodes, int nnodes,
                       const int32_t *children, int nchildren,
                       const double *invals, int ninvals,
                       const FvGpuSpan *spans, int nspans_total,
                       const uint64_t *spanoff, int nspanoff, uint32_t topb,
                       FvHit *out, uint32_t *out_cnt, uint8_t *truncated,
                       char *errbuf, size_t errlen);

/* gpu_engine.h:104-109 — batched filtered level-0 beam search from
 * qd[q].entry. Traversal is unrestricted; only predicate-passing nodes enter
 * the result list (capacity ef <= FV_GPU_MAX_EF). Results ascending by
 * (dist, node). exhausted[q] = 1 when a wider ef cannot add results
 * (frontier drained or max_visited reached). 0 / -1 + errbuf. */
int fv_gpu_graph_batch(FvGpuIndex *g, int nq, const float *queries,
                       const FvGpuQueryDesc *qd, const FvGpuPredNode *nodes,
                       int nnodes, const int32_t *children, int nchildren,
                       const double *invals, int ninvals, uint32_t ef,
                       uint64_t max_visited, FvHit *out, uint32_t *out_cnt,
                       uint8_t *exhausted, char *errbuf, size_t errlen);

/* engine_api.h:194 — query-tile width of the fast brute path for dim
 * (0 = no fast path). */
int fv_gpu_tile_variant_qt(int dim);
/* engine_api.h:195 — free/total bytes of the chosen device; returns the
 * device id or -1. */
int fv_gpu_mem_probe(uint64_t *free_bytes, uint64_t *total_bytes);

/* ================= extensions (no reference counterpart) ================= */

/* Per-call options for the brute path (a NULL pointer = all defaults). */
typedef struct FvGpuBruteOpts {
    int32_t precision;   /* FV_GPU_PREC_* to scan with; -1 = index default */
    int32_t route;       /* 0 auto, 1 force full scan, 2 force span gather */
    uint32_t overfetch;  /* quantized SQ8 scans (an SQ8 index over data
                            outside the exact INT8 window): the scan keeps
                            the overfetch x B rows nearest in code distance
                            and all are re-ranked exactly (0 = 4, at most 64,
                            at most 4096 rows). Exact scans ignore it: their
                            re-rank set is certified (every candidate within
                            2 eps of the B-th key) */
    uint32_t sample_stride; /* threshold pre-pass stride = over-fetch factor:
                               about stride x B candidates per query (0 = auto) */
    int32_t inline_pred; /* 1: evaluate one-comparison predicates inside the
                            BF16 scan instead of building bitmaps (batches
                            whose predicates all compare one column) */
    int32_t wg_per_cu;   /* main scan workgroups per CU and query tile (0 = auto) */
    int32_t qtw;         /* query tiles of 64 served per workgroup (1, 2, 4,
                            or 8: the query-resident kernel, one tile per
                            wave, predicate bitmaps only; 0 = auto; > 1
                            needs dims <= 128 and the BF16 scan) */
    int32_t graph;       /* 1: replay the run as a hipGraph captured for this
                            batch and these options (captured on first use) */
    int32_t order;       /* BF16 / INT8 scan tile order: 0 = auto (3 above
                            256 dims with several query tiles, else 1),
                            1 = each workgroup streams one contiguous chunk,
                            2 = all waves sweep the rows together, 3 = chunked
                            with a query-major grid (the query tiles reading
                            one chunk run together and share its reads) */
    int32_t waves;       /* waves per BF16 scan workgroup with one query tile
                            (4 or 8, 4 only up to 256 dims; 0 = auto: 8) */
} FvGpuBruteOpts;

/* Per-call statistics of the last batch run on this index. Kernel times are
 * measured with HIP events on the batch's stream when timing is enabled. */
typedef struct FvGpuStats {
    double t_bitmap_us, t_prepass_us, t_scan_us, t_merge_us, t_rerank_us;
    double t_gather_us, t_graph_us, t_total_us;
    uint64_t scan_rows, scan_bytes;   /* algorithmic bytes of the scan */
    uint64_t candidates;              /* keys handed to the merge */
    int32_t route;                    /* 1 scan, 2 gather, 3 graph */
    int32_t precision;
    uint32_t certified_fail;          /* queries re-run for certification */
    uint32_t nprog;                   /* distinct predicate bitmaps */
} FvGpuStats;

/* Enable (1) / disable (0) per-kernel event timing on an index. */
void fv_gpu_set_timing(FvGpuIndex *g, int on);
void fv_gpu_last_stats(const FvGpuIndex *g, FvGpuStats *st);

/* Service-time feedback of an index. No counterpart in the engine: the
 * table the reference worker keeps per index for the planner
 * (FvFeedbackEntry.gpu_q_us / gpu_batch_us / gpu_batch_n16, arbiter.h:
 * 75-104, filled at worker.c:296-333 from host wall time around each
 * batch), kept here by the library from device event time around every
 * fv_gpu_brute_batch / fv_gpu_graph_batch call, with each query's
 * selectivity measured from its predicate bitmap (or its span total). All
 * values are EWMAs, 0 = no sample yet. Strategy 0 = brute force, 1 =
 * graph; selectivity buckets (arbiter.h:50-60): >= 20%, 5-20%, 1-5%,
 * 0.2-1%, < 0.2%. */
#define FV_GPU_SEL_BUCKETS 5
typedef struct FvGpuFeedback {
    double q_us[2][FV_GPU_SEL_BUCKETS]; /* amortized per-query service time, us */
    double batch_us[2];                 /* whole-batch device time, us */
    double batch_n[2];                  /* queries per batch */
    uint64_t batches[2];                /* batches recorded */
} FvGpuFeedback;
/* Copy the index's table (0 on success, -1 for a NULL argument). */
int fv_gpu_feedback(const FvGpuIndex *g, FvGpuFeedback *out);
/* One batch into a table, the reference worker's rule: per query weight w =
 * want * 1000 / max(sel_permille, 1) for the graph (expected work), 1 for
 * brute force; per bucket b with n_b queries the sample (T / max w) x (mean
 * w over b) enters with weight n_b / (n_b + 64); the batch time and size
 * with weight 1/8; an empty field takes the sample. want may be NULL for
 * brute force. */
void fv_gpu_feedback_update(FvGpuFeedback *fb, int strategy, int n, const int32_t *sel_permille,
                            const uint32_t *want, double batch_us);
/* The selectivity bucket of a per-mille selectivity (arbiter.h:50-60). */
int fv_gpu_sel_bucket(int32_t sel_permille);

/* fv_gpu_brute_batch without any memory allocation. No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and errors as fv_gpu_brute_batch (entries past
 * out_cnt[q] in out are unspecified). Every device buffer of the call is
 * carved from one static 64 MiB scratch array per GPU (held until the
 * process ends) and host staging from a static 4 MiB array; a batch whose
 * buffers exceed them (e.g. 1024 queries at topb 4096, or more than about
 * 64 queries over 1M rows) fails with -1 and an errbuf naming the limit.
 * Calls are serialised by a process-wide lock (concurrent calls would
 * share the scratch). */
int fv_gpu_brute_batch_noalloc(FvGpuIndex *g, int nq, const float *queries,
                               const FvGpuQueryDesc *qd, const uint8_t *dedup,
                               const FvGpuPredNode *nodes, int nnodes,
                               const int32_t *children, int nchildren,
                               const double *invals, int ninvals,
                               const FvGpuSpan *spans, int nspans_total,
                               const uint64_t *spanoff, int nspanoff, uint32_t topb,
                               FvHit *out, uint32_t *out_cnt, uint8_t *truncated,
                               char *errbuf, size_t errlen);

/* fv_gpu_brute_batch with options (same arguments otherwise). */
int fv_gpu_brute_batch_ex(FvGpuIndex *g, int nq, const float *queries,
                          const FvGpuQueryDesc *qd, const uint8_t *dedup,
                          const FvGpuPredNode *nodes, int nnodes,
                          const int32_t *children, int nchildren,
                          const double *invals, int ninvals,
                          const FvGpuSpan *spans, int nspans_total,
                          const uint64_t *spanoff, int nspanoff, uint32_t topb,
                          const FvGpuBruteOpts *opts, FvHit *out,
                          uint32_t *out_cnt, uint8_t *truncated, char *errbuf,
                          size_t errlen);

/* fv_gpu_graph_batch with search knobs (no reference counterpart): wmax =
 * frontier heads expanded per iteration (1 = the CPU cursor's
 * one-at-a-time order), fcap = LDS frontier capacity (64..1024; entries
 * beyond it go to a per-query spill store in HBM and come back in order, so
 * no entry the CPU cursor's unbounded heap would pop is lost; a negative
 * fcap bounds the frontier at -fcap entries with no spill store, dropping
 * the overflow as the reference GPU kernel does), dbg = optional host array
 * [nq][4] receiving per query the frontier heads expanded (each reads one
 * neighbour list), the nodes visited (each reads one vector row), the
 * results, and the live frontier entries lost. fv_gpu_graph_batch is
 * wmax 8, fcap 1024. */
int fv_gpu_graph_batch_ex(FvGpuIndex *g, int nq, const float *queries,
                          const FvGpuQueryDesc *qd, const FvGpuPredNode *nodes,
                          int nnodes, const int32_t *children, int nchildren,
                          const double *invals, int ninvals, uint32_t ef,
                          uint64_t max_visited, int wmax, int fcap, FvHit *out,
                          uint32_t *out_cnt, uint8_t *exhausted, uint32_t *dbg,
                          char *errbuf, size_t errlen);

/*
 * Device-resident batches: stage a batch's inputs in HBM once, then run it
 * any number of times with no host<->device copy inside the run. Outputs
 * stay on the device until fv_gpu_batch_fetch. A batch belongs to one
 * index. Used by throughput measurement and by callers that keep queries
 * on the GPU. A staged batch holds up to 4096 queries (the synchronous
 * entry points above keep FV_GPU_MAX_BATCH).
 */
typedef struct FvGpuBatch FvGpuBatch;
FvGpuBatch *fv_gpu_batch_stage(FvGpuIndex *g, int nq, const float *queries,
                               const FvGpuQueryDesc *qd, const uint8_t *dedup,
                               const FvGpuPredNode *nodes, int nnodes,
                               const int32_t *children, int nchildren,
                               const double *invals, int ninvals,
                               const FvGpuSpan *spans, int nspans_total,
                               const uint64_t *spanoff, int nspanoff,
                               uint32_t topb, char *errbuf, size_t errlen);
/* Uneven batches: up to 4096 queries, each with its own k (topb_q[q] in
 * 1..topb; outputs have stride topb). Otherwise as fv_gpu_batch_stage. */
FvGpuBatch *fv_gpu_batch_stage_ex(FvGpuIndex *g, int nq, const float *queries,
                                  const FvGpuQueryDesc *qd, const uint8_t *dedup,
                                  const FvGpuPredNode *nodes, int nnodes,
                                  const int32_t *children, int nchildren,
                                  const double *invals, int ninvals,
                                  const FvGpuSpan *spans, int nspans_total,
                                  const uint64_t *spanoff, int nspanoff,
                                  const uint32_t *topb_q, uint32_t topb, char *errbuf,
                                  size_t errlen);
/* Restage (no reference counterpart; the reference re-uploads every batch,
 * worker.c:453-598): write a new set of queries, predicates and spans into
 * an existing batch's own buffers, after its pending work. topb_q may be
 * NULL (same k for every query). The staging layout is kept when the new
 * contents fit it (same query count and per-query-k mode, predicate and
 * span tables within the capacities reserved at staging: powers of two),
 * so the batch's captured run (FvGpuBruteOpts.graph) replays for the new
 * batch with no capture: a stream of batches pays one capture. Returns 1
 * when the captured run stays valid, 0 when there is none or the next graph
 * run captures again, -1 + errbuf on error. */
int fv_gpu_batch_restage(FvGpuIndex *g, FvGpuBatch *b, int nq, const float *queries,
                         const FvGpuQueryDesc *qd, const uint8_t *dedup,
                         const FvGpuPredNode *nodes, int nnodes, const int32_t *children,
                         int nchildren, const double *invals, int ninvals,
                         const FvGpuSpan *spans, int nspans_total, const uint64_t *spanoff,
                         int nspanoff, const uint32_t *topb_q, uint32_t topb, char *errbuf,
                         size_t errlen);
/* Enqueue the brute pipeline for a staged batch on the batch's stream
 * (asynchronous; 0 / -1 + errbuf). Each batch is given one of the index's
 * batch streams when it is staged, so different batches may overlap. */
int fv_gpu_batch_run_brute(FvGpuIndex *g, FvGpuBatch *b,
                           const FvGpuBruteOpts *opts, char *errbuf,
                           size_t errlen);

/* Autotuner (no reference counterpart): times the candidate launch shapes
 * of the full-scan route on this staged batch (query tiles per workgroup,
 * workgroups per CU, pre-pass sampling stride = over-fetch factor) and
 * writes the fastest whose runs certify every query to *best (precision,
 * route and inline_pred taken from base). *best_us = its time per run,
 * *ntried = shapes timed. Leaves the batch's outputs from the last trial. */
int fv_gpu_batch_tune(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *base,
                      int reps, FvGpuBruteOpts *best, double *best_us,
                      int *ntried, char *errbuf, size_t errlen);

/* Main-scan kernel timing (no reference counterpart): runs the batch once,
 * launches the same full-scan kernel reps times back to back between two
 * events on the batch's stream (*us = time per launch), then runs the
 * batch once more so its outputs are those of an ordinary run, and waits
 * for it. Fails when the batch takes the span-gather route. */
int fv_gpu_batch_time_scan(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *opts,
                           int reps, double *us, char *errbuf, size_t errlen);
/* Enqueue the graph pipeline for a staged batch (asynchronous); _ex with
 * the knobs of fv_gpu_graph_batch_ex. */
int fv_gpu_batch_run_graph(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef,
                           uint64_t max_visited, char *errbuf, size_t errlen);
int fv_gpu_batch_run_graph_ex(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef,
                              uint64_t max_visited, int wmax, int fcap, char *errbuf,
                              size_t errlen);
/* Wait for every stream of the index. */
int fv_gpu_sync(FvGpuIndex *g, char *errbuf, size_t errlen);
/* Copy a batch's results to the host (synchronous). */
int fv_gpu_batch_fetch(FvGpuIndex *g, FvGpuBatch *b, FvHit *out,
                       uint32_t *out_cnt, uint8_t *flags, char *errbuf,
                       size_t errlen);
void fv_gpu_batch_free(FvGpuBatch *b);

/* Device copy bandwidth: copies `bytes` (multiple of 16) between two fresh
 * buffers `reps` times with k_copy_calib and reports read + write GB/s.
 * The measured copy peak of roofline figures, and the known-size kernel
 * that scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings. */
int fv_gpu_copy_bench(int device, uint64_t bytes, int reps, double *gbs,
                      char *errbuf, size_t errlen);

/* Streaming read bandwidth: reads `bytes` (multiple of 16) `reps` times
 * with k_read_calib (16-byte loads, four in flight per lane,
 * nothing written) and reports GB/s read: the read-only peak a scan can
 * approach. Use a buffer well above the 256 MB last-level cache. */
int fv_gpu_read_bench(int device, uint64_t bytes, int reps, double *gbs,
                      char *errbuf, size_t errlen);

/*
 * HNSW graph construction on the host (the reference builds inside
 * fv_builder_finish, engine.cpp:698-1032; here it is a library call so a
 * GPU index can be given a graph without writing a file). Semantics are the
 * reference builder's: M links per new node, M0 = 2M at level 0, splitmix64
 * levels, efConstruction beam, heuristic pruning, node 0 first then
 * `threads` workers (threads = 1 gives the reference's single-threaded
 * graph exactly). The caller's vectors are read during the call only.
 */
typedef struct FvHnswGraph FvHnswGraph;
FvHnswGraph *fv_hnsw_build(const float *vectors, uint64_t count, int dim, int M,
                           int ef_construction, int threads, char *errbuf,
                           size_t errlen);
/* sizes: l0_stride = 2M+1 words per node; upper_slots words of upper slabs */
int fv_hnsw_info(const FvHnswGraph *h, uint32_t *M, uint32_t *l0_stride,
                 uint32_t *max_level, uint64_t *entry_node, uint64_t *upper_slots);
/* copy out (any pointer may be NULL): l0 [count][l0_stride] ([cnt, ids..]),
 * levels [count], upper_offs [count+1], upper [upper_slots] — the
 * reference index-file sections (engine.cpp:190-215) */
int fv_hnsw_export(const FvHnswGraph *h, uint32_t *l0, uint8_t *levels,
                   uint64_t *upper_offs, uint32_t *upper);
void fv_hnsw_free(FvHnswGraph *h);
/* Level-0 entry points of nq queries by greedy descent through the upper
 * levels (fv_index_graph_entry, engine.cpp:1539-1559), on `threads` host
 * threads; works on built graphs and on loaded index files alike. */
int fv_hnsw_entries(const float *vectors, int dim, const uint64_t *upper_offs,
                    const uint32_t *upper, uint32_t M, uint64_t entry_node,
                    uint32_t max_level, int nq, const float *queries,
                    uint32_t *out, int threads);
/* FP32 squared L2 distance in the reference's summation order
 * (engine.cpp:43-90), as the graph builder and entry descent use it. */
float fv_l2sq(const float *a, const float *b, int dim);

/*
 * HNSW construction with level 0 on the GPU (the reference builds on the
 * host only, engine.cpp:698-885; this call replaces fv_hnsw_build for
 * indexes already uploaded). Host phase: splitmix64 levels, node 0 and
 * every node with a level above 0 inserted with fv_hnsw_build's semantics
 * (so the upper levels are complete), then each remaining node's level-0
 * entry point by greedy descent. Device phase: the remaining nodes in index
 * order, in batches of min(max_batch, inserted / batch_div) nodes; per
 * batch, every node's efConstruction search over the level-0 graph as it
 * stood before the batch (the query cursor's beam, unfiltered), the
 * heuristic keeping M links, then per target the batch's reverse links in
 * node order: appended while the list has room, else the list and up to
 * 64 - cnt incoming links at a time re-pruned to M0 by the heuristic.
 * Deterministic for threads = 1. M <= 16 (M0 <= 32). g must hold the same
 * vectors (count x dim FP32 rows); `vectors` is read during the call only.
 */
typedef struct FvHnswGpuOpts {
    int heads;      /* frontier heads per search iteration, 1..8 (1: the
                       cursor's pop order; default 8) */
    int exact;      /* 1: unbounded frontier (spill store), 0: 1024 entries */
    int batch_div;  /* batch = inserted / batch_div (default 8) */
    int max_batch;  /* default 32768 (also capped by visited-set memory) */
    int keep;       /* 1: the built level 0 becomes g's graph (default 1;
                       work queued on the device finishes first) */
    int pad0;
    /* out: seconds of the host phase and of the device phase (wall); device
     * event time of the searches, the prunes and the link merges (0 when
     * more than 512 batches ran); nodes the searches visited and frontier
     * heads they expanded; batches */
    double t_host_s, t_device_s;
    double t_entries_s;  /* of t_host_s: the entry-point descents */
    double t_search_s, t_prune_s, t_link_s;
    uint64_t visited, heads_expanded;
    int batches;
    int pad;
} FvHnswGpuOpts;
FvHnswGraph *fv_hnsw_build_gpu(FvGpuIndex *g, const float *vectors, int M,
                               int ef_construction, int threads, FvHnswGpuOpts *opts,
                               char *errbuf, size_t errlen);
/* The host phase of fv_hnsw_build_gpu alone (the graph its device phase
 * starts from); with rest_on_host the remaining nodes are then inserted on
 * the host in index order with fv_hnsw_build's semantics (threads = 1: the
 * sequential result the device phase reproduces with batches of one). */
FvHnswGraph *fv_hnsw_build_seeded(const float *vectors, uint64_t count, int dim, int M,
                                  int ef_construction, int threads, int rest_on_host,
                                  char *errbuf, size_t errlen);

/*
 * Learned per-query beam width and brute-vs-graph routing (replaces the
 * fixed policy of exec.c:422-424 and the planner's cost race,
 * planner.c:185-346; no counterpart in the reference's GPU object).
 * Model: inputs [query (dim) | log10 selectivity | distances to the 16
 * nearest centroids] standardised, 64-64 ReLU hidden layers, outputs
 * (log ef, route logit > 0 = graph). FP32 parameters, trained on the GPU;
 * inference with BF16 weights. Parameter layout (FP32):
 *   W1[64][F] b1[64] W2[64][64] b2[64] W3[2][64] b3[2],  F = dim + 17.
 * Training targets per sample: [log ef, route (0 brute / 1 graph), ef
 * valid (0/1)]; loss = valid (y0 - log ef)^2 + lambda BCE(y1, route).
 */
typedef struct FvGpuModel FvGpuModel;
int fv_gpu_model_nparams(int dim);
FvGpuModel *fv_gpu_model_create(int device, int dim, const float *centroids, int ncent,
                                char *errbuf, size_t errlen);
void fv_gpu_model_free(FvGpuModel *m);
int fv_gpu_model_set_params(FvGpuModel *m, const float *params, const float *mean,
                            const float *std_, char *errbuf, size_t errlen);
int fv_gpu_model_get_params(FvGpuModel *m, float *params, float *mean, float *std_,
                            char *errbuf, size_t errlen);
int fv_gpu_model_grad(FvGpuModel *m, int n, const float *feats, const float *targets,
                      float lambda, float *grad, double *loss, char *errbuf, size_t errlen);
int fv_gpu_model_train(FvGpuModel *m, int n, const float *feats, const float *targets,
                       int epochs, int batch, float lr, float lambda, uint64_t seed,
                       float *loss_hist, char *errbuf, size_t errlen);
int fv_gpu_model_predict(FvGpuModel *m, int n, const float *feats, float *y,
                         char *errbuf, size_t errlen);
/* feature vectors [nq][dim+17] of a staged batch's queries */
int fv_gpu_batch_features(FvGpuIndex *g, FvGpuBatch *b, FvGpuModel *m, float *out,
                          char *errbuf, size_t errlen);

/* A routed run of a staged batch (staged with topb = k): features, BF16
 * MLP, then brute (BF16 scan + exact re-rank) for queries routed to it and
 * graph search with the per-query ef for the rest; k hits per query land in
 * the batch outputs (fv_gpu_batch_fetch). Modes: 0 learned route + learned
 * ef, 1 graph for all with learned ef, 2 graph for all with ef_fixed (the
 * fixed-ef fallback), 3 brute for all. */
typedef struct FvGpuLearnedOpts {
    uint32_t k;            /* hits per query (= staged topb) */
    int32_t mode;
    uint32_t ef_fixed;     /* mode 2 */
    uint32_t ef_min, ef_max; /* clamp of the predicted ef (ef_min >= k) */
    float ef_scale;        /* safety factor on the predicted ef (0 = 1) */
    float route_bias;      /* graph when the route logit exceeds it (+inf: never;
                              the run is then the brute pipeline alone) */
    uint64_t max_visited;  /* 0 = max(200000, 400 ef) */
    int32_t precision;     /* brute scan precision, -1 = index default */
    int32_t graph_cus;     /* reserved, must be 0: the graph-routed searches
                              share the chip with the packed brute scan (a
                              split of the CUs between the two is not
                              available; non-zero fails the call) */
} FvGpuLearnedOpts;
int fv_gpu_batch_run_learned(FvGpuIndex *g, FvGpuBatch *b, FvGpuModel *m,
                             const FvGpuLearnedOpts *o, char *errbuf, size_t errlen);
/* per-query decisions of the batch's last routed run (synchronous); any
 * output may be NULL; y [nq][2] = the model outputs (model modes only) */
int fv_gpu_batch_routes(FvGpuIndex *g, FvGpuBatch *b, uint8_t *route, uint32_t *ef,
                        float *y, char *errbuf, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif /* PAVES_AMD_GPU_H */
