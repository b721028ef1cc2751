/*
 * pv_internal.h — internal types of the MI355X query engine.
 *
 * Device data layout (one index, resident in HBM):
 *   vecs   fp32  [count_pad][dim]   row-major, rows >= count are zero
 *   vbf    bf16  [count_pad][dim]   RNE copy for the MFMA scan (dim % 16 == 0)
 *   norm   fp32  [count_pad]        ||x||^2 (0 for padding rows)
 *   cols   f64   [ncols][count]     scalar columns, NULL = NaN
 *   prows  u32   [ncols][count]     posting rows (rows sorted by value)
 *   nbr    u32   [count][nbr_stride] level-0 adjacency, 0xFFFFFFFF-padded
 * count_pad is count rounded up to 64 so 32-row tiles never read past the
 * allocation.
 */
#ifndef PV_INTERNAL_H
#define PV_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <mutex>
#include <vector>

#include "../../include/paves_gpu.h"

#define PV_TRY(call)                                                          \
    do {                                                                      \
        hipError_t e__ = (call);                                              \
        if (e__ != hipSuccess) {                                              \
            snprintf(errbuf, errlen, "gpu: %s failed: %s", #call,             \
                     hipGetErrorString(e__));                                 \
            return -1;                                                        \
        }                                                                     \
    } while (0)

#define PV_FAIL(...)                                                          \
    do {                                                                      \
        snprintf(errbuf, errlen, __VA_ARGS__);                                \
        return -1;                                                            \
    } while (0)

namespace pv {

constexpr uint32_t kNoNode = 0xFFFFFFFFu;
constexpr int kQT = 64;          /* queries per scan tile (2 MFMA column blocks) */
constexpr int kRT = 32;          /* rows per wave tile */
constexpr int kLdsCand = 32;     /* per-query LDS candidate slots in a scan WG */
constexpr int kSelMax = 4096;    /* max keys sorted in one select WG */
constexpr int kMaxProgDepth = 64;

constexpr int kMaxBatchEx = 4096; /* queries per batch through the extension API */

/* pool counters sit one 128-byte line apart: every scan WG reserves pool
 * space with an atomic per query, and counters sharing a line would
 * serialise all of them in one L2 channel */
constexpr uint32_t kPcntStride = 32;
/* Each query's pool is split into kPoolShards equal segments, each with its
 * own counter; scan WG x reserves in segment x % kPoolShards. All WGs flush
 * at the end of the scan, and one counter per query took the reservations
 * of every WG of the grid back to back (measured: 5 us of the flush at 256
 * WGs); the select stage reads the segments through a prefix of their
 * counts. */
constexpr uint32_t kPoolShards = 8;


/* one comparison evaluated inside the scan (col -1: no predicate) */
struct InlPred {
    int32_t col;
    int32_t mask;   /* kCmpPassMask of the comparison */
    double value;
    /* the same test on the row's rank in the column's posting order (non-NULL
     * values ascending, NULLs last): pass iff lo <= rank < hi (set when the
     * index holds ranks and the comparison is one rank range) */
    uint32_t lo, hi;
    uint32_t pad[2];
};
static_assert(sizeof(InlPred) == 32, "InlPred layout");

/* status bits per query */
constexpr uint32_t kStOverflow = 1u;   /* candidate pool overflowed */
constexpr uint32_t kStCertFail = 2u;   /* re-rank set exceeded capacity */
constexpr uint32_t kStPrepFail = 4u;   /* query outside the INT8 window of an INT8 scan */

/* predicate program op (postfix; see pv_pred.hip) */
enum : int32_t { OP_CMP = 0, OP_IN = 1, OP_AND2 = 2, OP_OR2 = 3, OP_NOT = 4,
                 OP_TRUE = 5, OP_FALSE = 6 };
struct PvOp {
    int32_t kind;
    int32_t col;   /* CMP/IN column */
    int32_t a;     /* CMP: FvCmpOp; IN: first value index */
    int32_t b;     /* IN: number of values */
    double value;  /* CMP constant */
};
static_assert(sizeof(PvOp) == 24, "PvOp layout");

/*
 * A CMP op carries in `b` the set of comparison outcomes that pass, bit o
 * for outcome o: 0 (v < c), 1 (v == c), 2 (v > c), 3 (v is NULL/NaN),
 * 4 (c is NaN, v is not). Indexed by FvCmpOp: LT, LE, EQ, GE, GT, NE;
 * NE passes v == v && v != c (engine.cpp:397-430).
 */
constexpr int32_t kCmpPassMask[6] = {0x01, 0x03, 0x02, 0x06, 0x04, 0x15};

/* A bump arena over caller-provided memory: buffers bound to one take
 * their space from it instead of allocating (the allocation-free entry
 * point binds a batch's buffers to the library's static device scratch). */
struct Arena {
    unsigned char *base = nullptr;
    size_t size = 0, used = 0;
    void *take(size_t n) {
        const size_t at = (used + 255) & ~(size_t)255;
        if (base == nullptr || at + n > size) return nullptr;
        used = at + n;
        return base + at;
    }
};

/* grow-only device buffer (from the arena when bound to one) */
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    Arena *arena = nullptr;
    hipError_t ensure(size_t need) {
        if (cap >= need && p) return hipSuccess;
        if (arena) {
            p = arena->take(need);
            cap = p ? need : 0;
            return p ? hipSuccess : hipErrorOutOfMemory;
        }
        if (p) hipFree(p);
        p = nullptr;
        cap = 0;
        size_t nc = need + need / 4 + 256;
        hipError_t e = hipMalloc(&p, nc);
        if (e == hipSuccess) cap = nc;
        return e;
    }
    void release() {
        if (p && !arena) hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T> T *as() const { return (T *)p; }
};

/* grow-only pinned host buffer (from a host arena when bound to one) */
struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    Arena *arena = nullptr;
    hipError_t ensure(size_t need) {
        if (cap >= need && p) return hipSuccess;
        if (arena) {
            p = arena->take(need);
            cap = p ? need : 0;
            return p ? hipSuccess : hipErrorOutOfMemory;
        }
        if (p) hipHostFree(p);
        p = nullptr;
        cap = 0;
        size_t nc = need + need / 4 + 256;
        hipError_t e = hipHostMalloc(&p, nc, hipHostMallocDefault);
        if (e == hipSuccess) cap = nc;
        return e;
    }
    void release() {
        if (p && !arena) hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T> T *as() const { return (T *)p; }
};

/* compiled predicate programs of one batch (host side) */
struct ProgSet {
    std::vector<PvOp> ops;
    std::vector<double> vals;
    std::vector<int32_t> off, len;   /* per distinct program */
    std::vector<int32_t> of_query;   /* per query: program index, -1 = none */
};

/* Packed staging area of a batch: byte offsets of its regions and the
 * element capacities of the regions whose size depends on the predicates
 * and spans. A batch restaged with the same query count, per-query-k mode
 * and contents that fit these capacities keeps every offset, so every
 * device pointer a captured run baked in stays valid. */
struct StageLayout {
    int nq = 0;              /* 0 = no layout yet */
    bool tq = false;         /* per-query k region present */
    size_t o_q = 0, o_qd = 0, o_hdr = 0, o_pofq = 0, o_inl = 0, o_tq = 0;
    size_t o_ops = 0, o_vals = 0, o_poff = 0, o_plen = 0, o_sp = 0, o_so = 0;
    size_t c_ops = 0, c_vals = 0, c_prog = 0, c_sp = 0, c_so = 0;
    size_t total = 0;
};

} // namespace pv

/* A staged batch: inputs resident in HBM plus device outputs. */
struct FvGpuBatch {
    FvGpuIndex *owner = nullptr;
    int nq = 0;
    int nqt = 0;            /* query tiles of kQT */
    uint32_t topb = 0;
    int nprog = 0;          /* distinct predicate programs */
    int nops = 0;           /* ops of all programs (contiguous, program order) */
    int nvals = 0;          /* IN values */
    bool all_spans = false; /* every query has spans and no dedup */
    bool any_dedup = false; /* some query's spans may repeat rows */
    uint64_t sum_total = 0; /* sum of span enumeration sizes */
    uint64_t max_total = 0;
    uint64_t max_entry = 0; /* largest qd[q].entry (graph runs) */
    bool has_topbq = false; /* per-query k (uneven batches); topb = the largest */
    const uint32_t *d_topbq_p = nullptr;   /* u32 [nq] in the staging area */
    /* inline predicates: every predicate is one comparison on column
     * inl_col, evaluated inside the scan instead of through a bitmap */
    bool inl_ok = false;
    int inl_col = -1;
    /* every inline predicate is one rank range of inl_col (InlPred.lo/hi) */
    bool inl_rank = false;
    const pv::InlPred *d_inl_p = nullptr;  /* [nqp] in the staging area */
    /* a captured run (FvGpuBruteOpts.graph): replayed while the options
     * it was captured with and every device buffer and size baked into its
     * kernel arguments (gsig) are unchanged */
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    FvGpuBruteOpts gopts = {};
    std::vector<uint64_t> gsig;
    bool warmed = false;    /* one eager run done: every buffer allocated */
    bool ran_bitmap = false;  /* the last run built predicate bitmaps */
    bool graph_bitmap = false; /* the captured run builds predicate bitmaps */
    bool arena_bound = false; /* buffers carved from an arena: keep no host records */
    void *last_scan = nullptr; /* launch record of the last main scan (timing) */
    bool capturing = false; /* a hipGraph capture of this batch's run is open */
    bool i8_ok = false;     /* the INT8 scan serves every query (mode 2: always;
                               mode 1: every query an integer in the window) */
    pv::DevBuf d_qi8;       /* [nqp][dim] int8 queries minus the index centre */
    pv::DevBuf d_qn8;       /* [nqp] ||q - centre||^2 */
    bool query_exact = true;
    std::vector<uint64_t> total;   /* host copy of qd[q].total */
    std::vector<int32_t> nspans;

    /* packed staging area: one H2D copy per stage */
    pv::PinBuf h_stage;
    pv::DevBuf d_stage;
    pv::StageLayout lay;
    const int32_t *d_hdr = nullptr; /* device [nprog, nops, nvals] (staging area):
                                       kernels read the program counts here, so
                                       a captured run serves restaged batches */
    float *d_q = nullptr;          /* [nqt*64][dim] fp32 (zero padded) */
    FvGpuQueryDesc *d_qd = nullptr;
    pv::PvOp *d_ops = nullptr;
    double *d_vals = nullptr;
    int32_t *d_poff = nullptr, *d_plen = nullptr, *d_pofq = nullptr;
    FvGpuSpan *d_spans = nullptr;
    uint64_t *d_spanoff = nullptr;

    /* derived per-query data */
    pv::DevBuf d_qbf;     /* [nqt*64][dim] bf16 */
    pv::DevBuf d_qaux;    /* qnorm f32 [nq_pad], eps f32 [nq_pad], tau f32 [nq_pad] */
    pv::DevBuf d_bitmap;  /* [nprog][words] u32 */
    pv::DevBuf d_gmin;    /* prepass group minima [nq][ngroups] (FP32 bits) */
    pv::DevBuf d_pool;    /* candidate pool [nq][cap_pool] u64 */
    pv::DevBuf d_pcnt;    /* [nq] u32 pool counts */
    uint64_t cap_pool = 0;
    uint32_t ngroups = 0;
    uint32_t stride = 0;

    /* outputs */
    pv::DevBuf d_out;     /* FvHit [nq][topb] */
    pv::DevBuf d_cnt;     /* u32 [nq] */
    pv::DevBuf d_flag;    /* u8 [nq] truncated / exhausted */
    pv::PinBuf h_out;     /* fetch staging */
    uint32_t out_stride = 0; /* hits per query in d_out */
    int last_route = 0;
    hipStream_t stream = nullptr; /* one of the index's batch streams */
    pv::DevBuf d_visit;   /* graph: overflow visited bitmaps */
    pv::DevBuf d_order;   /* graph: launch order [nq] (heaviest search first) and
                             passing rows per program [nprog] */
    pv::DevBuf d_fspill;  /* graph: frontier spill stores */
    pv::DevBuf d_progcnt; /* [nprog][bitmap blocks] passing rows per program and
                             bitmap block (learned features) */
    pv::DevBuf d_feat, d_y; /* learned run: features [nq][F], model outputs [nq][2] */
    pv::DevBuf d_route, d_efq; /* learned run: route u8 [nq] (1 = graph), ef u32 [nq] */
    /* learned run over several query tiles: the brute-routed queries packed
     * to the front of this companion batch (sharing the parent's predicate
     * programs and bitmaps), so the scan runs only the tiles they fill */
    FvGpuBatch *view = nullptr;
    pv::DevBuf d_vmap;    /* view slot -> parent query [nqp] (i32) */
    uint8_t *d_vskip = nullptr; /* a view's skip flags [nqp] (its staging area) */
    /* a routed run's graph search runs beside the packed brute pipeline on
     * this second stream (default priority), joined by events */
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

struct FvGpuIndex {
    int device = -1;
    int precision = FV_GPU_PREC_FP32;
    uint64_t count = 0;
    uint64_t count_pad = 0;
    int dim = 0;
    int dimb = 0;            /* dims of the BF16 tiled copy: dim rounded up to
                                a multiple of 128, zero-filled (0 = no copy) */
    int ncols = 0;
    uint32_t l0_stride = 0;
    int nbr_stride = 0;
    int ncu = 256;

    float *d_vecs = nullptr;
    uint16_t *d_vbf = nullptr;
    float *d_norm = nullptr;
    double *d_cols = nullptr;
    uint32_t *d_prows = nullptr;
    /* rank of every row in its column's posting order [ncols][count] (when
     * the postings are given and sorted: non-NULL values ascending, NULLs
     * last), and on the host each column's sorted non-NULL values, to turn
     * a one-comparison predicate into a rank range at staging */
    uint32_t *d_rank = nullptr;
    std::vector<std::vector<double>> h_sorted;
    uint32_t *d_nbr = nullptr;
    uint32_t *d_ones = nullptr;  /* all-ones bitmap row [count_pad/32] */
    /* INT8 copy of an SQ8 index, in the MFMA-tiled int8 layout over dimp
     * dims (dim rounded up to 64, 128 or a multiple of 256; zero codes past
     * dim), and the squared code norms. i8_mode 1 (exact): integer data
     * spanning <= 255 values, up to 256 dims — codes x - i8_center, the
     * scan is exact. i8_mode 2 (quantized): any other data up to 1024 dims
     * — codes clamp(rint((x - i8_mid[d]) * i8_inv), -127, 127) with the
     * per-dim midrange i8_mid and one scale; the scan ranks by the code
     * distance and the top overfetch x k rows are re-ranked exactly.
     * i8_mode 3 (FP8, FV_GPU_PREC_FP8): the same rows as E4M3 codes
     * e4m3_encode((x - i8_mid[d]) * i8_inv), one byte per dim in the same
     * tiled layout (v_mfma_f32_32x32x16_fp8_fp8 takes each lane's 16 bytes
     * as two 8-byte operands); squared code norms from the decoded codes;
     * ranked and re-ranked as mode 2. */
    int8_t *d_vi8 = nullptr;
    float *d_norm8 = nullptr;
    int i8_center = 0;
    int i8_mode = 0;
    int dimp = 0;
    float *d_i8mid = nullptr;   /* [dimp] (modes 2, 3) */
    float i8_inv = 1.f;

    float xmax_norm = 0.f;   /* max ||x|| */
    float vmax_abs = 0.f;    /* max |x_i| */
    float vmin = 0.f, vmax = 0.f; /* min / max coordinate value (0 included) */
    bool all_int = false;    /* every vector value is an integer */

    hipStream_t stream = nullptr;   /* upload and index-wide work */
    /* Batches are spread over a few streams (fixed per batch, so runs of
     * one batch stay ordered) so that the latency-bound small kernels of
     * one batch overlap the HBM-bound scan of another. */
    static constexpr int kMaxBatchStreams = 8;
    int nstreams = 3;                /* PAVES_BATCH_STREAMS (1..8), default 3 */
    hipStream_t bstreams[kMaxBatchStreams] = {};

    int next_stream = 0;
    hipStream_t ev_stream = nullptr; /* stream of the run being timed */
    bool timing = false;
    hipEvent_t ev[16] = {};
    FvGpuStats last = {};
    /* service-time feedback of the synchronous entry points: the table, its
     * lock, and the two events bracketing a call's device work */
    FvGpuFeedback fb = {};
    std::mutex fb_mu;
    hipEvent_t fb_ev[2] = {nullptr, nullptr};

    FvGpuBatch *sync_batch = nullptr; /* reused by the synchronous ABI */
    FvGpuBatch *last_batch = nullptr; /* batch of the last brute run (stats) */
};

namespace pv {
/* row length of a batch's BF16 query operands: the tiled copy's padded
 * dims, or dim for an index without the copy */
inline int qbf_dim(const FvGpuIndex *g) { return g->dimb > 0 ? g->dimb : g->dim; }
} // namespace pv

namespace pv {

/* per-query auxiliary array layout inside d_qaux (floats) */
struct QAux {
    float *qnorm, *eps, *tau;
    uint32_t *status;
};
inline QAux qaux_of(FvGpuBatch *b) {
    const int nqp = b->nqt * kQT;
    float *base = b->d_qaux.as<float>();
    QAux a;
    a.qnorm = base;
    a.eps = base + nqp;
    a.tau = base + 2 * nqp;
    a.status = (uint32_t *)(base + 3 * nqp);
    return a;
}

/* pv_pred.hip */
int compile_programs(int nq, const FvGpuQueryDesc *qd,
                     const FvGpuPredNode *nodes, int nnodes,
                     const int32_t *children, int nchildren,
                     const double *invals, int ninvals, int ncols,
                     ProgSet *ps, char *errbuf, size_t errlen);
hipError_t launch_bitmap(FvGpuIndex *g, FvGpuBatch *b, hipStream_t s);

/* pv_brute.hip */
int stage_batch(FvGpuIndex *g, FvGpuBatch *b, int nq, const float *queries,
                const FvGpuQueryDesc *qd, const uint8_t *dedup,
                const FvGpuPredNode *nodes, int nnodes,
                const int32_t *children, int nchildren, const double *invals,
                int ninvals, const FvGpuSpan *spans, int nspans_total,
                const uint64_t *spanoff, int nspanoff, uint32_t topb,
                char *errbuf, size_t errlen, const uint32_t *topb_q = nullptr,
                int max_nq = FV_GPU_MAX_BATCH, bool keep_layout = false);
void free_batch_buffers(FvGpuBatch *b);

/* pv_graph.hip */
int run_graph(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef, uint64_t max_visited,
              char *errbuf, size_t errlen);
/* options of a graph run inside a routed (learned) run */
struct GraphExtra {
    const uint8_t *route = nullptr;  /* device [nq]: 0 = brute, leave the query alone */
    const uint32_t *efq = nullptr;   /* device [nq]: per-query ef (<= ef) */
    uint32_t out_k = 0;              /* hits per query written (0 = ef) */
    bool bitmap_ready = false;       /* predicate bitmaps already built this run */
    hipStream_t stream = nullptr;    /* launch here instead of the batch's stream */
};
int run_graph_ex(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef, uint64_t max_visited, int wmax,
                 int fcap, uint32_t *dbg, const GraphExtra &x, char *errbuf, size_t errlen);
/* pv_graph.hip: unfiltered searches of the device graph builder over the
 * adjacency being built (nbr, nbr_stride ids per node, kNoNode-padded):
 * query i is qf[i] with entry point qd[i].entry; the top ef hits of query
 * i land in out[i * ef ...] (ascending (dist, node)), their number in cnt.
 * Buffers grow to nq_reserve queries on the first call. */
struct BuildSearchBufs {
    DevBuf visit, spill, status, out, cnt, exh;
    void release() {
        visit.release();
        spill.release();
        status.release();
        out.release();
        cnt.release();
        exh.release();
    }
};
int build_search(FvGpuIndex *g, BuildSearchBufs &B, const uint32_t *nbr, int nbr_stride,
                 const float *qf, const FvGpuQueryDesc *qd, const int32_t *pofq, int nq,
                 int nq_reserve, uint32_t ef, int wmax, bool exact, uint32_t *dbg,
                 hipStream_t s, char *errbuf, size_t errlen);
/* pv_brute.hip: service-time feedback of a synchronous call. begin before
 * its device work on stream s, end after it (also counts the programs'
 * passing rows when the run built bitmaps), record once the call's results
 * are on the host (strategy 0 brute, 1 graph; want = hits asked per query).
 * Failures leave the table as it was; they never fail the call. */
void feedback_begin(FvGpuIndex *g, hipStream_t s);
void feedback_end(FvGpuIndex *g, FvGpuBatch *b, hipStream_t s, bool bitmaps);
void feedback_record(FvGpuIndex *g, FvGpuBatch *b, int strategy, const FvGpuQueryDesc *qd,
                     uint32_t want, bool bitmaps);
/* pv_brute.hip: the brute pipeline; skip (device [nq], may be NULL) marks
 * queries it must leave alone; have_phase1 = bitmaps and query prep done;
 * decide (may be NULL): the routing decision, run in the threshold launch,
 * that writes skip */
struct LearnArgs;
int run_brute_ex(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *o,
                 uint32_t stride_override, const uint8_t *skip, bool have_phase1,
                 char *errbuf, size_t errlen, const LearnArgs *decide = nullptr);
/* the routing decision's blocks alone (features already in A.feat_out) */
hipError_t launch_route_blocks(const LearnArgs &A, int nq, hipStream_t s);
/* router features computed by the query prep of phase 1 (dim <= 256) */
struct PrepFeatures {
    const float *cent_t;
    int ncent, nfeat;
    float *feat;
};
hipError_t launch_phase1(FvGpuIndex *g, FvGpuBatch *b, int prec, bool bitmap, hipStream_t s,
                         uint32_t *progpart, const PrepFeatures *pf = nullptr);
/* precision the brute scan runs in for these options (FP32 or BF16) */
int scan_precision(const FvGpuIndex *g, const FvGpuBatch *b, const FvGpuBruteOpts *o);

/* timing helpers (pv_index.hip) */
void ev_record(FvGpuIndex *g, int i);
double ev_elapsed_us(FvGpuIndex *g, int a, int b);

} // namespace pv

#endif
