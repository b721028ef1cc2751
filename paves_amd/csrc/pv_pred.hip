/*
 * pv_pred.hip — predicate programs and the predicate-to-bitmap kernel.
 *
 * The ABI hands over predicate trees in the serialized form of
 * engine_api.h:95-117 (nodes with absolute child/value ranges, one root per
 * query in FvGpuQueryDesc.proot, gpu_engine.h:46-51). The host turns each
 * query's tree into a postfix program (n-ary AND/OR folded into binary ops
 * so the evaluation stack depth is bounded by the tree depth), validates
 * every index, and merges byte-identical programs: queries with the same
 * predicate share one bitmap. The bitmap kernel evaluates each distinct
 * program for every row — the waves of a block split the programs, one
 * ballot per 64 rows — and stores bit (row & 31) of word [prog][row >> 5].
 *
 * Semantics follow FvPred::eval_node (engine.cpp:397-430) and the CUDA
 * interpreter (gpu_engine.cu:238-279): comparisons in double on the f64
 * columns, NULL = NaN, NE = v == v && v != c.
 */
#include <string.h>


#include "pv_device.h"

namespace pv {

namespace {

struct Emitter {
    const FvGpuPredNode *nodes;
    int nnodes;
    const int32_t *children;
    int nchildren;
    int ninvals;
    int ncols;
    std::vector<PvOp> *ops;
    int sp = 0, maxsp = 0;
    const char *err = nullptr;

    void push() {
        if (++sp > maxsp) maxsp = sp;
    }
    bool emit(int32_t ni, int depth) {
        if (depth > kMaxProgDepth) { err = "predicate nesting too deep"; return false; }
        if (ni < 0 || ni >= nnodes) { err = "predicate node index out of range"; return false; }
        const FvGpuPredNode &n = nodes[ni];
        PvOp op = {};
        switch (n.kind) {
            case FV_NODE_CMP:
                if (n.col < 0 || n.col >= ncols) { err = "predicate column out of range"; return false; }
                if (n.op < FV_CMP_LT || n.op > FV_CMP_NE) { err = "bad comparison op"; return false; }
                op.kind = OP_CMP; op.col = n.col; op.a = n.op; op.value = n.value;
                op.b = kCmpPassMask[n.op];
                ops->push_back(op);
                push();
                return true;
            case FV_NODE_IN:
                if (n.col < 0 || n.col >= ncols) { err = "predicate column out of range"; return false; }
                if (n.nvalues < 0 || n.val0 < 0 || (int64_t)n.val0 + n.nvalues > ninvals) {
                    err = "IN value range out of bounds"; return false;
                }
                op.kind = OP_IN; op.col = n.col; op.a = n.val0; op.b = n.nvalues;
                ops->push_back(op);
                push();
                return true;
            case FV_NODE_AND:
            case FV_NODE_OR: {
                if (n.nchildren < 0 || n.child0 < 0 ||
                    (int64_t)n.child0 + n.nchildren > nchildren) {
                    err = "child range out of bounds"; return false;
                }
                const bool is_and = n.kind == FV_NODE_AND;
                if (n.nchildren == 0) {
                    /* empty conjunction is true, empty disjunction false */
                    op.kind = is_and ? OP_TRUE : OP_FALSE;
                    ops->push_back(op);
                    push();
                    return true;
                }
                for (int i = 0; i < n.nchildren; i++) {
                    if (!emit(children[n.child0 + i], depth + 1)) return false;
                    if (i > 0) {
                        op.kind = is_and ? OP_AND2 : OP_OR2;
                        ops->push_back(op);
                        sp--;
                    }
                }
                return true;
            }
            case FV_NODE_NOT:
                if (n.nchildren < 1 || n.child0 < 0 || n.child0 >= nchildren) {
                    err = "NOT without child"; return false;
                }
                if (!emit(children[n.child0], depth + 1)) return false;
                op.kind = OP_NOT;
                ops->push_back(op);
                return true;
        }
        err = "unknown predicate node kind";
        return false;
    }
};

uint64_t fnv(const void *p, size_t n, uint64_t h) {
    const unsigned char *c = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= c[i];
        h *= 1099511628211ull;
    }
    return h;
}

/* content equality of two programs (IN values compared by value) */
bool same_prog(const ProgSet &ps, int32_t a_off, int32_t a_len,
               const std::vector<PvOp> &ops, int32_t b_off, int32_t b_len,
               const double *invals) {
    if (a_len != b_len) return false;
    for (int i = 0; i < a_len; i++) {
        const PvOp &x = ps.ops[a_off + i], &y = ops[b_off + i];
        if (x.kind != y.kind || x.col != y.col) return false;
        if (x.kind == OP_CMP) {
            if (x.a != y.a || memcmp(&x.value, &y.value, 8) != 0) return false;
        } else if (x.kind == OP_IN) {
            if (x.b != y.b) return false;
            if (memcmp(invals + x.a, invals + y.a, 8 * (size_t)x.b) != 0) return false;
        }
    }
    return true;
}

} // namespace

int compile_programs(int nq, const FvGpuQueryDesc *qd,
                     const FvGpuPredNode *nodes, int nnodes,
                     const int32_t *children, int nchildren,
                     const double *invals, int ninvals, int ncols,
                     ProgSet *ps, char *errbuf, size_t errlen) {
    ps->ops.clear();
    ps->vals.assign(invals, invals + (ninvals > 0 ? ninvals : 0));
    ps->off.clear();
    ps->len.clear();
    ps->of_query.assign(nq, -1);
    /* distinct programs by content hash: a flat open-addressed table of
     * program indices, at most half full (at most nq programs) */
    size_t cap = 16;
    while (cap < 2 * (size_t)(nq > 0 ? nq : 1)) cap <<= 1;
    std::vector<int32_t> slot(cap, -1);
    std::vector<uint64_t> phash;
    std::vector<PvOp> tmp;
    for (int q = 0; q < nq; q++) {
        if (qd[q].proot < 0) continue;
        if (nodes == nullptr) PV_FAIL("gpu: predicate root without nodes (query %d)", q);
        tmp.clear();
        Emitter em;
        em.nodes = nodes;
        em.nnodes = nnodes;
        em.children = children;
        em.nchildren = children ? nchildren : 0;
        em.ninvals = invals ? ninvals : 0;
        em.ncols = ncols;
        em.ops = &tmp;
        if (!em.emit(qd[q].proot, 0)) PV_FAIL("gpu: query %d: %s", q, em.err);
        if (em.maxsp > 63) PV_FAIL("gpu: query %d: predicate stack too deep", q);
        /* content hash: op fields + IN values */
        uint64_t h = 1469598103934665603ull;
        for (const PvOp &op : tmp) {
            h = fnv(&op.kind, 8, h);
            if (op.kind == OP_CMP) h = fnv(&op.a, 4, fnv(&op.value, 8, h));
            if (op.kind == OP_IN) h = fnv(invals + op.a, 8 * (size_t)op.b, fnv(&op.b, 4, h));
        }
        int32_t found = -1;
        size_t i = (size_t)(h ^ (h >> 29)) & (cap - 1);
        for (; slot[i] >= 0; i = (i + 1) & (cap - 1)) {
            const int32_t p = slot[i];
            if (phash[(size_t)p] == h &&
                same_prog(*ps, ps->off[p], ps->len[p], tmp, 0, (int32_t)tmp.size(), invals)) {
                found = p;
                break;
            }
        }
        if (found < 0) {
            /* programs are numbered in order of first use */
            found = (int32_t)ps->off.size();
            ps->off.push_back((int32_t)ps->ops.size());
            ps->len.push_back((int32_t)tmp.size());
            ps->ops.insert(ps->ops.end(), tmp.begin(), tmp.end());
            phash.push_back(h);
            slot[i] = found;
        }
        ps->of_query[q] = found;
    }
    return 0;
}

/* All distinct programs over kBmRows rows per block (bitmap_rows). */
__global__ __launch_bounds__(256) void k_bitmap(
    const int32_t *__restrict__ hdr, const PvOp *__restrict__ ops,
    const int32_t *__restrict__ poff, const int32_t *__restrict__ plen,
    const double *__restrict__ vals, const double *__restrict__ cols, uint64_t count,
    int ncols, uint64_t words, uint32_t *__restrict__ bitmap) {
    bitmap_rows(blockIdx.x, hdr, ops, poff, plen, vals, cols, count, ncols, words, bitmap,
                nullptr, 0);
}

hipError_t launch_bitmap(FvGpuIndex *g, FvGpuBatch *b, hipStream_t s) {
    if (b->nprog == 0) return hipSuccess;
    if (b->nprog > kBmProgs || b->d_bitmap.p == nullptr) return hipErrorInvalidValue;
    const uint64_t words = g->count_pad / 32;
    const unsigned blocks = bitmap_blocks(g);
    hipLaunchKernelGGL(k_bitmap, dim3(blocks), dim3(256), bitmap_lds_bytes(g->ncols), s,
                       b->d_hdr, b->d_ops, b->d_poff, b->d_plen, b->d_vals, g->d_cols,
                       g->count, g->ncols, words, b->d_bitmap.as<uint32_t>());
    return hipGetLastError();
}

} // namespace pv
