/*
 * pv_graph.hip — batched filtered level-0 beam search on MI355X.
 *
 * Replaces fv_gpu_graph_batch / fv_graph_kernel (gpu_engine.cu:1118-1485,
 * 2338-2542). Semantics follow the CPU HnswCursor (engine.cpp:1631-1712):
 * traversal is unrestricted, only predicate-passing nodes enter the result
 * list (capacity ef); a neighbour enters the frontier when the result list
 * is not full or it beats the worst result; stop when the best frontier
 * entry is worse than the worst result with the list full, when visited
 * reaches max_visited (exhausted), or when the frontier drains (exhausted).
 *
 * CDNA4 design: one 64-lane wave per query (a block is one wave, so every
 * barrier is wave-local and cheap; many queries per CU hide HBM latency).
 *   - frontier and results are sorted key arrays in LDS (key = dist bits
 *     << 32 | node, so key order = (dist, node) order, the CPU heaps'
 *     pair order); new entries are merged wave-parallel (bitonic sort of
 *     the new keys, then merge-path ranks by binary search), truncated to
 *     the frontier capacity / ef;
 *   - visited: an LDS open-addressing hash (linear probing, bounded probe
 *     length); a node whose probe window is full goes to a per-query
 *     global bitmap that the wave zero-fills the first time it needs it —
 *     exact membership, no pre-launch memset;
 *   - each iteration pops up to W frontier heads that pass the stop test,
 *     gathers their padded neighbour rows (one 128-byte line per node for
 *     M0 = 32), and computes exact FP32 distances in the CPU cursor's
 *     summation order, 16 lanes per node, 32 nodes per memory round trip;
 *   - the predicate is a bit test in the batch's predicate bitmap.
 * With W = 1 and a frontier that never overflows, the search visits the
 * nodes the CPU cursor visits and returns the same list.
 */
#include <math.h>
#include <string.h>

#include <algorithm>

#include "pv_device.h"

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

namespace pv {

constexpr int kHashBits = 11;
constexpr int kHashSize = 1 << kHashBits; /* 2048 slots */
constexpr int kProbe = 16;
constexpr int kMaxNew = 256;              /* W * nbr_stride <= 256 (W <= 8 at degree 32) */
constexpr int kNS = kMaxNew / 64;         /* new-node slots per lane */

struct GraphArgs {
    const float *vecs;
    const uint32_t *nbr;
    int nbr_stride;
    uint64_t count;
    int dim;
    const float *qf;
    const FvGpuQueryDesc *qd;
    const int32_t *pofq;
    const uint32_t *bitmap;
    uint64_t words;
    uint64_t vwords;     /* words per overflow bitmap, multiple of 4 */
    uint32_t ef;
    uint64_t max_visited;
    int wmax;
    int fcap;
    uint32_t *visit;     /* [nq][words] overflow bitmaps */
    FvHit *out;
    uint32_t *out_cnt;
    uint8_t *out_exh;
    uint32_t *status;    /* per-query status words, cleared by the search */
    const uint8_t *route; /* optional: queries with route 0 are left alone */
    const uint32_t *efq;  /* optional per-query ef (<= ef, the LDS sizing) */
    uint32_t out_k;       /* hits written per query (stride); 0 = ef */
    uint32_t *dbg;       /* [nq][4] iterations, visited, results, overflow */
};

__device__ __forceinline__ uint32_t hash_slot(uint32_t node) {
    return (node * 2654435761u) >> (32 - kHashBits);
}

/* Test-and-set in the LDS hash; returns 1 if newly inserted, 0 if already
 * present, 2 if the probe window is full (caller uses the global bitmap). */
__device__ __forceinline__ int hash_insert(uint32_t *h, uint32_t node) {
    uint32_t s = hash_slot(node);
    for (int p = 0; p < kProbe; p++) {
        uint32_t *slot = &h[(s + p) & (kHashSize - 1)];
        uint32_t old = atomicCAS(slot, kNoNode, node);
        if (old == kNoNode) return 1;
        if (old == node) return 0;
    }
    return 2;
}

/* ascending sort of the first m (<= kMaxNew) keys in LDS by one wave,
 * padding with ~0 up to the next power of two >= 64: up to 64 keys in
 * registers (shuffle bitonic, no barrier), else a bitonic pass over 128 or
 * 256 entries in LDS. Entries m.. hold ~0 on entry. */
__device__ __forceinline__ void wave_sort_new(uint64_t *a, int m) {
    const int l = threadIdx.x;
    if (m <= 64) {
        uint64_t x = a[l];
        x = wave_sort64(x);
        a[l] = x;
        __syncthreads();
        return;
    }
    const int n = m <= 128 ? 128 : 256;
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int h = 0; h < kNS; h++) {
                const int i = l + 64 * h;
                const int ixj = i ^ j;
                if (i < n && ixj > i) {
                    const bool up = (i & k) == 0;
                    const uint64_t x = a[i], y = a[ixj];
                    if ((x > y) == up) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

/* number of entries of sorted a[0..n) strictly less than key */
__device__ __forceinline__ int lower_rank(const uint64_t *a, int n, uint64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/*
 * Merge: sorted A[skip .. n) with the m unsorted keys nw[0..m) (m <= kMaxNew,
 * slots m..kMaxNew-1 must hold ~0) into A[0 .. min(cap, n-skip+m)), ascending.
 * Keys are unique. Returns the new count. A may hold up to cap (<= 1024)
 * entries; every lane reads its elements before any write.
 */
__device__ int wave_merge(uint64_t *A, int n, int skip, uint64_t *nw, int m,
                          int cap) {
    const int l = threadIdx.x;
    if (m > 0) wave_sort_new(nw, m);
    const int nold = n - skip;
    uint64_t ov[16];
    int op[16];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int i = l + 64 * k;
        op[k] = -1;
        if (i < nold) {
            ov[k] = A[skip + i];
            op[k] = i + lower_rank(nw, m, ov[k]);
        }
    }
    uint64_t nv[kNS];
    int np[kNS];
#pragma unroll
    for (int k = 0; k < kNS; k++) {
        const int j = l + 64 * k;
        np[k] = -1;
        if (j < m) {
            nv[k] = nw[j];
            np[k] = j + lower_rank(A + skip, nold, nv[k]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++)
        if (op[k] >= 0 && op[k] < cap) A[op[k]] = ov[k];
#pragma unroll
    for (int k = 0; k < kNS; k++)
        if (np[k] >= 0 && np[k] < cap) A[np[k]] = nv[k];
    __syncthreads();
    return min(cap, nold + m);
}

/*
 * Exact CPU-order distances for up to NPG nodes per 16-lane group (4 NPG
 * nodes per call): all row loads are issued before any arithmetic so one
 * call costs one memory round trip. The summation order is the one of
 * cpu_order_l2 (engine.cpp:43-63).
 */
template <int NPG>
__device__ __forceinline__ void dist_nodes(const GraphArgs &A, const float *qsh,
                                           const uint32_t *nodes, int nn, int base,
                                           float *outd) {
    const int l = threadIdx.x, g = l >> 4, l16 = l & 15;
    const int D = A.dim;
    float acc0[NPG], acc1[NPG];
    const float *rp[NPG];
    bool ok[NPG];
#pragma unroll
    for (int k = 0; k < NPG; k++) {
        const int j = base + g + 4 * k;
        ok[k] = j < nn;
        rp[k] = A.vecs + (uint64_t)(ok[k] ? nodes[j] : 0) * D;
        acc0[k] = 0.f;
        acc1[k] = 0.f;
    }
    int i = 0;
    for (; i + 128 <= D; i += 128) {
        float x[NPG][8];
#pragma unroll
        for (int k = 0; k < NPG; k++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                x[k][2 * b] = rp[k][i + 32 * b + l16];
                x[k][2 * b + 1] = rp[k][i + 32 * b + 16 + l16];
            }
#pragma unroll
        for (int k = 0; k < NPG; k++)
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const float d0 = __fsub_rn(x[k][2 * b], qsh[i + 32 * b + l16]);
                const float d1 = __fsub_rn(x[k][2 * b + 1], qsh[i + 32 * b + 16 + l16]);
                acc0[k] = __fmaf_rn(d0, d0, acc0[k]);
                acc1[k] = __fmaf_rn(d1, d1, acc1[k]);
            }
    }
    for (; i + 32 <= D; i += 32) {
#pragma unroll
        for (int k = 0; k < NPG; k++) {
            const float d0 = __fsub_rn(rp[k][i + l16], qsh[i + l16]);
            const float d1 = __fsub_rn(rp[k][i + 16 + l16], qsh[i + 16 + l16]);
            acc0[k] = __fmaf_rn(d0, d0, acc0[k]);
            acc1[k] = __fmaf_rn(d1, d1, acc1[k]);
        }
    }
    for (; i + 16 <= D; i += 16) {
#pragma unroll
        for (int k = 0; k < NPG; k++) {
            const float d = __fsub_rn(rp[k][i + l16], qsh[i + l16]);
            acc0[k] = __fmaf_rn(d, d, acc0[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < NPG; k++) {
        float t = __fadd_rn(acc0[k], acc1[k]);
        t = __fadd_rn(t, lane_xor(t, 8));
        t = __fadd_rn(t, lane_xor(t, 4));
        t = __fadd_rn(t, lane_xor(t, 2));
        t = __fadd_rn(t, lane_xor(t, 1));
        t = tail_l2(rp[k], qsh, i, D, t);
        const int j = base + g + 4 * k;
        if (ok[k] && l16 == 0) outd[j] = t;
    }
}

__global__ __launch_bounds__(64) void k_graph(GraphArgs A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int q = blockIdx.x, l = threadIdx.x;
    if (A.route && A.route[q] == 0) return; /* routed to the brute path */
    const uint32_t efq = A.efq ? max(1u, min(A.efq[q], A.ef)) : A.ef;
    const int D = A.dim;
    float *qsh = (float *)smem;                                   /* [D] */
    uint64_t *F = (uint64_t *)(smem + (((size_t)D * 4 + 15) & ~(size_t)15)); /* [fcap] */
    uint64_t *R = F + A.fcap;                                     /* [ef] */
    uint64_t *nwF = R + A.ef;                                     /* [kMaxNew] */
    uint64_t *nwR = nwF + kMaxNew;                                /* [kMaxNew] */
    uint32_t *H = (uint32_t *)(nwR + kMaxNew);                    /* [2048] */
    uint32_t *nb = H + kHashSize;                                 /* [kMaxNew] new nodes */
    float *nd = (float *)(nb + kMaxNew);                          /* [kMaxNew] */
    uint32_t *pops = (uint32_t *)(nd + kMaxNew);                  /* [8] */

    const FvGpuQueryDesc qd = A.qd[q];
    const int prog = A.pofq[q];
    const uint32_t *bm = prog >= 0 ? A.bitmap + (uint64_t)prog * A.words : nullptr;
    uint32_t *vb = A.visit + (uint64_t)q * A.vwords;

    for (int j = l; j < D; j += 64) qsh[j] = A.qf[(uint64_t)q * D + j];
    for (int j = l; j < kHashSize; j += 64) H[j] = kNoNode;
    __syncthreads();

    uint32_t entry = qd.entry;
    if (entry >= A.count) entry = 0;
    /* entry point: distance by lanes 0..15 (others compute a masked copy) */
    nb[0] = entry;
    __syncthreads();
    dist_nodes<4>(A, qsh, nb, 1, 0, nd);
    if (l == 0) hash_insert(H, entry);
    __syncthreads();
    int nF = 1, nR = 0;
    const float ed = nd[0];
    const bool epass = bm == nullptr || ((bm[entry >> 5] >> (entry & 31)) & 1u);
    if (l == 0) {
        F[0] = make_key(ed, entry);
        if (epass) R[0] = make_key(ed, entry);
    }
    if (epass) nR = 1;
    __syncthreads();

    uint64_t visited = 1;
    int exh = 0;
    bool vb_ready = false;
    uint32_t iters = 0, ovf = 0;
    const int nbs = A.nbr_stride;

    for (;;) {
        if (nF == 0) {
            exh = 1;
            break;
        }
        const uint64_t c0 = F[0];
        const uint64_t worst = nR > 0 ? R[nR - 1] : ~0ull;
        const bool full = nR >= (int)efq;
        if (full && key_dist(c0) > key_dist(worst)) break;
        if (visited >= A.max_visited) {
            exh = 1;
            break;
        }
        if (iters > (uint32_t)A.count + 8) { /* cannot happen; bounds the loop */
            exh = 1;
            break;
        }
        iters++;
        /* pop up to W heads that pass the stop test */
        int w = 1;
        while (w < A.wmax && w < nF && (!full || key_dist(F[w]) <= key_dist(worst))) w++;
        if (l < w) pops[l] = key_row(F[l]);
        __syncthreads();

        /* gather neighbours of the popped nodes: w * nbs <= kMaxNew slots */
        const int nslot = w * nbs;
        int isnew[kNS];
        uint32_t cand[kNS];
#pragma unroll
        for (int h = 0; h < kNS; h++) {
            const int s = l + 64 * h;
            cand[h] = kNoNode;
            isnew[h] = 0;
            if (s < nslot) cand[h] = A.nbr[(uint64_t)pops[s / nbs] * nbs + (s % nbs)];
        }
        /* visited test-and-set */
        bool need_vb = false;
#pragma unroll
        for (int h = 0; h < kNS; h++) {
            if (cand[h] == kNoNode) continue;
            const int r = hash_insert(H, cand[h]);
            if (r == 1) isnew[h] = 1;
            else if (r == 2) {
                isnew[h] = 2;
                need_vb = true;
            }
        }
        if (__any(need_vb)) {
            if (!vb_ready) {
                /* first overflow: zero this query's bitmap */
                u32x4_t z = {0u, 0u, 0u, 0u};
                const uint64_t n4 = A.words / 4;
                for (uint64_t i = l; i < n4; i += 64) ((u32x4_t *)vb)[i] = z;
                for (uint64_t i = n4 * 4 + l; i < A.words; i += 64) vb[i] = 0;
                __threadfence_block();
                __syncthreads();
                vb_ready = true;
            }
            /* all test-and-sets in flight together */
            uint32_t old[kNS];
#pragma unroll
            for (int h = 0; h < kNS; h++)
                if (isnew[h] == 2)
                    old[h] = atomicOr(&vb[cand[h] >> 5], 1u << (cand[h] & 31));
#pragma unroll
            for (int h = 0; h < kNS; h++)
                if (isnew[h] == 2) {
                    isnew[h] = (old[h] & (1u << (cand[h] & 31))) ? 0 : 1;
                    ovf++;
                }
        }
        /* compact new nodes */
        const uint64_t lt = (l == 0) ? 0ull : (~0ull >> (64 - l));
        int nn = 0;
#pragma unroll
        for (int h = 0; h < kNS; h++) {
            const uint64_t bb = __ballot(isnew[h] == 1);
            if (isnew[h] == 1) nb[nn + __popcll(bb & lt)] = cand[h];
            nn += __popcll(bb);
        }
        __syncthreads();
        visited += (uint64_t)nn;

        /* predicate words of the new nodes, requested with their rows (the
         * admission below reads them after the distances) */
        uint32_t pw[kNS];
#pragma unroll
        for (int h = 0; h < kNS; h++) {
            const int j = l + 64 * h;
            pw[h] = ~0u;
            if (bm != nullptr && j < nn) pw[h] = bm[nb[j] >> 5];
        }
        /* distances: 32 nodes per round trip (16 for the last few) */
        int base = 0;
        for (; base + 16 < nn; base += 32) dist_nodes<8>(A, qsh, nb, nn, base, nd);
        if (base < nn) dist_nodes<4>(A, qsh, nb, nn, base, nd);
        __syncthreads();

        /* admission against the worst result at iteration start */
        const float wd = key_dist(worst);
        int mF = 0, mR = 0;
        {
            bool adm[kNS], pas[kNS];
            uint64_t key[kNS];
#pragma unroll
            for (int h = 0; h < kNS; h++) {
                const int j = l + 64 * h;
                adm[h] = false;
                pas[h] = false;
                key[h] = ~0ull;
                if (j < nn) {
                    const float d = nd[j];
                    const uint32_t node = nb[j];
                    key[h] = make_key(d, node);
                    adm[h] = !full || d < wd;
                    if (adm[h]) pas[h] = ((pw[h] >> (node & 31)) & 1u) != 0u;
                }
            }
            /* fill with ~0 first, then scatter */
#pragma unroll
            for (int h = 0; h < kNS; h++) {
                nwF[l + 64 * h] = ~0ull;
                nwR[l + 64 * h] = ~0ull;
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < kNS; h++) {
                const uint64_t ab = __ballot(adm[h]), pb = __ballot(pas[h]);
                if (adm[h]) nwF[mF + __popcll(ab & lt)] = key[h];
                if (pas[h]) nwR[mR + __popcll(pb & lt)] = key[h];
                mF += __popcll(ab);
                mR += __popcll(pb);
            }
            __syncthreads();
        }
        nF = wave_merge(F, nF, w, nwF, mF, A.fcap);
        if (mR > 0) nR = wave_merge(R, nR, 0, nwR, mR, (int)efq);
    }

    const uint32_t ostride = A.out_k ? A.out_k : A.ef;
    const int nout = min(nR, (int)ostride);
    for (int i = l; i < nout; i += 64) {
        FvHit h;
        h.node = key_row(R[i]);
        h.dist = key_dist(R[i]);
        A.out[(uint64_t)q * ostride + i] = h;
    }
    if (l == 0) {
        A.out_cnt[q] = (uint32_t)nout;
        A.out_exh[q] = (uint8_t)exh;
        A.status[q] = 0;
        if (A.dbg) {
            A.dbg[(uint64_t)q * 4 + 0] = iters;
            A.dbg[(uint64_t)q * 4 + 1] = (uint32_t)min<uint64_t>(visited, 0xFFFFFFFFull);
            A.dbg[(uint64_t)q * 4 + 2] = (uint32_t)nR;
            A.dbg[(uint64_t)q * 4 + 3] = ovf;
        }
    }
}

size_t graph_shmem(int dim, int fcap, uint32_t ef) {
    return (((size_t)dim * 4 + 15) & ~(size_t)15) + (size_t)fcap * 8 + (size_t)ef * 8 +
           2 * kMaxNew * 8 + kHashSize * 4 + kMaxNew * 4 + kMaxNew * 4 + 12 * 4;
}

int run_graph_ex(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef, uint64_t max_visited, int wmax,
                 int fcap, uint32_t *dbg, const GraphExtra &x, char *errbuf, size_t errlen) {
    if (g->d_nbr == nullptr) PV_FAIL("gpu: graph not uploaded");
    if (ef < 1 || ef > FV_GPU_MAX_EF) PV_FAIL("gpu: bad graph batch (nq=%d ef=%u)", b->nq, ef);
    if (require_bitmaps(b, errbuf, errlen) != 0) return -1;
    if (b->max_entry >= g->count)
        PV_FAIL("gpu: entry point %llu out of range", (unsigned long long)b->max_entry);
    if (wmax < 1) wmax = 1;
    while (wmax > 1 && wmax * g->nbr_stride > kMaxNew) wmax--;
    if (g->nbr_stride > kMaxNew) PV_FAIL("gpu: graph degree %d too large", g->nbr_stride);
    if (fcap < 64) fcap = 64;
    if (fcap > 1024) fcap = 1024;
    const size_t shm = graph_shmem(g->dim, fcap, ef);
    if (shm > 160 * 1024) PV_FAIL("gpu: graph search LDS %zu bytes exceeds the CU", shm);
    const uint64_t words = g->count_pad / 32;
    const uint64_t vwords = (words + 3) & ~3ull;
    PV_TRY(b->d_visit.ensure((size_t)b->nq * vwords * 4));
    const uint32_t ostride = x.out_k ? x.out_k : ef;
    b->out_stride = ostride;
    b->last_route = 3;
    PV_TRY(b->d_out.ensure((size_t)b->nq * ostride * sizeof(FvHit)));
    GraphArgs A;
    A.vecs = g->d_vecs;
    A.nbr = g->d_nbr;
    A.nbr_stride = g->nbr_stride;
    A.count = g->count;
    A.dim = g->dim;
    A.qf = b->d_q;
    A.qd = b->d_qd;
    A.pofq = b->d_pofq;
    A.bitmap = b->d_bitmap.as<uint32_t>();
    A.words = words;
    A.vwords = vwords;
    A.ef = ef;
    A.max_visited = max_visited;
    A.wmax = wmax;
    A.fcap = fcap;
    A.visit = b->d_visit.as<uint32_t>();
    A.out = b->d_out.as<FvHit>();
    A.out_cnt = b->d_cnt.as<uint32_t>();
    A.out_exh = b->d_flag.as<uint8_t>();
    A.status = qaux_of(b).status;
    A.route = x.route;
    A.efq = x.efq;
    A.out_k = x.out_k;
    A.dbg = dbg;
    if (shm > 64 * 1024)
        PV_TRY(hipFuncSetAttribute((const void *)k_graph,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    g->ev_stream = b->stream;
    ev_record(g, 0);
    if (!x.bitmap_ready) PV_TRY(launch_bitmap(g, b, b->stream));
    hipLaunchKernelGGL(k_graph, dim3(b->nq), dim3(64), shm, b->stream, A);
    PV_TRY(hipGetLastError());
    ev_record(g, 5);
    memset(&g->last, 0, sizeof(g->last));
    g->last.route = 3;
    return 0;
}

int run_graph_opts(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef, uint64_t max_visited,
                   int wmax, int fcap, uint32_t *dbg, char *errbuf, size_t errlen) {
    return run_graph_ex(g, b, ef, max_visited, wmax, fcap, dbg, GraphExtra(), errbuf, errlen);
}

int run_graph(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef, uint64_t max_visited,
              char *errbuf, size_t errlen) {
    return run_graph_opts(g, b, ef, max_visited, 8, 512, nullptr, errbuf, errlen);
}

} // namespace pv

using namespace pv;

static int graph_sync(FvGpuIndex *g, int nq, const float *queries,
                      const FvGpuQueryDesc *qd, const FvGpuPredNode *nodes,
                      int nnodes, const int32_t *children, int nchildren,
                      const double *invals, int ninvals, uint32_t ef,
                      uint64_t max_visited, int wmax, int fcap, FvHit *out,
                      uint32_t *out_cnt, uint8_t *exhausted, uint32_t *dbg_host,
                      char *errbuf, size_t errlen) {
    if (g == nullptr) PV_FAIL("gpu: NULL index");
    if (g->d_nbr == nullptr) PV_FAIL("gpu: graph not uploaded");
    if (nq < 1 || nq > FV_GPU_MAX_BATCH || ef < 1 || ef > FV_GPU_MAX_EF)
        PV_FAIL("gpu: bad graph batch (nq=%d ef=%u)", nq, ef);
    if (out == nullptr || out_cnt == nullptr || exhausted == nullptr)
        PV_FAIL("gpu: NULL output arrays");
    for (int q = 0; q < nq; q++)
        if ((uint64_t)qd[q].entry >= g->count)
            PV_FAIL("gpu: query %d: entry point %u out of range", q, qd[q].entry);
    PV_TRY(hipSetDevice(g->device));
    if (g->sync_batch == nullptr) g->sync_batch = new FvGpuBatch();
    FvGpuBatch *b = g->sync_batch;
    /* graph calls carry no spans (gpu_engine.h:104-109) */
    std::vector<FvGpuQueryDesc> qd2(qd, qd + nq);
    for (auto &d : qd2) {
        d.nspans = 0;
        d.span0 = 0;
        d.soff0 = 0;
    }
    if (stage_batch(g, b, nq, queries, qd2.data(), nullptr, nodes, nnodes, children,
                    nchildren, invals, ninvals, nullptr, 0, nullptr, 0, 1, errbuf,
                    errlen) != 0)
        return -1;
    uint32_t *d_dbg = nullptr;
    if (dbg_host) PV_TRY(hipMalloc(&d_dbg, (size_t)nq * 16));
    int rc = run_graph_opts(g, b, ef, max_visited, wmax, fcap, d_dbg, errbuf, errlen);
    if (rc == 0) {
        b->out_stride = ef;
        std::vector<FvHit> tmp((size_t)nq * ef);
        hipError_t e = hipMemcpyAsync(tmp.data(), b->d_out.p, (size_t)nq * ef * sizeof(FvHit),
                                      hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(out_cnt, b->d_cnt.p, (size_t)nq * 4, hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(exhausted, b->d_flag.p, (size_t)nq, hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess && dbg_host)
            e = hipMemcpyAsync(dbg_host, d_dbg, (size_t)nq * 16, hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
            snprintf(errbuf, errlen, "gpu: graph fetch: %s", hipGetErrorString(e));
            rc = -1;
        } else {
            for (int q = 0; q < nq; q++) {
                const uint32_t c = std::min(out_cnt[q], ef);
                memcpy(out + (uint64_t)q * ef, tmp.data() + (uint64_t)q * ef, c * sizeof(FvHit));
            }
            if (g->timing) {
                g->last.t_graph_us = ev_elapsed_us(g, 0, 5);
                g->last.t_total_us = g->last.t_graph_us;
            }
        }
    }
    if (d_dbg) hipFree(d_dbg);
    return rc;
}

extern "C" int fv_gpu_graph_batch(FvGpuIndex *g, int nq, const float *queries,
                                  const FvGpuQueryDesc *qd, const FvGpuPredNode *nodes,
                                  int nnodes, const int32_t *children, int nchildren,
                                  const double *invals, int ninvals, uint32_t ef,
                                  uint64_t max_visited, FvHit *out, uint32_t *out_cnt,
                                  uint8_t *exhausted, char *errbuf, size_t errlen) {
    try {
        return graph_sync(g, nq, queries, qd, nodes, nnodes, children, nchildren, invals,
                          ninvals, ef, max_visited, 8, 512, out, out_cnt, exhausted,
                          nullptr, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

extern "C" int fv_gpu_graph_batch_ex(FvGpuIndex *g, int nq, const float *queries,
                                     const FvGpuQueryDesc *qd, const FvGpuPredNode *nodes,
                                     int nnodes, const int32_t *children, int nchildren,
                                     const double *invals, int ninvals, uint32_t ef,
                                     uint64_t max_visited, int wmax, int fcap,
                                     FvHit *out, uint32_t *out_cnt, uint8_t *exhausted,
                                     uint32_t *dbg, char *errbuf, size_t errlen) {
    try {
        return graph_sync(g, nq, queries, qd, nodes, nnodes, children, nchildren, invals,
                          ninvals, ef, max_visited, wmax, fcap, out, out_cnt, exhausted,
                          dbg, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

extern "C" int fv_gpu_batch_run_graph(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef,
                                      uint64_t max_visited, char *errbuf, size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g) PV_FAIL("gpu: bad batch");
    try {
        return run_graph(g, b, ef, max_visited, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}
