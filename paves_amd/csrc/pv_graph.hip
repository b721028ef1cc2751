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
 * CDNA4 design: one 256-thread workgroup (four waves) per query, so a
 * query's per-iteration work is spread over 256 lanes and a 64-query batch
 * still puts 256 waves on the chip.
 *   - frontier and results are sorted key arrays in LDS (key = dist bits
 *     << 32 | node, so key order = (dist, node) order, the CPU heaps' pair
 *     order). Each iteration's admitted keys are sorted per wave in
 *     registers (64-key bitonic on DPP / swizzle / permlane exchanges: the
 *     wave's lane holds its own node's key, so no compaction is needed) and
 *     merged by rank: an entry's new position is its own index plus the
 *     number of keys below it in every other sorted run (binary searches in
 *     LDS, all independent), so a merge costs two barriers;
 *   - the frontier keeps at most fcap entries (the reference GPU kernel's
 *     FV_FRONTIER_CAP, gpu_engine.cu:1069); entries pushed past it are
 *     dropped and counted (debug word 3), the only way the search can
 *     depart from the unbounded CPU cursor;
 *   - visited: a 4096-slot LDS open-addressing hash (linear probing,
 *     bounded probe length); a node whose probe window is full goes to a
 *     per-query global bitmap that the workgroup zero-fills the first time
 *     it needs it — exact membership, no pre-launch memset;
 *   - each iteration pops up to W frontier heads that pass the stop test;
 *     thread t reads neighbour slot t of the popped heads' padded lists (one
 *     128-byte line per head for M0 = 32), test-and-sets it, and the new
 *     nodes' rows are read with exact FP32 distances in the CPU cursor's
 *     summation order, 16 lanes per node, all row loads of a pass issued
 *     before the arithmetic;
 *   - the predicate is a bit test in the batch's predicate bitmap.
 * With W = 1 and a frontier that never drops an entry, the search visits
 * the nodes the CPU cursor visits and returns the same list.
 */
#include <math.h>
#include <string.h>

#include <algorithm>

#include "pv_device.h"

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

namespace pv {

constexpr int kGW = 4;                    /* waves per query workgroup */
constexpr int kGT = kGW * 64;             /* threads per query workgroup */
constexpr int kHashBits = 12;
constexpr int kHashSize = 1 << kHashBits; /* 4096 slots */
constexpr int kProbe = 16;
constexpr int kMaxNew = kGT;              /* W * nbr_stride <= 256 (W <= 8 at degree 32) */

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
    uint64_t *spill;     /* [nq][scap + fcap] frontier spill stores */
    uint32_t scap;       /* spilled entries a store accepts */
    uint32_t *dbg;       /* [nq][4] heads expanded, visited, results, frontier entries lost */
    const uint32_t *order; /* optional [nq]: query of workgroup b (heaviest search first) */
};

__device__ __forceinline__ uint32_t hash_slot(uint32_t node) {
    return (node * 2654435761u) >> (32 - kHashBits);
}

/* Test-and-set in the LDS hash; returns 1 if newly inserted, 0 if already
 * present, 2 if the probe window is full (caller uses the global bitmap).
 * Slots are never cleared, so a node whose window is full stays out of the
 * hash for the whole search and is always found in the bitmap. */
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

/* number of keys below `key` in a sorted run of 64 padded with ~0 (no
 * key equals ~0), by binary lifting: six fixed steps, no data-dependent
 * loop, so the many independent searches of a merge interleave their LDS
 * reads */
__device__ __forceinline__ int run_rank(const uint64_t *run, uint64_t key) {
    int pos = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1)
        if (run[pos + step - 1] < key) pos += step;
    /* the steps reach at most 63: a 64th key below `key` is counted here
     * (for pos < 63, run[pos] >= key already) */
    return pos + (run[pos] < key ? 1 : 0);
}

/* number of entries of sorted a[0..n) below key, n <= 1024: binary
 * lifting in ten fixed steps */
__device__ __forceinline__ int lift_rank(const uint64_t *a, int n, uint64_t key) {
    int pos = 0;
#pragma unroll
    for (int step = 512; step > 0; step >>= 1)
        if (pos + step <= n && a[pos + step - 1] < key) pos += step;
    /* the steps reach at most 1023: with n = 1024 the last entry is
     * counted here (otherwise a[pos] >= key or pos == n) */
    return pos + ((pos < n && a[pos] < key) ? 1 : 0);
}

/* per-workgroup scalars shared through LDS */
struct GraphShared {
    int wcnt[kGW];       /* new nodes per wave (compaction) */
    int rcF[kGW];        /* admitted keys per wave */
    int rcR[kGW];        /* passing admitted keys per wave */
    uint64_t smin;       /* smallest key a merge spilled */
    uint64_t wmin[kGW];  /* per-wave minima (refill) */
    uint32_t wsum[kGW];  /* per-wave sums (refill) */
    uint32_t nsel, nrest;
    uint32_t hist[256], tmp[8], bc[8];  /* radix select (refill) */
};

/*
 * Merge the sorted entries A[skip .. n) with kGW sorted runs of new keys
 * (run r: runs[64 r .. 64 r + rc[r]), wave r's keys, ~0-padded). Keys are
 * unique across A and the runs. Keys above lim are dead (they can never be
 * popped) and are discarded. Of the `live` remaining entries, in ascending
 * order, the first cap go to A[0 ..) and the rest to spill[nspill ..)
 * (those beyond scap are lost); S.smin receives the first spilled key. An
 * entry's position is its index in its own sequence plus its rank in each
 * of the others. Called by all kGT threads with identical arguments; every
 * thread reads its entries before any write. Returns live.
 */
__device__ int block_merge(uint64_t *A, int n, int skip, const uint64_t *runs, const int (&rc)[kGW],
                           int cap, uint64_t lim, uint64_t *spill, uint32_t nspill, uint32_t scap,
                           GraphShared &S) {
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int nold = n - skip;
    /* live entries of each sequence (uniform) */
    const bool all_live = lim == ~0ull;
    const int lold = all_live ? nold : lift_rank(A + skip, nold, lim + 1);
    int live = lold;
    int lrc[kGW];
#pragma unroll
    for (int r = 0; r < kGW; r++) {
        lrc[r] = (all_live || rc[r] == 0) ? rc[r] : run_rank(runs + 64 * r, lim + 1);
        live += lrc[r];
    }
    constexpr int kOld = 1024 / kGT;   /* A holds at most 1024 entries */
    uint64_t ov[kOld];
    int op[kOld];
#pragma unroll
    for (int k = 0; k < kOld; k++) {
        const int i = t + kGT * k;
        op[k] = -1;
        if (i < lold) {
            ov[k] = A[skip + i];
            int p = i;
#pragma unroll
            for (int r = 0; r < kGW; r++)
                if (lrc[r] > 0) p += run_rank(runs + 64 * r, ov[k]);
            op[k] = p;
        }
    }
    uint64_t nv = ~0ull;
    int np = -1;
    if (l < lrc[w]) {
        nv = runs[64 * w + l];
        int p = l + lift_rank(A + skip, lold, nv);
#pragma unroll
        for (int r = 0; r < kGW; r++)
            if (r != w && lrc[r] > 0) p += run_rank(runs + 64 * r, nv);
        np = p;
    }
    __syncthreads();
    auto put = [&](int p, uint64_t key) {
        if (p < cap) {
            A[p] = key;
        } else if (spill != nullptr) {
            const uint32_t s = nspill + (uint32_t)(p - cap);
            if (s < scap) spill[s] = key;
            if (p == cap) S.smin = key;
        }
    };
#pragma unroll
    for (int k = 0; k < kOld; k++)
        if (op[k] >= 0) put(op[k], ov[k]);
    if (np >= 0) put(np, nv);
    __threadfence_block();
    __syncthreads();
    return live;
}

/* sum / minimum over the workgroup (uniform result) */
__device__ __forceinline__ uint32_t block_sum(uint32_t v, GraphShared &S) {
#pragma unroll
    for (int o = 32; o; o >>= 1) v += lane_xor(v, o);
    if ((threadIdx.x & 63) == 0) S.wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < kGW; r++) s += S.wsum[r];
    __syncthreads();
    return s;
}
__device__ __forceinline__ uint64_t block_min(uint64_t v, GraphShared &S) {
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        const uint64_t u = lane_xor(v, o);
        v = u < v ? u : v;
    }
    if ((threadIdx.x & 63) == 0) S.wmin[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t m = ~0ull;
#pragma unroll
    for (int r = 0; r < kGW; r++) m = S.wmin[r] < m ? S.wmin[r] : m;
    __syncthreads();
    return m;
}

/*
 * The frontier's spill store O holds live entries that did not fit in the
 * LDS frontier; omin is its smallest key. When omin beats the LDS head (or
 * the LDS frontier is empty), the LDS entries join the store and the
 * smallest min(live, cap / 2) entries of the store come back to LDS,
 * sorted: the k-th smallest by radix select over the store, a partition
 * pass (in chunks read before they are overwritten: the store is compacted
 * in place), dead entries (above lim) dropped, then a bitonic sort in LDS.
 * cap is a power of two; the store has room for cap entries beyond the
 * spills it accepts. Rare: only a frontier that outgrows LDS (a result
 * list that fills slowly under a selective filter) spills.
 */
__device__ void frontier_refill(uint64_t *F, int &nF, uint64_t *O, uint32_t &nO, uint64_t &omin,
                                uint64_t lim, int cap, bool &pruned, GraphShared &S) {
    const int t = threadIdx.x;
    for (int i = t; i < nF; i += kGT) O[nO + i] = F[i];
    const uint32_t n = nO + (uint32_t)nF;
    __threadfence_block();
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t i = t; i < n; i += kGT) c += O[i] <= lim ? 1u : 0u;
    const uint32_t live = block_sum(c, S);
    if (live < n) pruned = true;
    if (live == 0) {
        nF = 0;
        nO = 0;
        omin = ~0ull;
        return;
    }
    const uint32_t k = min(live, (uint32_t)cap / 2);
    const uint64_t kth = block_kth([&](uint32_t i) { return O[i]; }, n, k, S.hist, S.tmp, S.bc);
    if (t == 0) {
        S.nsel = 0;
        S.nrest = 0;
    }
    __syncthreads();
    uint64_t rmin = ~0ull;
    for (uint32_t b = 0; b < n; b += 4 * kGT) {
        uint64_t v[4];
        bool ok[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t i = b + t + kGT * j;
            ok[j] = i < n;
            v[j] = ok[j] ? O[i] : ~0ull;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (!ok[j]) continue;
            if (v[j] <= kth) {
                F[atomicAdd(&S.nsel, 1u)] = v[j];
            } else if (v[j] <= lim) {
                O[atomicAdd(&S.nrest, 1u)] = v[j];
                rmin = v[j] < rmin ? v[j] : rmin;
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    nF = (int)S.nsel;
    nO = S.nrest;
    omin = block_min(rmin, S);
    uint32_t np2 = 2;
    while (np2 < (uint32_t)nF) np2 <<= 1;
    for (uint32_t i = (uint32_t)nF + t; i < np2; i += kGT) F[i] = ~0ull;
    __syncthreads();
    block_bitonic(F, np2);
}

/*
 * Exact CPU-order distances of nodes[base + g + 16 k] (k < NPG) by the
 * sixteen 16-lane groups g of the workgroup: all row loads are issued
 * before any arithmetic, so one call costs one memory round trip. The
 * summation order is the one of cpu_order_l2 (engine.cpp:43-63).
 */
template <int NPG>
__device__ __forceinline__ void dist_nodes(const GraphArgs &A, const float *qsh,
                                           const uint32_t *nodes, int nn, int base,
                                           float *outd) {
    const int t = threadIdx.x, g = t >> 4, l16 = t & 15;
    const int D = A.dim;
    float acc0[NPG], acc1[NPG];
    const float *rp[NPG];
    bool ok[NPG];
#pragma unroll
    for (int k = 0; k < NPG; k++) {
        const int j = base + g + 16 * k;
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
        float s = __fadd_rn(acc0[k], acc1[k]);
        s = __fadd_rn(s, lane_xor(s, 8));
        s = __fadd_rn(s, lane_xor(s, 4));
        s = __fadd_rn(s, lane_xor(s, 2));
        s = __fadd_rn(s, lane_xor(s, 1));
        s = tail_l2(rp[k], qsh, i, D, s);
        const int j = base + g + 16 * k;
        if (ok[k] && l16 == 0) outd[j] = s;
    }
}

/*
 * 128-dim rows (the common case): the same arithmetic as dist_nodes<4>, with
 * the query slice in registers and two row buffers, so the loads of the
 * next 64 nodes are in flight while the current 64 are reduced.
 */
__device__ __forceinline__ void dist_all_128(const GraphArgs &A, const float *qsh,
                                             const uint32_t *nodes, int nn, float *outd) {
    const int t = threadIdx.x, g = t >> 4, l16 = t & 15;
    float qv[8];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        qv[2 * b] = qsh[32 * b + l16];
        qv[2 * b + 1] = qsh[32 * b + 16 + l16];
    }
    float xa[4][8], xb[4][8];
    auto load = [&](float (&x)[4][8], int base) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = base + g + 16 * k;
            const float *rp = A.vecs + (uint64_t)(j < nn ? nodes[j] : 0) * 128;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                x[k][2 * b] = rp[32 * b + l16];
                x[k][2 * b + 1] = rp[32 * b + 16 + l16];
            }
        }
    };
    auto reduce = [&](const float (&x)[4][8], int base) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float a0 = 0.f, a1 = 0.f;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const float d0 = __fsub_rn(x[k][2 * b], qv[2 * b]);
                const float d1 = __fsub_rn(x[k][2 * b + 1], qv[2 * b + 1]);
                a0 = __fmaf_rn(d0, d0, a0);
                a1 = __fmaf_rn(d1, d1, a1);
            }
            float s = __fadd_rn(a0, a1);
            s = __fadd_rn(s, lane_xor(s, 8));
            s = __fadd_rn(s, lane_xor(s, 4));
            s = __fadd_rn(s, lane_xor(s, 2));
            s = __fadd_rn(s, lane_xor(s, 1));
            const int j = base + g + 16 * k;
            if (j < nn && l16 == 0) outd[j] = s;
        }
    };
    int base = 0;
    load(xa, 0);
    for (;;) {
        if (base + 64 < nn) load(xb, base + 64);
        reduce(xa, base);
        base += 64;
        if (base >= nn) break;
        if (base + 64 < nn) load(xa, base + 64);
        reduce(xb, base);
        base += 64;
        if (base >= nn) break;
    }
}

/* distances of nodes[0 .. nn): 64 per round trip */
__device__ __forceinline__ void dist_all(const GraphArgs &A, const float *qsh,
                                         const uint32_t *nodes, int nn, float *outd) {
    if (A.dim == 128) {
        if (nn > 0) dist_all_128(A, qsh, nodes, nn, outd);
        return;
    }
    for (int base = 0; base < nn; base += 64) dist_nodes<4>(A, qsh, nodes, nn, base, outd);
}


__global__ __launch_bounds__(kGT) __attribute__((amdgpu_waves_per_eu(3)))
void k_graph(GraphArgs A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int q = A.order ? (int)A.order[blockIdx.x] : (int)blockIdx.x;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    if (A.route && A.route[q] == 0) return; /* routed to the brute path */
    const uint32_t efq = A.efq ? max(1u, min(A.efq[q], A.ef)) : A.ef;
    const int D = A.dim;
    float *qsh = (float *)smem;                                   /* [D] */
    uint64_t *F = (uint64_t *)(smem + (((size_t)D * 4 + 15) & ~(size_t)15)); /* [fcap] */
    uint64_t *R = F + A.fcap;                                     /* [ef] */
    uint64_t *runF = R + A.ef;                                    /* [kGW][64] */
    uint64_t *runR = runF + kGT;                                  /* [kGW][64] */
    uint32_t *H = (uint32_t *)(runR + kGT);                       /* [kHashSize] */
    uint32_t *nb = H + kHashSize;                                 /* [kMaxNew] new nodes */
    float *nd = (float *)(nb + kMaxNew);                          /* [kMaxNew] */
    GraphShared &S = *(GraphShared *)(nd + kMaxNew);

    const FvGpuQueryDesc qd = A.qd[q];
    const int prog = A.pofq[q];
    const uint32_t *bm = prog >= 0 ? A.bitmap + (uint64_t)prog * A.words : nullptr;
    uint32_t *vb = A.visit + (uint64_t)q * A.vwords;
    uint64_t *O = A.spill + (uint64_t)q * (A.scap + (uint32_t)A.fcap); /* frontier spill store */

    for (int j = t; j < D; j += kGT) qsh[j] = A.qf[(uint64_t)q * D + j];
    for (int j = t; j < kHashSize; j += kGT) H[j] = kNoNode;
    uint32_t entry = qd.entry;
    if (entry >= A.count) entry = 0;
    if (t == 0) nb[0] = entry;
    __syncthreads();
    dist_nodes<1>(A, qsh, nb, 1, 0, nd);
    if (t == 0) hash_insert(H, entry);
    __syncthreads();
    int nF = 1, nR = 0;
    const float ed = nd[0];
    const bool epass = bm == nullptr || ((bm[entry >> 5] >> (entry & 31)) & 1u);
    if (t == 0) {
        F[0] = make_key(ed, entry);
        if (epass) R[0] = make_key(ed, entry);
    }
    if (epass) nR = 1;
    __syncthreads();

    uint64_t visited = 1;
    int exh = 0;
    bool vb_ready = false;
    bool pruned = false;        /* dead frontier entries were discarded */
    uint32_t nO = 0, lost = 0;  /* spilled entries; live entries beyond the store */
    uint64_t omin = ~0ull;      /* smallest spilled key */
    uint32_t iters = 0, heads = 0;
    const int nbs = A.nbr_stride;

    for (;;) {
        const bool full = nR >= (int)efq;
        const uint64_t worst = nR > 0 ? R[nR - 1] : ~0ull;
        /* frontier entries above lim can never be popped: when the list is
         * full the CPU cursor stops at them (engine.cpp:1666) */
        const uint64_t lim = full ? make_key(key_dist(worst), kNoNode) : ~0ull;
        if (nO > 0 && (nF == 0 || omin < F[0]))
            frontier_refill(F, nF, O, nO, omin, lim, A.fcap, pruned, S);
        if (nF == 0) {
            /* frontier drained; with discarded dead entries the CPU cursor's
             * heap still holds one, and it stops on the full-list test */
            exh = pruned ? 0 : 1;
            break;
        }
        const uint64_t c0 = F[0];
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
        /* pop up to W heads that pass the stop test and precede every
         * spilled entry (uniform) */
        int nh;
        {
            /* lane j tests head j; the heads are the leading run of passes */
            bool ok = l == 0;
            if (l > 0 && l < A.wmax && l < nF) {
                const uint64_t fj = F[l];
                ok = fj < omin && (!full || key_dist(fj) <= key_dist(worst));
            }
            const uint64_t okb = __ballot(ok);
            nh = __ffsll((unsigned long long)~okb) - 1;   /* okb bit 0 is set */
        }
        heads += (uint32_t)nh;

        /* neighbour slot t of the popped heads, then its visited test-and-set */
        uint32_t cand = kNoNode;
        if (t < nh * nbs) {
            const uint32_t row = key_row(F[t / nbs]);
            if (row < A.count) cand = A.nbr[(uint64_t)row * nbs + (t % nbs)];
        }
        int isnew = 0;
        if (cand != kNoNode) isnew = hash_insert(H, cand);
        if (__syncthreads_or(isnew == 2)) {
            if (!vb_ready) {
                /* first overflow of this search: zero this query's bitmap */
                u32x4_t z = {0u, 0u, 0u, 0u};
                const uint64_t n4 = A.words / 4;
                for (uint64_t i = t; i < n4; i += kGT) ((u32x4_t *)vb)[i] = z;
                for (uint64_t i = n4 * 4 + t; i < A.words; i += kGT) vb[i] = 0;
                __threadfence();
                __syncthreads();
                vb_ready = true;
            }
            if (isnew == 2) {
                const uint32_t old = atomicOr(&vb[cand >> 5], 1u << (cand & 31));
                isnew = (old & (1u << (cand & 31))) ? 0 : 1;
            }
        }
        /* compact the new nodes into nb[] (wave ballots + wave offsets) */
        const uint64_t lt = (l == 0) ? 0ull : (~0ull >> (64 - l));
        const uint64_t bb = __ballot(isnew == 1);
        if (l == 0) S.wcnt[w] = __popcll(bb);
        __syncthreads();
        int woff = 0, nn = 0;
#pragma unroll
        for (int r = 0; r < kGW; r++) {
            woff += r < w ? S.wcnt[r] : 0;
            nn += S.wcnt[r];
        }
        if (isnew == 1) nb[woff + __popcll(bb & lt)] = cand;
        __syncthreads();
        visited += (uint64_t)nn;

        /* predicate word of thread t's new node, requested with the rows */
        const bool mine = t < nn;
        const uint32_t node = mine ? nb[t] : 0u;
        const uint32_t pw = (mine && bm != nullptr) ? bm[node >> 5] : ~0u;
        dist_all(A, qsh, nb, nn, nd);
        __syncthreads();

        /* admission against the worst result at iteration start; wave w's
         * admitted (and passing) keys sorted in registers form run w */
        const float wd = key_dist(worst);
        bool adm = false, pas = false;
        uint64_t key = ~0ull;
        if (mine) {
            const float d = nd[t];
            key = make_key(d, node);
            adm = !full || d < wd;
            pas = adm && (((pw >> (node & 31)) & 1u) != 0u);
        }
        const uint64_t sF = wave_sort64(adm ? key : ~0ull);
        const uint64_t sR = wave_sort64(pas ? key : ~0ull);
        runF[t] = sF;
        runR[t] = sR;
        const int cF = __popcll(__ballot(adm)), cR = __popcll(__ballot(pas));
        if (l == 0) {
            S.rcF[w] = cF;
            S.rcR[w] = cR;
        }
        if (t == 0) S.smin = ~0ull;
        __syncthreads();
        int rcF[kGW], rcR[kGW], mF = 0, mR = 0;
#pragma unroll
        for (int r = 0; r < kGW; r++) {
            rcF[r] = S.rcF[r];
            rcR[r] = S.rcR[r];
            mF += rcF[r];
            mR += rcR[r];
        }
        /* results first: the frontier merge discards what the new worst
         * result makes dead */
        if (mR > 0)
            nR = min((int)efq, block_merge(R, nR, 0, runR, rcR, (int)efq, ~0ull, nullptr, 0, 0, S));
        const bool full2 = nR >= (int)efq;
        const uint64_t lim2 = full2 ? make_key(key_dist(R[nR - 1]), kNoNode) : ~0ull;
        const int nold = nF - nh;
        const int live = block_merge(F, nF, nh, runF, rcF, A.fcap, lim2, O, nO, A.scap, S);
        if (live < nold + mF) pruned = true;
        nF = min(A.fcap, live);
        if (live > A.fcap) {
            /* the merge stored the first min(add, room) of them, the
             * smallest one first */
            const uint32_t add = (uint32_t)(live - A.fcap);
            const uint32_t room = A.scap - min(nO, A.scap);
            lost += add > room ? add - room : 0u;
            nO += min(add, room);
            if (room > 0) omin = S.smin < omin ? S.smin : omin;
        }
        if (full2 && nO > 0 && omin > lim2) {
            /* every spilled entry is dead */
            nO = 0;
            omin = ~0ull;
            pruned = true;
        }
    }

    const uint32_t ostride = A.out_k ? A.out_k : A.ef;
    const int nout = min(nR, (int)ostride);
    for (int i = t; i < nout; i += kGT) {
        FvHit h;
        h.node = key_row(R[i]);
        h.dist = key_dist(R[i]);
        A.out[(uint64_t)q * ostride + i] = h;
    }
    if (t == 0) {
        A.out_cnt[q] = (uint32_t)nout;
        A.out_exh[q] = (uint8_t)exh;
        A.status[q] = 0;
        if (A.dbg) {
            A.dbg[(uint64_t)q * 4 + 0] = heads;
            A.dbg[(uint64_t)q * 4 + 1] = (uint32_t)min<uint64_t>(visited, 0xFFFFFFFFull);
            A.dbg[(uint64_t)q * 4 + 2] = (uint32_t)nR;
            A.dbg[(uint64_t)q * 4 + 3] = lost;
        }
    }
}

/*
 * Launch order of a graph batch, heaviest search first (the longest-
 * processing-time rule of the reference worker, worker.c:206-232: one query
 * is one workgroup and workgroups start in index order, so descending
 * expected work makes the last wave of workgroups finish together instead
 * of leaving a long search to start late). Expected work ~ ef / selectivity
 * (worker.c:221: want * 1000 / sel_permille); the selectivity is measured:
 * passing rows of each program from its bitmap (k_prog_count).
 */
/* popcount of each program's bitmap row, one block per program (pv_learn.hip) */
__global__ void k_prog_count(const uint32_t *__restrict__ bitmap, uint64_t words,
                             uint32_t *__restrict__ cnt);

constexpr int kOrderMax = 4096;
__global__ __launch_bounds__(1024) void k_lpt_order(const int32_t *__restrict__ pofq,
                                                    const uint32_t *__restrict__ pcnt,
                                                    const uint32_t *__restrict__ efq, uint32_t ef,
                                                    uint64_t count, int nq,
                                                    uint32_t *__restrict__ order) {
    __shared__ uint64_t key[kOrderMax];
    int n2 = 1;
    while (n2 < nq) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        uint64_t k = ~0ull;
        if (i < nq) {
            const int p = pofq[i];
            const double pass = p < 0 ? (double)count : (double)max(pcnt[p], 1u);
            const double e = efq ? (double)max(1u, min(efq[i], ef)) : (double)ef;
            const float w = (float)(e * (double)count / pass);
            /* descending work, then ascending query: ascending key order */
            k = ((uint64_t)(~__float_as_uint(w)) << 32) | (uint32_t)i;
        }
        key[i] = k;
    }
    __syncthreads();
    for (int size = 2; size <= n2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const uint64_t a = key[i], b = key[j];
                    if ((a > b) == up) {
                        key[i] = b;
                        key[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < nq; i += blockDim.x) order[i] = (uint32_t)key[i];
}

size_t graph_shmem(int dim, int fcap, uint32_t ef) {
    return (((size_t)dim * 4 + 15) & ~(size_t)15) + (size_t)fcap * 8 + (size_t)ef * 8 +
           2 * (size_t)kGT * 8 + kHashSize * 4 + kMaxNew * 4 + kMaxNew * 4 +
           sizeof(GraphShared);
}

int run_graph_ex(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef, uint64_t max_visited, int wmax,
                 int fcap, uint32_t *dbg, const GraphExtra &x, char *errbuf, size_t errlen) {
    if (g->d_nbr == nullptr) PV_FAIL("gpu: graph not uploaded");
    if (ef < 1 || ef > FV_GPU_MAX_EF) PV_FAIL("gpu: bad graph batch (nq=%d ef=%u)", b->nq, ef);
    if (require_bitmaps(b, errbuf, errlen) != 0) return -1;
    if (b->max_entry >= g->count)
        PV_FAIL("gpu: entry point %llu out of range", (unsigned long long)b->max_entry);
    if (g->nbr_stride > kMaxNew) PV_FAIL("gpu: graph degree %d too large", g->nbr_stride);
    if (wmax < 1) wmax = 1;
    while (wmax > 1 && wmax * g->nbr_stride > kMaxNew) wmax--;
    /* fcap < 0: a bounded frontier of -fcap entries with no spill store —
     * live entries beyond it are dropped (counted in debug word 3), as in
     * the reference GPU kernel's fixed frontier (gpu_engine.cu:1396-1420);
     * deep filtered searches then end sooner at some cost in recall */
    const bool spill = fcap > 0;
    if (fcap < 0) fcap = -fcap;
    /* the LDS frontier: a power of two in [64, 1024] (its refill sorts up
     * to half of it with a bitonic network) */
    fcap = std::min(1024, std::max(64, fcap));
    while (fcap & (fcap - 1)) fcap &= fcap - 1;
    const size_t shm = graph_shmem(g->dim, fcap, ef);
    if (shm > 160 * 1024) PV_FAIL("gpu: graph search LDS %zu bytes exceeds the CU", shm);
    const uint64_t words = g->count_pad / 32;
    const uint64_t vwords = (words + 3) & ~3ull;
    PV_TRY(b->d_visit.ensure((size_t)b->nq * vwords * 4));
    /* frontier spill stores: every node enters a frontier at most once, so
     * min(max_visited, count) entries (plus one iteration's overshoot of
     * max_visited) never overflow; the total is capped at 4 GiB, beyond
     * which entries are lost and counted (debug word 3) */
    uint64_t scap = spill ? std::min<uint64_t>(max_visited, g->count) + kMaxNew : 0;
    const uint64_t kStoreBytes = (uint64_t)4 << 30;
    if (spill && (uint64_t)b->nq * (scap + fcap) * 8 > kStoreBytes)
        scap = std::max<uint64_t>(kStoreBytes / ((uint64_t)b->nq * 8), (uint64_t)fcap + 64) - fcap;
    PV_TRY(b->d_fspill.ensure((size_t)b->nq * (scap + fcap) * 8));
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
    A.spill = b->d_fspill.as<uint64_t>();
    A.scap = (uint32_t)scap;
    A.dbg = dbg;
    A.order = nullptr;
    if (shm > 64 * 1024)
        PV_TRY(hipFuncSetAttribute((const void *)k_graph,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipStream_t gs = x.stream ? x.stream : b->stream;
    g->ev_stream = gs;
    ev_record(g, 0);
    if (!x.bitmap_ready) PV_TRY(launch_bitmap(g, b, gs));
    /* heaviest searches first when the batch has predicates and more
     * queries than fill the chip at once */
    if (b->nprog > 0 && b->nq > 1 && b->nq <= kOrderMax && !b->capturing) {
        PV_TRY(b->d_order.ensure(((size_t)b->nq + (size_t)b->nprog) * 4));
        uint32_t *ord = b->d_order.as<uint32_t>();
        uint32_t *pc = ord + b->nq;
        hipLaunchKernelGGL(k_prog_count, dim3((unsigned)b->nprog), dim3(256), 0, gs,
                           b->d_bitmap.as<uint32_t>(), words, pc);
        hipLaunchKernelGGL(k_lpt_order, dim3(1), dim3(1024), 0, gs, b->d_pofq, pc, x.efq, ef,
                           g->count, b->nq, ord);
        PV_TRY(hipGetLastError());
        A.order = ord;
    }
    hipLaunchKernelGGL(k_graph, dim3(b->nq), dim3(kGT), shm, gs, A);
    PV_TRY(hipGetLastError());
    ev_record(g, 5);
    memset(&g->last, 0, sizeof(g->last));
    g->last.route = 3;
    return 0;
}

int build_search(FvGpuIndex *g, BuildSearchBufs &B, const uint32_t *nbr, int nbr_stride,
                 const float *qf, const FvGpuQueryDesc *qd, const int32_t *pofq, int nq,
                 int nq_reserve, uint32_t ef, int wmax, bool exact, uint32_t *dbg,
                 hipStream_t s, char *errbuf, size_t errlen) {
    if (nq < 1) return 0;
    if (ef < 1 || ef > FV_GPU_MAX_EF) PV_FAIL("hnsw: bad efConstruction %u", ef);
    if (nbr_stride > kMaxNew) PV_FAIL("hnsw: graph degree %d too large", nbr_stride);
    while (wmax > 1 && wmax * nbr_stride > kMaxNew) wmax--;
    if (wmax < 1) wmax = 1;
    const int fcap = 1024;
    const size_t shm = graph_shmem(g->dim, fcap, ef);
    if (shm > 160 * 1024) PV_FAIL("hnsw: search LDS %zu bytes exceeds the CU", shm);
    const uint64_t nr = (uint64_t)std::max(nq, nq_reserve);
    const uint64_t words = g->count_pad / 32;
    const uint64_t vwords = (words + 3) & ~3ull;
    /* the search never stops on max_visited (visited <= count) */
    const uint64_t max_visited = g->count + 1;
    uint64_t scap = exact ? g->count + kMaxNew : 0;
    const uint64_t kStoreBytes = (uint64_t)4 << 30;
    if (exact && nr * (scap + fcap) * 8 > kStoreBytes)
        scap = std::max<uint64_t>(kStoreBytes / (nr * 8), (uint64_t)fcap + 64) - fcap;
    PV_TRY(B.visit.ensure(nr * vwords * 4));
    PV_TRY(B.spill.ensure(exact ? nr * (scap + fcap) * 8 : 64));
    PV_TRY(B.status.ensure(nr * 4));
    PV_TRY(B.out.ensure(nr * ef * sizeof(FvHit)));
    PV_TRY(B.cnt.ensure(nr * 4));
    PV_TRY(B.exh.ensure(nr));
    GraphArgs A;
    memset(&A, 0, sizeof(A));
    A.vecs = g->d_vecs;
    A.nbr = nbr;
    A.nbr_stride = nbr_stride;
    A.count = g->count;
    A.dim = g->dim;
    A.qf = qf;
    A.qd = qd;
    A.pofq = pofq;
    A.bitmap = nullptr;
    A.words = words;
    A.vwords = vwords;
    A.ef = ef;
    A.max_visited = max_visited;
    A.wmax = wmax;
    A.fcap = fcap;
    A.visit = B.visit.as<uint32_t>();
    A.out = B.out.as<FvHit>();
    A.out_cnt = B.cnt.as<uint32_t>();
    A.out_exh = B.exh.as<uint8_t>();
    A.status = B.status.as<uint32_t>();
    A.out_k = ef;
    A.spill = B.spill.as<uint64_t>();
    A.scap = (uint32_t)scap;
    A.dbg = dbg;
    if (shm > 64 * 1024)
        PV_TRY(hipFuncSetAttribute((const void *)k_graph,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(k_graph, dim3(nq), dim3(kGT), shm, s, A);
    PV_TRY(hipGetLastError());
    return 0;
}

int run_graph_opts(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef, uint64_t max_visited,
                   int wmax, int fcap, uint32_t *dbg, char *errbuf, size_t errlen) {
    return run_graph_ex(g, b, ef, max_visited, wmax, fcap, dbg, GraphExtra(), errbuf, errlen);
}

int run_graph(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef, uint64_t max_visited,
              char *errbuf, size_t errlen) {
    return run_graph_opts(g, b, ef, max_visited, 8, 1024, nullptr, errbuf, errlen);
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
    feedback_begin(g, b->stream);
    int rc = run_graph_opts(g, b, ef, max_visited, wmax, fcap, d_dbg, errbuf, errlen);
    if (rc == 0) feedback_end(g, b, b->stream, b->nprog > 0);
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
            feedback_record(g, b, 1, qd2.data(), ef, b->nprog > 0);
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
                          ninvals, ef, max_visited, 8, 1024, out, out_cnt, exhausted,
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

extern "C" int fv_gpu_batch_run_graph_ex(FvGpuIndex *g, FvGpuBatch *b, uint32_t ef,
                                         uint64_t max_visited, int wmax, int fcap,
                                         char *errbuf, size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g) PV_FAIL("gpu: bad batch");
    try {
        return run_graph_opts(g, b, ef, max_visited, wmax, fcap, nullptr, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}
