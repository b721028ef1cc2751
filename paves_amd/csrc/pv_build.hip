/*
 * pv_build.hip — HNSW construction with level 0 on MI355X.
 *
 * The reference builds its graph on the host, one insertion at a time on a
 * thread pool (HnswBuild, engine.cpp:698-885: efConstruction beam, the
 * neighbour heuristic keeping M links, reverse links re-pruned once a list
 * holds M0 = 2M). Level 0 is where nearly all of that work lies (every
 * node, the widest lists), so this builder keeps the host's insertion for
 * the upper levels (about 1/M of the nodes) and inserts the rest at level 0
 * in batches on the GPU, each batch against the graph as it stood before
 * it:
 *   1. k_gather_batch: the batch's rows and query descriptors (entry point
 *      from the host's greedy descent through the finished upper levels);
 *   2. k_graph (pv_graph.hip): one unfiltered efConstruction search per
 *      node, the query cursor's beam (with one head per iteration and the
 *      unbounded frontier it visits what the reference beam visits);
 *   3. k_prune: one wave per new node applies the heuristic to its ranked
 *      hits (skip a hit when a kept node is closer to it than the new node
 *      is; at most M kept; all kept when there are no more than M hits —
 *      the reference prune's early return), writes the node's list and one
 *      reverse-link key (target << 32 | node) per kept link;
 *   4. radix sort of the keys: each target's incoming links in node order;
 *   5. k_heads / k_link: one wave per target merges its incoming links —
 *      appended while the list has room, else the list plus up to 64 - cnt
 *      incoming links at a time are ranked by distance to the target
 *      (64-key register sort) and re-pruned to M0 by the heuristic.
 * Every distance is the exact FP32 distance in the reference's summation
 * order (cpu_order_l2, 16 lanes per pair, four pairs per wave), so the
 * result is deterministic and, with one head per search iteration and the
 * unbounded frontier, equals a sequential restatement of these batch
 * semantics bit for bit.
 *
 * Adjacency under construction: [count][nbr_stride] u32, kNoNode-padded,
 * lists compact — the layout k_graph reads, so the finished buffer can stay
 * on the device as the index's graph.
 */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "pv_device.h"
#include "pv_hnsw.h"

namespace pv {

constexpr int kBW = 4;           /* waves per workgroup of the prune and link kernels */
constexpr int kLinkPool = 64;    /* candidates ranked at once for one target */
constexpr int kLinkGrid = 2048;  /* workgroups of the persistent link kernel */

/* rows and descriptors of batch nodes ids[0..nb): queries of k_graph */
__global__ void k_gather_batch(const float *__restrict__ vecs, int dim,
                               const uint32_t *__restrict__ ids,
                               const uint32_t *__restrict__ entries, int nb,
                               float *__restrict__ qf, FvGpuQueryDesc *__restrict__ qd) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)nb * dim;
    if (i < total) {
        const uint64_t q = i / dim, d = i % dim;
        qf[i] = vecs[(uint64_t)ids[q] * dim + d];
    }
    if (i < (uint64_t)nb) {
        FvGpuQueryDesc x;
        x.proot = -1;
        x.need_eval = 0;
        x.span0 = 0;
        x.nspans = 0;
        x.soff0 = 0;
        x.entry = entries[i];
        x.total = 0;
        qd[i] = x;
    }
}

/*
 * Whether a kept node is closer to candidate x than the base node is
 * (dx = d(base, x)): kept ids are in lanes 0..nk-1 of kreg; four kept nodes
 * per pass, one per 16-lane group. Wave-uniform; every lane takes part.
 */
__device__ __forceinline__ bool heur_reject(const float *__restrict__ vecs, int dim,
                                            uint32_t x, float dx, uint32_t kreg, int nk) {
    const int lane = threadIdx.x & 63, grp = lane >> 4, l16 = lane & 15;
    const float *xr = vecs + (uint64_t)x * dim;
    for (int k0 = 0; k0 < nk; k0 += 4) {
        const int kk = k0 + grp;
        const uint32_t kid = (uint32_t)__shfl((int)kreg, kk < nk ? kk : nk - 1);
        const float d = cpu_order_l2(xr, vecs + (uint64_t)kid * dim, dim, l16);
        if (__any(kk < nk && d < dx)) return true;
    }
    return false;
}

/*
 * One wave per new node: the heuristic over its search hits (ascending
 * (dist, node)), the node's level-0 list and its reverse-link keys.
 */
__global__ __launch_bounds__(kBW * 64) void k_prune(
    const float *__restrict__ vecs, int dim, const uint32_t *__restrict__ ids, int nb,
    const FvHit *__restrict__ hits, const uint32_t *__restrict__ hcnt, uint32_t hstride,
    int M, uint32_t *__restrict__ nbr, int nbs, uint64_t *__restrict__ keys) {
    const int wv = blockIdx.x * kBW + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wv >= nb) return;
    const uint32_t u = ids[wv];
    const FvHit *h = hits + (uint64_t)wv * hstride;
    const int cnt = (int)min(hcnt[wv], hstride);
    uint32_t kreg = kNoNode;
    int nk = 0;
    if (cnt <= M) {
        /* no more hits than links: all kept (the reference prune returns) */
        if (lane < cnt) kreg = (uint32_t)h[lane].node;
        nk = cnt;
    } else {
        for (int c = 0; c < cnt && nk < M; c++) {
            const uint32_t x = (uint32_t)h[c].node;
            const float dx = h[c].dist;
            if (!heur_reject(vecs, dim, x, dx, kreg, nk)) {
                if (lane == nk) kreg = x;
                nk++;
            }
        }
    }
    if (lane < nk) nbr[(uint64_t)u * nbs + lane] = kreg;
    if (lane < M)
        keys[(uint64_t)wv * M + lane] =
            lane < nk ? (((uint64_t)kreg << 32) | u) : ~0ull;
}

/* sorted keys: index of the first key of every target's run */
__global__ void k_heads(const uint64_t *__restrict__ keys, uint64_t n,
                        uint32_t *__restrict__ heads, uint32_t *__restrict__ nheads) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    if (k == ~0ull) return;
    if (i > 0 && (keys[i - 1] >> 32) == (k >> 32)) return;
    heads[atomicAdd(nheads, 1u)] = (uint32_t)i;
}

/*
 * One wave per target (persistent over the head list): merge the batch's
 * incoming links (sorted by node) into the target's list.
 */
__global__ __launch_bounds__(kBW * 64) void k_link(
    const float *__restrict__ vecs, int dim, const uint64_t *__restrict__ keys, uint64_t nkeys,
    const uint32_t *__restrict__ heads, const uint32_t *__restrict__ nheads, int M0,
    uint32_t *__restrict__ nbr, int nbs) {
    const int lane = threadIdx.x & 63, grp = lane >> 4, l16 = lane & 15;
    const uint32_t nh = *nheads;
    for (uint32_t hi = blockIdx.x * kBW + (threadIdx.x >> 6); hi < nh;
         hi += gridDim.x * kBW) {
        const uint64_t s = heads[hi];
        const uint32_t v = (uint32_t)(keys[s] >> 32);
        /* end of the run */
        uint64_t e = s + 1;
        for (;;) {
            const uint64_t j = e + lane;
            const bool in = j < nkeys && (uint32_t)(keys[j] >> 32) == v && keys[j] != ~0ull;
            const uint64_t m = __ballot(in);
            if (m == ~0ull) {
                e += 64;
                continue;
            }
            e += (uint64_t)__ffsll((unsigned long long)~m) - 1;
            break;
        }
        const uint64_t r = e - s;
        uint32_t *lst = nbr + (uint64_t)v * nbs;
        uint32_t mine = lane < M0 ? lst[lane] : kNoNode;
        int c = __popcll(__ballot(mine != kNoNode));
        const float *vr = vecs + (uint64_t)v * dim;
        uint64_t pos = 0;
        while (pos < r) {
            if ((uint64_t)c + (r - pos) <= (uint64_t)M0) {
                /* room for the rest: appended in node order */
                const int nadd = (int)(r - pos);
                if (lane >= c && lane < c + nadd) mine = (uint32_t)keys[s + pos + (lane - c)];
                c += nadd;
                pos = r;
                break;
            }
            const int take = (int)min<uint64_t>(r - pos, (uint64_t)(kLinkPool - c));
            const uint32_t cid =
                lane < c ? mine
                         : (lane < c + take ? (uint32_t)keys[s + pos + (lane - c)] : kNoNode);
            pos += take;
            const int n = c + take;
            /* d(v, candidate of lane j): pair 4t+grp in pass t */
            float dme = 0.f;
            for (int t = 0; t < kLinkPool / 4; t++) {
                if (4 * t >= n) break;
                const uint32_t x = (uint32_t)__shfl((int)cid, 4 * t + grp);
                const uint32_t xs = x == kNoNode ? v : x;
                const float d = cpu_order_l2(vr, vecs + (uint64_t)xs * dim, dim, l16);
                const float got = __shfl(d, 16 * (lane & 3));
                if ((lane >> 2) == t) dme = got;
            }
            const uint64_t key = cid == kNoNode ? ~0ull : make_key(dme, cid);
            const uint64_t sk = wave_sort64(key);
            uint32_t kreg = kNoNode;
            int nk = 0;
            for (int j = 0; j < n && nk < M0; j++) {
                const uint64_t kj =
                    ((uint64_t)(uint32_t)__shfl((int)(sk >> 32), j) << 32) |
                    (uint32_t)__shfl((int)(uint32_t)sk, j);
                const uint32_t x = key_row(kj);
                if (!heur_reject(vecs, dim, x, key_dist(kj), kreg, nk)) {
                    if (lane == nk) kreg = x;
                    nk++;
                }
            }
            mine = lane < nk ? kreg : kNoNode;
            c = nk;
        }
        if (lane < M0) lst[lane] = lane < c ? mine : kNoNode;
    }
}

/* nodes visited (acc[0]) and heads expanded (acc[1]) by a batch's
 * searches: debug words 1 and 0 of each query */
__global__ void k_sum_visited(const uint32_t *__restrict__ dbg, int nq,
                              unsigned long long *__restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) {
        atomicAdd(&acc[0], (unsigned long long)dbg[(uint64_t)i * 4 + 1]);
        atomicAdd(&acc[1], (unsigned long long)dbg[(uint64_t)i * 4 + 0]);
    }
}

} // namespace pv

using namespace pv;

namespace {

struct BuildBufs {
    uint32_t *nbr = nullptr;
    uint32_t *ids = nullptr, *ents = nullptr;
    float *qf = nullptr;
    FvGpuQueryDesc *qd = nullptr;
    int32_t *pofq = nullptr;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    uint32_t *heads = nullptr, *nheads = nullptr;
    uint32_t *dbg = nullptr;
    unsigned long long *visited = nullptr;
    void *tmp = nullptr;
    hipStream_t s = nullptr;
    BuildSearchBufs sb;
    std::vector<hipEvent_t> ev;
    ~BuildBufs() {
        if (s) hipStreamSynchronize(s);
        for (hipEvent_t e : ev) hipEventDestroy(e);
        hipFree(dbg);
        hipFree(visited);
        hipFree(nbr);
        hipFree(ids);
        hipFree(ents);
        hipFree(qf);
        hipFree(qd);
        hipFree(pofq);
        hipFree(k0);
        hipFree(k1);
        hipFree(heads);
        hipFree(nheads);
        hipFree(tmp);
        sb.release();
        if (s) hipStreamDestroy(s);
    }
};

double secs_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

/* the device phase; fills l0 ([count][M0 + 1]) from the device lists */
int device_phase(FvGpuIndex *g, BuildBufs &B, const std::vector<uint32_t> &rest,
                 const std::vector<uint32_t> &entries, uint32_t *l0, uint32_t l0s, int M,
                 int efc, FvHnswGpuOpts &o, char *errbuf, size_t errlen) {
    const uint64_t n = g->count;
    const int dim = g->dim;
    const int M0 = 2 * M;
    const int nbs = (M0 + 7) & ~7;
    /* batch sizes: the visited bitmaps of one batch stay within 8 GiB */
    const uint64_t vbytes = (((g->count_pad / 32) + 3) & ~3ull) * 4;
    int bmax = std::max(1, o.max_batch);
    bmax = (int)std::min<uint64_t>((uint64_t)bmax, std::max<uint64_t>(1, ((uint64_t)8 << 30) / vbytes));
    const int div = std::max(1, o.batch_div);
    const uint64_t nrest = rest.size();
    std::vector<int> sizes;
    for (uint64_t done = 0, have = n - nrest; done < nrest;) {
        const int b = (int)std::min<uint64_t>(
            nrest - done, std::max<uint64_t>(1, std::min<uint64_t>(have / div, (uint64_t)bmax)));
        sizes.push_back(b);
        done += b;
        have += b;
    }
    const int bpeak = sizes.empty() ? 1 : *std::max_element(sizes.begin(), sizes.end());
    o.batches = (int)sizes.size();

    PV_TRY(hipStreamCreateWithFlags(&B.s, hipStreamNonBlocking));
    /* adjacency: the host phase's level-0 lists, kNoNode-padded */
    {
        std::vector<uint32_t> pad((size_t)n * nbs, kNoNode);
        for (uint64_t v = 0; v < n; v++) {
            const uint32_t *lst = l0 + v * l0s;
            const uint32_t c = std::min<uint32_t>(lst[0], (uint32_t)M0);
            memcpy(&pad[v * nbs], lst + 1, (size_t)c * 4);
        }
        PV_TRY(hipMalloc(&B.nbr, pad.size() * 4));
        PV_TRY(hipMemcpy(B.nbr, pad.data(), pad.size() * 4, hipMemcpyHostToDevice));
    }
    if (nrest == 0) return 0;
    PV_TRY(hipMalloc(&B.ids, nrest * 4));
    PV_TRY(hipMalloc(&B.ents, nrest * 4));
    PV_TRY(hipMemcpy(B.ids, rest.data(), nrest * 4, hipMemcpyHostToDevice));
    PV_TRY(hipMemcpy(B.ents, entries.data(), nrest * 4, hipMemcpyHostToDevice));
    PV_TRY(hipMalloc(&B.qf, (size_t)bpeak * dim * 4));
    PV_TRY(hipMalloc(&B.qd, (size_t)bpeak * sizeof(FvGpuQueryDesc)));
    PV_TRY(hipMalloc(&B.pofq, (size_t)bpeak * 4));
    PV_TRY(hipMemset(B.pofq, 0xFF, (size_t)bpeak * 4)); /* -1: no predicate */
    const uint64_t pk = (uint64_t)bpeak * M;
    PV_TRY(hipMalloc(&B.k0, pk * 8));
    PV_TRY(hipMalloc(&B.k1, pk * 8));
    PV_TRY(hipMalloc(&B.heads, pk * 4));
    PV_TRY(hipMalloc(&B.nheads, 4));
    size_t tmpb = 0;
    PV_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tmpb, B.k0, B.k1, (int)pk, 0, 64, B.s));
    PV_TRY(hipMalloc(&B.tmp, std::max<size_t>(tmpb, 256)));
    PV_TRY(hipMalloc(&B.dbg, (size_t)bpeak * 16));
    PV_TRY(hipMalloc(&B.visited, 16));
    PV_TRY(hipMemset(B.visited, 0, 16));
    /* stage timing: four events per batch while there are few batches */
    const bool timed = sizes.size() <= 512;
    if (timed) {
        B.ev.resize(4 * sizes.size());
        for (hipEvent_t &e : B.ev) PV_TRY(hipEventCreate(&e));
    }

    const int heads = std::min(8, std::max(1, o.heads));
    uint64_t off = 0;
    for (size_t bi = 0; bi < sizes.size(); bi++) {
        const int nb = sizes[bi];
        const uint64_t tot = (uint64_t)nb * dim;
        if (timed) PV_TRY(hipEventRecord(B.ev[4 * bi], B.s));
        hipLaunchKernelGGL(k_gather_batch, dim3((unsigned)((std::max<uint64_t>(tot, nb) + 255) / 256)),
                           dim3(256), 0, B.s, g->d_vecs, dim, B.ids + off, B.ents + off, nb, B.qf,
                           B.qd);
        PV_TRY(hipGetLastError());
        if (build_search(g, B.sb, B.nbr, nbs, B.qf, B.qd, B.pofq, nb, bpeak, (uint32_t)efc,
                         heads, o.exact != 0, B.dbg, B.s, errbuf, errlen) != 0)
            return -1;
        hipLaunchKernelGGL(k_sum_visited, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, B.s,
                           B.dbg, nb, B.visited);
        if (timed) PV_TRY(hipEventRecord(B.ev[4 * bi + 1], B.s));
        hipLaunchKernelGGL(k_prune, dim3((unsigned)((nb + kBW - 1) / kBW)), dim3(kBW * 64), 0, B.s,
                           g->d_vecs, dim, B.ids + off, nb, B.sb.out.as<FvHit>(),
                           B.sb.cnt.as<uint32_t>(), (uint32_t)efc, M, B.nbr, nbs, B.k0);
        PV_TRY(hipGetLastError());
        if (timed) PV_TRY(hipEventRecord(B.ev[4 * bi + 2], B.s));
        const uint64_t nk = (uint64_t)nb * M;
        size_t tb = tmpb;
        PV_TRY(hipcub::DeviceRadixSort::SortKeys(B.tmp, tb, B.k0, B.k1, (int)nk, 0, 64, B.s));
        PV_TRY(hipMemsetAsync(B.nheads, 0, 4, B.s));
        hipLaunchKernelGGL(k_heads, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, B.s, B.k1, nk,
                           B.heads, B.nheads);
        PV_TRY(hipGetLastError());
        const unsigned lg = (unsigned)std::min<uint64_t>(kLinkGrid, (nk + kBW - 1) / kBW);
        hipLaunchKernelGGL(k_link, dim3(lg), dim3(kBW * 64), 0, B.s, g->d_vecs, dim, B.k1, nk,
                           B.heads, B.nheads, M0, B.nbr, nbs);
        PV_TRY(hipGetLastError());
        if (timed) PV_TRY(hipEventRecord(B.ev[4 * bi + 3], B.s));
        off += nb;
    }
    PV_TRY(hipStreamSynchronize(B.s));
    if (timed) {
        double st[3] = {0, 0, 0};
        for (size_t bi = 0; bi < sizes.size(); bi++)
            for (int k = 0; k < 3; k++) {
                float ms = 0.f;
                PV_TRY(hipEventElapsedTime(&ms, B.ev[4 * bi + k], B.ev[4 * bi + k + 1]));
                st[k] += ms * 1e-3;
            }
        o.t_search_s = st[0];
        o.t_prune_s = st[1];
        o.t_link_s = st[2];
    }
    {
        unsigned long long vis[2] = {0, 0};
        PV_TRY(hipMemcpy(vis, B.visited, 16, hipMemcpyDeviceToHost));
        o.visited = vis[0];
        o.heads_expanded = vis[1];
    }
    /* device lists -> [cnt, ids...] */
    std::vector<uint32_t> pad((size_t)n * nbs);
    PV_TRY(hipMemcpy(pad.data(), B.nbr, pad.size() * 4, hipMemcpyDeviceToHost));
    for (uint64_t v = 0; v < n; v++) {
        uint32_t *lst = l0 + v * l0s;
        uint32_t c = 0;
        while (c < (uint32_t)M0 && pad[v * nbs + c] != kNoNode) c++;
        lst[0] = c;
        memcpy(lst + 1, &pad[v * nbs], (size_t)c * 4);
        for (uint32_t j = c; j < (uint32_t)M0; j++) lst[1 + j] = 0;
    }
    return 0;
}

} // namespace

extern "C" FvHnswGraph *fv_hnsw_build_gpu(FvGpuIndex *g, const float *vectors, int M,
                                          int ef_construction, int threads, FvHnswGpuOpts *opts,
                                          char *errbuf, size_t errlen) {
    FvHnswGpuOpts o;
    memset(&o, 0, sizeof(o));
    o.heads = 8;
    o.batch_div = 8;
    o.max_batch = 32768;
    o.keep = 1;
    if (opts) {
        o.heads = opts->heads > 0 ? opts->heads : o.heads;
        o.exact = opts->exact;
        o.batch_div = opts->batch_div > 0 ? opts->batch_div : o.batch_div;
        o.max_batch = opts->max_batch > 0 ? opts->max_batch : o.max_batch;
        o.keep = opts->keep;
    }
    if (g == nullptr || vectors == nullptr || g->d_vecs == nullptr) {
        snprintf(errbuf, errlen, "hnsw: NULL index or vectors");
        return nullptr;
    }
    if (M < 2 || M > 16 || ef_construction < 1 || ef_construction > FV_GPU_MAX_EF) {
        snprintf(errbuf, errlen,
                 "hnsw: device build needs 2 <= M <= 16 and 1 <= efConstruction <= %d "
                 "(M=%d efc=%d)", FV_GPU_MAX_EF, M, ef_construction);
        return nullptr;
    }
    if (hipSetDevice(g->device) != hipSuccess) {
        snprintf(errbuf, errlen, "hnsw: hipSetDevice(%d) failed", g->device);
        return nullptr;
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> rest, entries;
    double te = 0.0;
    FvHnswGraph *h = hnsw_seed(vectors, g->count, g->dim, M, ef_construction, threads, rest,
                               entries, errbuf, errlen, &te);
    if (h == nullptr) return nullptr;
    const double th = secs_since(t0);
    if (opts) {
        opts->t_host_s = th;
        opts->t_entries_s = te;
    }
    const auto t1 = std::chrono::steady_clock::now();
    uint32_t l0s = 0;
    uint32_t *l0 = hnsw_l0(h, &l0s);
    int rc;
    {
        BuildBufs B;
        rc = device_phase(g, B, rest, entries, l0, l0s, M, ef_construction, o, errbuf, errlen);
        if (rc == 0 && o.keep) {
            /* the built lists become the index's graph, once no queued
             * search can still read the old one */
            hipDeviceSynchronize();
            hipFree(g->d_nbr);
            g->d_nbr = B.nbr;
            g->nbr_stride = (2 * M + 7) & ~7;
            g->l0_stride = l0s;
            B.nbr = nullptr;
        }
    }
    hnsw_detach(h);
    if (rc != 0) {
        fv_hnsw_free(h);
        return nullptr;
    }
    if (opts) {
        opts->t_device_s = secs_since(t1);
        opts->t_search_s = o.t_search_s;
        opts->t_prune_s = o.t_prune_s;
        opts->t_link_s = o.t_link_s;
        opts->visited = o.visited;
        opts->heads_expanded = o.heads_expanded;
        opts->batches = o.batches;
    }
    return h;
}
