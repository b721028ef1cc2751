/*
 * pv_hnsw.cpp — host-side HNSW graph construction and entry-point descent.
 *
 * Builds the graph the GPU beam search walks, in the reference's on-disk
 * layout: level-0 lists [cnt, n1..nM0] with stride 2M+1, per-node upper
 * slabs of (M+1) words per level addressed by an offsets array, the global
 * entry node and top level (engine.cpp:190-215 file sections).
 *
 * Construction semantics are the reference builder's (HnswBuild,
 * engine.cpp:698-885): deterministic levels from splitmix64 (Steele, Lea &
 * Flood 2014) with mL = 1/ln M, greedy descent above the node's level, an
 * efConstruction beam per level, the HNSW neighbour heuristic (Malkov &
 * Yashunin 2018, alg. 4 without extension) keeping M links for the new
 * node, reverse links appended up to M0 = 2M (level 0) or M, re-pruned with
 * the same heuristic once full; node 0 first, then the rest in parallel.
 * With one thread the graph is bit-identical to the reference builder's
 * (insertion order and every distance equal), which the tests check; with
 * several threads, like the reference, the result depends on scheduling.
 *
 * Distances use the reference's FP32 summation order (engine.cpp:43-90):
 * two 16-lane FMA accumulators over 32-dim blocks, one more 16-lane block
 * into the first, the 16->8->4->2->1 reduction tree, then the tail (an
 * 8-wide separately-rounded block when 8 or more remain, then FMAs). An
 * AVX-512, an AVX2 and a portable version compute identical results and are
 * chosen at run time.
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <queue>
#include <thread>
#include <utility>
#include <vector>

#include "../../../include/paves_gpu.h"
#include "../pv_hnsw.h"

namespace {

/* ------------------------------------------------------------------ */
/* distance in the reference's FP32 order                              */

inline float tree16(const float *t16) {
    float t8[8], t4[4];
    for (int j = 0; j < 8; j++) t8[j] = t16[8 + j] + t16[j];
    for (int j = 0; j < 4; j++) t4[j] = t8[4 + j] + t8[j];
    const float s0 = t4[0] + t4[2];
    const float s1 = t4[1] + t4[3];
    return s0 + s1;
}

inline float tail(const float *a, const float *b, int i, int dim, float sum) {
    if (dim - i >= 8) {
        for (int j = 0; j < 8; j++, i++) {
            const float d = a[i] - b[i];
            const float p = d * d;
            sum = sum + p;
        }
    }
    for (; i < dim; i++) {
        const float d = a[i] - b[i];
        sum = fmaf(d, d, sum);
    }
    return sum;
}

float l2_portable(const float *a, const float *b, int dim) {
    float acc0[16] = {}, acc1[16] = {};
    int i = 0;
    for (; i + 32 <= dim; i += 32)
        for (int j = 0; j < 16; j++) {
            const float d0 = a[i + j] - b[i + j];
            const float d1 = a[i + 16 + j] - b[i + 16 + j];
            acc0[j] = fmaf(d0, d0, acc0[j]);
            acc1[j] = fmaf(d1, d1, acc1[j]);
        }
    for (; i + 16 <= dim; i += 16)
        for (int j = 0; j < 16; j++) {
            const float d = a[i + j] - b[i + j];
            acc0[j] = fmaf(d, d, acc0[j]);
        }
    float t[16];
    for (int j = 0; j < 16; j++) t[j] = acc0[j] + acc1[j];
    return tail(a, b, i, dim, tree16(t));
}

__attribute__((target("avx2,fma"))) float l2_avx2(const float *a, const float *b, int dim) {
    /* a 16-lane accumulator as two 8-lane halves: the same per-lane ops */
    __m256 a0l = _mm256_setzero_ps(), a0h = a0l, a1l = a0l, a1h = a0l;
    int i = 0;
    for (; i + 32 <= dim; i += 32) {
        const __m256 d0l = _mm256_sub_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i));
        const __m256 d0h = _mm256_sub_ps(_mm256_loadu_ps(a + i + 8), _mm256_loadu_ps(b + i + 8));
        const __m256 d1l = _mm256_sub_ps(_mm256_loadu_ps(a + i + 16), _mm256_loadu_ps(b + i + 16));
        const __m256 d1h = _mm256_sub_ps(_mm256_loadu_ps(a + i + 24), _mm256_loadu_ps(b + i + 24));
        a0l = _mm256_fmadd_ps(d0l, d0l, a0l);
        a0h = _mm256_fmadd_ps(d0h, d0h, a0h);
        a1l = _mm256_fmadd_ps(d1l, d1l, a1l);
        a1h = _mm256_fmadd_ps(d1h, d1h, a1h);
    }
    for (; i + 16 <= dim; i += 16) {
        const __m256 dl = _mm256_sub_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i));
        const __m256 dh = _mm256_sub_ps(_mm256_loadu_ps(a + i + 8), _mm256_loadu_ps(b + i + 8));
        a0l = _mm256_fmadd_ps(dl, dl, a0l);
        a0h = _mm256_fmadd_ps(dh, dh, a0h);
    }
    float t[16];
    _mm256_storeu_ps(t, _mm256_add_ps(a0l, a1l));
    _mm256_storeu_ps(t + 8, _mm256_add_ps(a0h, a1h));
    return tail(a, b, i, dim, tree16(t));
}

__attribute__((target("avx512f"))) float l2_avx512(const float *a, const float *b, int dim) {
    __m512 acc0 = _mm512_setzero_ps(), acc1 = acc0;
    int i = 0;
    for (; i + 32 <= dim; i += 32) {
        const __m512 d0 = _mm512_sub_ps(_mm512_loadu_ps(a + i), _mm512_loadu_ps(b + i));
        const __m512 d1 = _mm512_sub_ps(_mm512_loadu_ps(a + i + 16), _mm512_loadu_ps(b + i + 16));
        acc0 = _mm512_fmadd_ps(d0, d0, acc0);
        acc1 = _mm512_fmadd_ps(d1, d1, acc1);
    }
    for (; i + 16 <= dim; i += 16) {
        const __m512 d = _mm512_sub_ps(_mm512_loadu_ps(a + i), _mm512_loadu_ps(b + i));
        acc0 = _mm512_fmadd_ps(d, d, acc0);
    }
    float t[16];
    _mm512_storeu_ps(t, _mm512_add_ps(acc0, acc1));
    return tail(a, b, i, dim, tree16(t));
}

using L2Fn = float (*)(const float *, const float *, int);

L2Fn pick_l2() {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f")) return l2_avx512;
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return l2_avx2;
    return l2_portable;
}

const L2Fn l2 = pick_l2();

/* ------------------------------------------------------------------ */

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* level of node i: floor(-ln(u) / ln M), u in (0, 1] from splitmix64 of a
 * per-node seed (engine.cpp:716-730), capped at 30 */
int node_level(uint64_t i, double mL) {
    double u = (double)(splitmix64(i * 0x5851F42D4C957F2Dull + 12345) >> 11) *
               (1.0 / 9007199254740992.0);
    if (u < 1e-18) u = 1e-18;
    const int lvl = (int)(-log(u) * mL);
    return lvl > 30 ? 30 : lvl;
}

struct SpinLock {
    std::atomic_flag f = ATOMIC_FLAG_INIT;
    void lock() {
        while (f.test_and_set(std::memory_order_acquire)) {
#if defined(__x86_64__)
            _mm_pause();
#endif
        }
    }
    void unlock() { f.clear(std::memory_order_release); }
};

using DN = std::pair<float, uint32_t>; /* (distance, node): ordered as the reference's pairs */

struct Graph {
    const float *X = nullptr;
    int dim = 0;
    uint64_t n = 0;
    uint32_t M = 0, M0 = 0, efc = 0;
    std::vector<uint8_t> level;
    std::vector<uint32_t> l0;        /* n x (M0 + 1) */
    std::vector<uint64_t> uoff;      /* n + 1 slab offsets into up */
    std::vector<uint32_t> up;        /* level slabs of (M + 1) */
    std::unique_ptr<SpinLock[]> lk;
    std::atomic<uint64_t> entry{0};
    std::atomic<int> top{-1};
    std::mutex top_mu;

    const float *row(uint64_t i) const { return X + i * (uint64_t)dim; }
    uint32_t *links(uint64_t node, int lev) {
        if (lev == 0) return &l0[node * (uint64_t)(M0 + 1)];
        return &up[uoff[node] + (uint64_t)(lev - 1) * (M + 1)];
    }
    /* copy of a node's list at a level, taken under its lock */
    int snapshot(uint64_t node, int lev, uint32_t *dst) {
        lk[node].lock();
        const uint32_t *s = links(node, lev);
        const int c = (int)s[0];
        memcpy(dst, s + 1, (size_t)c * 4);
        lk[node].unlock();
        return c;
    }

    uint64_t descend(const float *q, uint64_t ep, int lev) {
        float best = l2(q, row(ep), dim);
        uint32_t nb[64];
        for (bool moved = true; moved;) {
            moved = false;
            const int c = snapshot(ep, lev, nb);
            for (int i = 0; i < c; i++) {
                const float d = l2(q, row(nb[i]), dim);
                if (d < best) {
                    best = d;
                    ep = nb[i];
                    moved = true;
                }
            }
        }
        return ep;
    }

    /* efc-wide beam at one level; result ascending by (dist, node) */
    void beam(const float *q, uint64_t ep, int lev, std::vector<uint32_t> &tag, uint32_t stamp,
              std::vector<DN> &res_out) {
        std::priority_queue<DN, std::vector<DN>, std::greater<DN>> open;
        std::priority_queue<DN> best;
        const float d0 = l2(q, row(ep), dim);
        open.push({d0, (uint32_t)ep});
        best.push({d0, (uint32_t)ep});
        tag[ep] = stamp;
        uint32_t nb[64];
        while (!open.empty()) {
            const DN c = open.top();
            if (best.size() >= efc && c.first > best.top().first) break;
            open.pop();
            const int cnt = snapshot(c.second, lev, nb);
            for (int i = 0; i < cnt; i++) {
                const uint32_t v = nb[i];
                if (tag[v] == stamp) continue;
                tag[v] = stamp;
                const float d = l2(q, row(v), dim);
                if (best.size() < efc || d < best.top().first) {
                    open.push({d, v});
                    best.push({d, v});
                    if (best.size() > efc) best.pop();
                }
            }
        }
        res_out.resize(best.size());
        for (size_t k = best.size(); k-- > 0;) {
            res_out[k] = best.top();
            best.pop();
        }
    }

    /* neighbour heuristic over candidates sorted ascending: keep c unless a
     * kept node is closer to c than the base node is; stop at m kept */
    void prune(std::vector<DN> &c, uint32_t m) const {
        if (c.size() <= m) return;
        std::vector<DN> keep;
        keep.reserve(m);
        for (const DN &x : c) {
            if (keep.size() >= m) break;
            bool ok = true;
            for (const DN &k : keep)
                if (l2(row(x.second), row(k.second), dim) < x.first) {
                    ok = false;
                    break;
                }
            if (ok) keep.push_back(x);
        }
        c.swap(keep);
    }

    /* reverse link node -> target, re-pruning a full list */
    void add_link(uint64_t node, uint32_t target, int lev) {
        const uint32_t cap = lev == 0 ? M0 : M;
        lk[node].lock();
        uint32_t *s = links(node, lev);
        if (s[0] < cap) {
            s[1 + s[0]++] = target;
            lk[node].unlock();
            return;
        }
        std::vector<DN> c;
        c.reserve(cap + 1);
        c.push_back({l2(row(node), row(target), dim), target});
        for (uint32_t i = 0; i < s[0]; i++) c.push_back({l2(row(node), row(s[1 + i]), dim), s[1 + i]});
        std::sort(c.begin(), c.end());
        prune(c, cap);
        s[0] = (uint32_t)c.size();
        for (size_t i = 0; i < c.size(); i++) s[1 + i] = c[i].second;
        lk[node].unlock();
    }

    void insert(uint64_t node, std::vector<uint32_t> &tag, uint32_t &stamp) {
        const float *q = row(node);
        const int lv = level[node];
        int t = top.load(std::memory_order_acquire);
        uint64_t ep = entry.load(std::memory_order_acquire);
        if (t < 0) {
            std::lock_guard<std::mutex> g(top_mu);
            if (top.load() < 0) {
                entry.store(node);
                top.store(lv);
                return;
            }
            t = top.load();
            ep = entry.load();
        }
        for (int l = t; l > lv; l--) ep = descend(q, ep, l);
        std::vector<DN> found, mine;
        for (int l = std::min(t, lv); l >= 0; l--) {
            beam(q, ep, l, tag, ++stamp, found);
            if (!found.empty()) ep = found[0].second;
            mine = found;
            prune(mine, M);
            lk[node].lock();
            uint32_t *s = links(node, l);
            s[0] = (uint32_t)mine.size();
            for (size_t i = 0; i < mine.size(); i++) s[1 + i] = mine[i].second;
            lk[node].unlock();
            for (const DN &x : mine) add_link(x.second, (uint32_t)node, l);
        }
        if (lv > top.load(std::memory_order_acquire)) {
            std::lock_guard<std::mutex> g(top_mu);
            if (lv > top.load()) {
                entry.store(node);
                top.store(lv);
            }
        }
    }
};

} // namespace

struct FvHnswGraph {
    Graph g;
};

namespace {

/* levels, slab offsets and empty lists for count nodes */
void init_graph(Graph &g, const float *vectors, uint64_t count, int dim, int M,
                int ef_construction) {
    g.X = vectors;
    g.dim = dim;
    g.n = count;
    g.M = (uint32_t)M;
    g.M0 = 2 * (uint32_t)M;
    g.efc = (uint32_t)ef_construction;
    const double mL = 1.0 / log((double)M);
    g.level.resize(count);
    g.uoff.resize(count + 1);
    uint64_t slots = 0;
    for (uint64_t i = 0; i < count; i++) {
        g.level[i] = (uint8_t)node_level(i, mL);
        g.uoff[i] = slots;
        slots += (uint64_t)g.level[i] * (g.M + 1);
    }
    g.uoff[count] = slots;
    g.up.assign(slots, 0);
    g.l0.assign(count * (uint64_t)(g.M0 + 1), 0);
    g.lk.reset(new SpinLock[count]);
}

/* nodes[0..n) inserted by T threads, each taking the next index */
void insert_all(Graph &g, const uint32_t *nodes, uint64_t n, int T) {
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        std::vector<uint32_t> tag(g.n, 0);
        uint32_t stamp = 0;
        for (;;) {
            const uint64_t i = next.fetch_add(1, std::memory_order_relaxed);
            if (i >= n) break;
            g.insert(nodes[i], tag, stamp);
        }
    };
    if (T <= 1 || n < 2) {
        work();
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < T; t++) pool.emplace_back(work);
    for (auto &th : pool) th.join();
}

bool bad_build_args(const float *vectors, uint64_t count, int dim, int M, int ef_construction,
                    char *errbuf, size_t errlen) {
    if (vectors == nullptr || count == 0 || count >= 0xFFFFFFFFull || dim <= 0 || M < 2 ||
        M > 32 || ef_construction < 1) {
        snprintf(errbuf, errlen, "hnsw: bad build arguments (count=%llu dim=%d M=%d efc=%d)",
                 (unsigned long long)count, dim, M, ef_construction);
        return true;
    }
    return false;
}

} // namespace

extern "C" FvHnswGraph *fv_hnsw_build(const float *vectors, uint64_t count, int dim, int M,
                                      int ef_construction, int threads, char *errbuf,
                                      size_t errlen) {
    if (bad_build_args(vectors, count, dim, M, ef_construction, errbuf, errlen)) return nullptr;
    try {
        std::unique_ptr<FvHnswGraph> h(new FvHnswGraph());
        Graph &g = h->g;
        init_graph(g, vectors, count, dim, M, ef_construction);
        {
            std::vector<uint32_t> tag(count, 0);
            uint32_t stamp = 0;
            g.insert(0, tag, stamp);
        }
        std::vector<uint32_t> rest(count - 1);
        for (uint64_t i = 1; i < count; i++) rest[i - 1] = (uint32_t)i;
        insert_all(g, rest.data(), rest.size(), std::max(1, threads));
        g.X = nullptr; /* the caller's array is not kept */
        return h.release();
    } catch (...) {
        snprintf(errbuf, errlen, "hnsw: out of memory");
        return nullptr;
    }
}

namespace pv {

/* Host phase of the device builder: levels of every node, then node 0 and
 * every node with a level above 0 inserted with the reference semantics in
 * index order (their upper levels are then complete: the device phase only
 * inserts level-0 nodes), then the level-0 entry point of every remaining
 * node by greedy descent from the top. rest lists the remaining nodes in
 * index order. */
FvHnswGraph *hnsw_seed(const float *vectors, uint64_t count, int dim, int M,
                       int ef_construction, int threads, std::vector<uint32_t> &rest,
                       std::vector<uint32_t> &entries, char *errbuf, size_t errlen,
                       double *t_entries_s) {
    if (bad_build_args(vectors, count, dim, M, ef_construction, errbuf, errlen)) return nullptr;
    try {
        std::unique_ptr<FvHnswGraph> h(new FvHnswGraph());
        Graph &g = h->g;
        init_graph(g, vectors, count, dim, M, ef_construction);
        {
            std::vector<uint32_t> tag(count, 0);
            uint32_t stamp = 0;
            g.insert(0, tag, stamp);
        }
        std::vector<uint32_t> up;
        rest.clear();
        for (uint64_t i = 1; i < count; i++) (g.level[i] > 0 ? up : rest).push_back((uint32_t)i);
        const int T = std::max(1, threads);
        insert_all(g, up.data(), up.size(), T);
        const auto te = std::chrono::steady_clock::now();
        entries.assign(rest.size(), 0);
        const int top = std::max(0, g.top.load());
        const uint64_t ep0 = g.entry.load();
        std::atomic<uint64_t> next{0};
        auto work = [&]() {
            for (;;) {
                const uint64_t i0 = next.fetch_add(1024, std::memory_order_relaxed);
                if (i0 >= rest.size()) break;
                const uint64_t i1 = std::min<uint64_t>(rest.size(), i0 + 1024);
                for (uint64_t i = i0; i < i1; i++) {
                    uint64_t ep = ep0;
                    for (int l = top; l >= 1; l--) ep = g.descend(g.row(rest[i]), ep, l);
                    entries[i] = (uint32_t)ep;
                }
            }
        };
        if (T <= 1 || rest.size() < 4096) {
            work();
        } else {
            std::vector<std::thread> pool;
            for (int t = 0; t < T; t++) pool.emplace_back(work);
            for (auto &th : pool) th.join();
        }
        if (t_entries_s)
            *t_entries_s =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - te).count();
        return h.release();
    } catch (...) {
        snprintf(errbuf, errlen, "hnsw: out of memory");
        return nullptr;
    }
}

/* the graph's level-0 lists [count][M0 + 1] ([cnt, ids...]) */
uint32_t *hnsw_l0(FvHnswGraph *h, uint32_t *stride) {
    *stride = h->g.M0 + 1;
    return h->g.l0.data();
}

/* forget the caller's vector array (the graph no longer reads it) */
void hnsw_detach(FvHnswGraph *h) { h->g.X = nullptr; }

} // namespace pv

extern "C" FvHnswGraph *fv_hnsw_build_seeded(const float *vectors, uint64_t count, int dim,
                                             int M, int ef_construction, int threads,
                                             int rest_on_host, char *errbuf, size_t errlen) {
    std::vector<uint32_t> rest, entries;
    FvHnswGraph *h = pv::hnsw_seed(vectors, count, dim, M, ef_construction, threads, rest,
                                   entries, errbuf, errlen);
    if (h == nullptr) return nullptr;
    if (rest_on_host) {
        try {
            insert_all(h->g, rest.data(), rest.size(), std::max(1, threads));
        } catch (...) {
            delete h;
            snprintf(errbuf, errlen, "hnsw: out of memory");
            return nullptr;
        }
    }
    h->g.X = nullptr;
    return h;
}

extern "C" int fv_hnsw_info(const FvHnswGraph *h, uint32_t *M, uint32_t *l0_stride,
                            uint32_t *max_level, uint64_t *entry_node, uint64_t *upper_slots) {
    if (h == nullptr) return -1;
    if (M) *M = h->g.M;
    if (l0_stride) *l0_stride = h->g.M0 + 1;
    if (max_level) *max_level = (uint32_t)std::max(0, h->g.top.load());
    if (entry_node) *entry_node = h->g.entry.load();
    if (upper_slots) *upper_slots = h->g.uoff.empty() ? 0 : h->g.uoff.back();
    return 0;
}

extern "C" int fv_hnsw_export(const FvHnswGraph *h, uint32_t *l0, uint8_t *levels,
                              uint64_t *upper_offs, uint32_t *upper) {
    if (h == nullptr) return -1;
    const Graph &g = h->g;
    if (l0) memcpy(l0, g.l0.data(), g.l0.size() * 4);
    if (levels) memcpy(levels, g.level.data(), g.level.size());
    if (upper_offs) memcpy(upper_offs, g.uoff.data(), g.uoff.size() * 8);
    if (upper) memcpy(upper, g.up.data(), g.up.size() * 4);
    return 0;
}

extern "C" void fv_hnsw_free(FvHnswGraph *h) { delete h; }

/* greedy descent through levels max_level..1 from entry_node
 * (fv_index_graph_entry, engine.cpp:1539-1559), for nq queries */
extern "C" int fv_hnsw_entries(const float *vectors, int dim, const uint64_t *upper_offs,
                               const uint32_t *upper, uint32_t M, uint64_t entry_node,
                               uint32_t max_level, int nq, const float *queries,
                               uint32_t *out, int threads) {
    if (vectors == nullptr || queries == nullptr || out == nullptr || nq < 0 || dim <= 0)
        return -1;
    if (max_level > 0 && (upper_offs == nullptr || upper == nullptr)) return -1;
    auto one = [&](int qi) {
        const float *q = queries + (uint64_t)qi * dim;
        uint64_t ep = entry_node;
        for (int lev = (int)max_level; lev >= 1; lev--) {
            float best = l2(q, vectors + ep * (uint64_t)dim, dim);
            for (bool moved = true; moved;) {
                moved = false;
                const uint32_t *s = upper + upper_offs[ep] + (uint64_t)(lev - 1) * (M + 1);
                const uint32_t c = s[0];
                for (uint32_t i = 0; i < c; i++) {
                    const float d = l2(q, vectors + (uint64_t)s[1 + i] * dim, dim);
                    if (d < best) {
                        best = d;
                        ep = s[1 + i];
                        moved = true;
                    }
                }
            }
        }
        out[qi] = (uint32_t)ep;
    };
    const int T = std::max(1, std::min(threads, nq));
    if (T <= 1) {
        for (int i = 0; i < nq; i++) one(i);
        return 0;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < T; t++)
        pool.emplace_back([&]() {
            for (int i; (i = next.fetch_add(1)) < nq;) one(i);
        });
    for (auto &th : pool) th.join();
    return 0;
}

/* exact FP32 distance in the reference order (test and tool use) */
extern "C" float fv_l2sq(const float *a, const float *b, int dim) { return l2(a, b, dim); }
