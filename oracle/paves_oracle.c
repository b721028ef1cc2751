/*
 * paves_oracle.c — CPU restatement of PAVES's query arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP engine in
 * paves_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it. The product never links or calls it.
 *
 * Each function cites the reference code it restates (paths relative to the
 * PAVES source tree). Built with -ffp-contract=off so every rounding step is
 * the one written here (fused multiply-adds are explicit fmaf calls).
 *
 * Pinning (see tests/test_oracle.py):
 *   or_l2sq          — bit-identical to the compiled reference l2sq (via
 *                      oracle/_ref) on random non-integer data; exact against
 *                      FP64 on integer data (closed form).
 *   or_pred_eval     — known-answer tests (NaN/NULL, NE, IN, AND/OR/NOT) and
 *                      the compiled reference's filtered cursors.
 *   or_brute_topb    — compiled reference BruteCursor (same ids, distances)
 *                      and an FP64 brute force on tiny inputs.
 *   or_hnsw_search   — compiled reference HnswCursor on the same graph file.
 *   or_graph_entry   — compiled reference fv_index_graph_entry.
 *   or_hnsw_batch_insert — parity unpinned as a whole (the reference has no
 *                      batched builder): its search is or_hnsw_search, its
 *                      heuristic is checked against brute force on tiny
 *                      inputs and the host builder's reference-identical
 *                      graph reduces to it with batches of one node
 *                      (tests/test_oracle_build.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "paves_oracle.h"

/* ---------------------------------------------------------------------
 * Distance. engine.cpp:43-63 (AVX-512F branch): two 16-lane accumulators
 * over 32-dim blocks (d = a - b, acc = fma(d, d, acc)), one more 16-lane
 * block into acc0, then _mm512_reduce_add_ps(acc0 + acc1) — GCC's tree:
 * 8 adds (hi half + lo half), 4 adds, 2 adds, 1 add — then a scalar tail
 * sum = fma(d, d, sum) (GCC contracts the reference loop at -O3).
 * ------------------------------------------------------------------- */
float or_l2sq(const float *a, const float *b, int dim) {
    float acc0[16], acc1[16];
    int i = 0, j;
    for (j = 0; j < 16; j++) acc0[j] = acc1[j] = 0.f;
    for (; i + 32 <= dim; i += 32) {
        for (j = 0; j < 16; j++) {
            float d0 = a[i + j] - b[i + j];
            float d1 = a[i + 16 + j] - b[i + 16 + j];
            acc0[j] = fmaf(d0, d0, acc0[j]);
            acc1[j] = fmaf(d1, d1, acc1[j]);
        }
    }
    for (; i + 16 <= dim; i += 16) {
        for (j = 0; j < 16; j++) {
            float d = a[i + j] - b[i + j];
            acc0[j] = fmaf(d, d, acc0[j]);
        }
    }
    float t[16], t8[8], t4[4];
    for (j = 0; j < 16; j++) t[j] = acc0[j] + acc1[j];
    for (j = 0; j < 8; j++) t8[j] = t[8 + j] + t[j];
    for (j = 0; j < 4; j++) t4[j] = t8[4 + j] + t8[j];
    float s0 = t4[0] + t4[2];
    float s1 = t4[1] + t4[3];
    float sum = s0 + s1;
    /* The reference's scalar tail `sum += d * d` as its compiler emits it
     * (GCC 11, -O3, AVX-512 target): when >= 8 elements remain, one 8-wide
     * in-order (fold-left) block — products rounded separately, added in
     * order — then a fused multiply-add per remaining element (the
     * scalar epilogue is contracted). Pinned against oracle/_ref for tail
     * lengths 1..15. */
    if (dim - i >= 8) {
        for (j = 0; j < 8; j++, i++) {
            float d = a[i] - b[i];
            float p = d * d;
            sum = sum + p;
        }
    }
    for (; i < dim; i++) {
        float d = a[i] - b[i];
        sum = fmaf(d, d, sum);
    }
    return sum;
}

/* FP64 recomputation (the stated reference for distance tolerances). */
double or_l2sq_f64(const float *a, const float *b, int dim) {
    double s = 0.0;
    for (int i = 0; i < dim; i++) {
        double d = (double)a[i] - (double)b[i];
        s += d * d;
    }
    return s;
}

/* Round-to-nearest-even float -> bfloat16 bits (NaN kept quiet). */
uint16_t or_bf16_bits(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

/* ---------------------------------------------------------------------
 * Predicate. engine.cpp:397-430 (FvPred::eval_node) over the serialized
 * arrays of engine_api.h:95-117 (as interpreted by gpu_engine.cu:238-279).
 * NULL is stored as NaN: every comparison with NaN is false; NE is
 * v == v && v != c; IN is true when v equals any listed value.
 * ------------------------------------------------------------------- */
int or_pred_eval(const OrPredNode *nodes, const int32_t *children,
                 const double *invals, const double *cols, uint64_t count,
                 int32_t ni, uint32_t row) {
    const OrPredNode *n = &nodes[ni];
    switch (n->kind) {
        case OR_NODE_CMP: {
            double v = cols[(uint64_t)n->col * count + row];
            switch (n->op) {
                case OR_CMP_LT: return v < n->value;
                case OR_CMP_LE: return v <= n->value;
                case OR_CMP_EQ: return v == n->value;
                case OR_CMP_GE: return v >= n->value;
                case OR_CMP_GT: return v > n->value;
                case OR_CMP_NE: return v == v && v != n->value;
            }
            return 0;
        }
        case OR_NODE_IN: {
            double v = cols[(uint64_t)n->col * count + row];
            for (int i = 0; i < n->nvalues; i++)
                if (v == invals[n->val0 + i]) return 1;
            return 0;
        }
        case OR_NODE_AND:
            for (int i = 0; i < n->nchildren; i++)
                if (!or_pred_eval(nodes, children, invals, cols, count,
                                  children[n->child0 + i], row))
                    return 0;
            return 1;
        case OR_NODE_OR:
            for (int i = 0; i < n->nchildren; i++)
                if (or_pred_eval(nodes, children, invals, cols, count,
                                 children[n->child0 + i], row))
                    return 1;
            return 0;
        case OR_NODE_NOT:
            return !or_pred_eval(nodes, children, invals, cols, count,
                                 children[n->child0], row);
    }
    return 0;
}

/* ---------------------------------------------------------------------
 * Exact filtered top-B. Order is (dist, row) lexicographic — the order of
 * BruteCursor's final sort (engine.cpp:1114-1116) and of the GPU packed
 * keys (gpu_engine.cu:435-438). A max-heap on (dist, row) over rows in
 * increasing order keeps exactly the B smallest pairs. out_cnt = min(B,
 * passing); truncated = (out_cnt == B), as the reference's tiled GPU path
 * reports it (gpu_engine.cu:1059-1064).
 * ------------------------------------------------------------------- */
static inline int pair_less(float da, uint64_t na, float db, uint64_t nb) {
    return da < db || (da == db && na < nb);
}

static void heap_sift_up(OrHit *h, uint32_t i) {
    while (i) {
        uint32_t p = (i - 1) / 2;
        if (!pair_less(h[p].dist, h[p].node, h[i].dist, h[i].node)) break;
        OrHit t = h[p]; h[p] = h[i]; h[i] = t;
        i = p;
    }
}

static void heap_sift_down(OrHit *h, uint32_t n) {
    uint32_t i = 0;
    for (;;) {
        uint32_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < n && pair_less(h[m].dist, h[m].node, h[l].dist, h[l].node)) m = l;
        if (r < n && pair_less(h[m].dist, h[m].node, h[r].dist, h[r].node)) m = r;
        if (m == i) break;
        OrHit t = h[i]; h[i] = h[m]; h[m] = t;
        i = m;
    }
}

static int hit_cmp(const void *pa, const void *pb) {
    const OrHit *a = (const OrHit *)pa, *b = (const OrHit *)pb;
    if (pair_less(a->dist, a->node, b->dist, b->node)) return -1;
    if (pair_less(b->dist, b->node, a->dist, a->node)) return 1;
    return 0;
}

void or_sort_hits(OrHit *h, uint32_t n) { qsort(h, n, sizeof(OrHit), hit_cmp); }

uint64_t or_brute_topb(const float *vectors, uint64_t count, int dim,
                       const double *cols, const float *query,
                       const OrPredNode *nodes, const int32_t *children,
                       const double *invals, int32_t proot, uint32_t topb,
                       OrHit *out, uint32_t *out_cnt, uint8_t *truncated) {
    uint32_t n = 0;
    uint64_t passing = 0;
    for (uint64_t r = 0; r < count; r++) {
        if (proot >= 0 &&
            !or_pred_eval(nodes, children, invals, cols, count, proot, (uint32_t)r))
            continue;
        passing++;
        float d = or_l2sq(query, vectors + r * (uint64_t)dim, dim);
        if (n < topb) {
            out[n].node = r;
            out[n].dist = d;
            heap_sift_up(out, n);
            n++;
        } else if (pair_less(d, r, out[0].dist, out[0].node)) {
            out[0].node = r;
            out[0].dist = d;
            heap_sift_down(out, n);
        }
    }
    or_sort_hits(out, n);
    *out_cnt = n;
    *truncated = (n == topb) ? 1 : 0;
    return passing;
}

/* Batch form over the ABI's per-query descriptors (gpu_engine.h:52-60),
 * parallel over queries. Returns 0. */
int or_brute_batch(const float *vectors, uint64_t count, int dim,
                   const double *cols, int nq, const float *queries,
                   const int32_t *proots, const OrPredNode *nodes,
                   const int32_t *children, const double *invals,
                   uint32_t topb, OrHit *out, uint32_t *out_cnt,
                   uint8_t *truncated, uint64_t *passing) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < nq; q++) {
        uint64_t p = or_brute_topb(vectors, count, dim, cols,
                                   queries + (uint64_t)q * dim, nodes, children,
                                   invals, proots[q], topb,
                                   out + (uint64_t)q * topb, &out_cnt[q],
                                   &truncated[q]);
        if (passing) passing[q] = p;
    }
    return 0;
}

/* ---------------------------------------------------------------------
 * SQ8 over arbitrary float data: the scan ranks rows by the distance of
 * int8 codes and the top overfetch x B are re-ranked exactly — the role of
 * the reference's SQ8 tiled path (quantize, int8 scan to 2 topb candidates,
 * re-rank: gpu_engine.cu:643-760, 1911-2080). This library's quantizer:
 * per-dim midrange centre mid[d] = (lo[d] + hi[d]) * 0.5f and one scale
 * inv = 127 / max_d (hi[d] - lo[d]) * 0.5f (1 when every dim is constant);
 * code = clamp(rint((x - mid) * inv), -127, 127), the same map for rows and
 * queries. Code distance |xc|^2 + |qc|^2 - 2 xc.qc in integers, as FP32;
 * the R = min(over x B, 4096, passing) passing rows smallest by (code
 * distance, row) are re-ranked with or_l2sq and sorted by (dist, row); the
 * first min(B, passing) are the result, truncated = (count == B).
 * ------------------------------------------------------------------- */
/* (code distance, row) order of the FP8 candidate sort */
static int dr_cmp(const void *a, const void *b) {
    const double da = *(const double *)a, db = *(const double *)b;
    const uint64_t ra = *(const uint64_t *)((const char *)a + 8);
    const uint64_t rb = *(const uint64_t *)((const char *)b + 8);
    if (da != db) return da < db ? -1 : 1;
    return ra < rb ? -1 : (ra > rb ? 1 : 0);
}

/* per-dim midrange centres, and the largest half-range */
static float midrange_hmax(const float *vectors, uint64_t count, int dim, float *mid) {
    float hmax = 0.f;
    for (int d = 0; d < dim; d++) {
        float lo = INFINITY, hi = -INFINITY;
        for (uint64_t r = 0; r < count; r++) {
            const float x = vectors[r * (uint64_t)dim + d];
            lo = x < lo ? x : lo;
            hi = x > hi ? x : hi;
        }
        mid[d] = (lo + hi) * 0.5f;
        const float h = (hi - lo) * 0.5f;
        hmax = h > hmax ? h : hmax;
    }
    return hmax;
}

void or_sq8_params(const float *vectors, uint64_t count, int dim, float *mid, float *inv) {
    const float hmax = midrange_hmax(vectors, count, dim, mid);
    *inv = hmax > 0.f ? 127.0f / hmax : 1.0f;
}

void or_sq8_codes(const float *x, uint64_t n, int dim, const float *mid, float inv,
                  int8_t *codes) {
    for (uint64_t r = 0; r < n; r++)
        for (int d = 0; d < dim; d++) {
            float c = rintf((x[r * (uint64_t)dim + d] - mid[d]) * inv);
            c = c < -127.f ? -127.f : (c > 127.f ? 127.f : c);
            codes[r * (uint64_t)dim + d] = (int8_t)(int)c;
        }
}

int or_sq8_brute_batch(const float *vectors, uint64_t count, int dim, const double *cols,
                       int nq, const float *queries, const int32_t *proots,
                       const OrPredNode *nodes, const int32_t *children,
                       const double *invals, uint32_t topb, uint32_t over, OrHit *out,
                       uint32_t *out_cnt, uint8_t *truncated) {
    float *mid = (float *)malloc(sizeof(float) * (size_t)dim);
    int8_t *xc = (int8_t *)malloc((size_t)count * dim);
    int32_t *xn = (int32_t *)malloc(sizeof(int32_t) * (size_t)count);
    if (!mid || !xc || !xn) {
        free(mid); free(xc); free(xn);
        return -1;
    }
    float inv;
    or_sq8_params(vectors, count, dim, mid, &inv);
    or_sq8_codes(vectors, count, dim, mid, inv, xc);
    for (uint64_t r = 0; r < count; r++) {
        int32_t s = 0;
        for (int d = 0; d < dim; d++) s += (int32_t)xc[r * (uint64_t)dim + d] * xc[r * (uint64_t)dim + d];
        xn[r] = s;
    }
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < nq; q++) {
        int8_t *qc = (int8_t *)malloc((size_t)dim);
        OrHit *cand = (OrHit *)malloc(sizeof(OrHit) * ((size_t)count + 1));
        if (!qc || !cand) {
            free(qc); free(cand);
            rc = -1;
            continue;
        }
        const float *qv = queries + (uint64_t)q * dim;
        or_sq8_codes(qv, 1, dim, mid, inv, qc);
        int32_t qn = 0;
        for (int d = 0; d < dim; d++) qn += (int32_t)qc[d] * qc[d];
        uint64_t n = 0;
        for (uint64_t r = 0; r < count; r++) {
            if (proots[q] >= 0 &&
                !or_pred_eval(nodes, children, invals, cols, count, proots[q], (uint32_t)r))
                continue;
            const int8_t *x = xc + r * (uint64_t)dim;
            int32_t dot = 0;
            for (int d = 0; d < dim; d++) dot += (int32_t)x[d] * qc[d];
            cand[n].node = r;
            cand[n].dist = (float)(xn[r] + qn - 2 * dot);
            n++;
        }
        qsort(cand, n, sizeof(OrHit), hit_cmp);
        uint64_t R = (uint64_t)topb * over;
        if (R > 4096) R = 4096;
        if (R > n) R = n;
        for (uint64_t i = 0; i < R; i++)
            cand[i].dist = or_l2sq(qv, vectors + cand[i].node * (uint64_t)dim, dim);
        qsort(cand, R, sizeof(OrHit), hit_cmp);
        const uint32_t B = (uint32_t)(n < topb ? n : topb);
        memcpy(out + (uint64_t)q * topb, cand, B * sizeof(OrHit));
        out_cnt[q] = B;
        truncated[q] = (B == topb) ? 1 : 0;
        free(qc);
        free(cand);
    }
    free(mid); free(xc); free(xn);
    return rc;
}

/* ---------------------------------------------------------------------
 * FP8 index (this library's reduced-precision index beside SQ8, in the
 * reference SQ8's role, gpu_engine.cu:643-760 and 1911-2080): the rows'
 * midrange-centred values under one scale (the largest half-range maps to
 * 448) as FP8 E4M3 codes (OCP e4m3fn: bias 7, 3 mantissa bits, largest
 * finite 448, subnormals m x 2^-9), ranked by the code distance and the top
 * over x topb re-ranked with the exact FP32 distance.
 *
 * Pinning (tests/test_oracle_fp8.py): encoder known answers worked by hand
 * (ties to even, carries into the next binade, subnormals, saturation);
 * decode of all 256 codes against the format's definition; encode equals
 * the nearest code by exhaustive search; with over x topb >= the passing
 * rows the batch equals the exact brute force bit for bit.
 * ------------------------------------------------------------------- */
uint8_t or_fp8_encode(float v) {
    const float a = fabsf(v);
    const uint32_t s = v < 0.f ? 0x80u : 0u;
    if (!(a < 448.f)) return (uint8_t)(s | 0x7Eu);
    if (a < 0.015625f) {                            /* below 2^-6: m x 2^-9 */
        const uint32_t m = (uint32_t)rintf(a * 512.f);
        return (uint8_t)(m == 0 ? 0u : (s | m));
    }
    int e;
    (void)frexpf(a, &e);                            /* a in [2^(e-1), 2^e) */
    const float frac = ldexpf(a, 1 - e) - 1.f;
    uint32_t m = (uint32_t)rintf(frac * 8.f);
    int E = e - 1 + 7;
    if (m == 8u) {
        m = 0u;
        E++;
    }
    return (uint8_t)(s | ((uint32_t)E << 3) | m);
}

float or_fp8_decode(uint8_t c) {
    const int E = (c >> 3) & 15, m = c & 7;
    const float a = E == 0 ? ldexpf((float)m, -9) : ldexpf(1.f + (float)m * 0.125f, E - 7);
    return (c & 0x80u) ? -a : a;
}

/* SQ8's per-dim midrange and largest half-range, mapped to 448 */
void or_fp8_params(const float *vectors, uint64_t count, int dim, float *mid, float *scale) {
    const float hmax = midrange_hmax(vectors, count, dim, mid);
    *scale = hmax > 0.f ? 448.0f / hmax : 1.0f;
}

void or_fp8_codes(const float *x, uint64_t n, int dim, const float *mid, float scale,
                  uint8_t *codes) {
    for (uint64_t r = 0; r < n; r++)
        for (int d = 0; d < dim; d++)
            codes[r * (uint64_t)dim + d] = or_fp8_encode((x[r * (uint64_t)dim + d] - mid[d]) * scale);
}

int or_fp8_brute_batch(const float *vectors, uint64_t count, int dim, const double *cols,
                       int nq, const float *queries, const int32_t *proots,
                       const OrPredNode *nodes, const int32_t *children,
                       const double *invals, uint32_t topb, uint32_t over, OrHit *out,
                       uint32_t *out_cnt, uint8_t *truncated) {
    float *mid = (float *)malloc(sizeof(float) * (size_t)dim);
    double *xd = (double *)malloc(sizeof(double) * (size_t)count * dim);
    double *xn = (double *)malloc(sizeof(double) * (size_t)count);
    if (!mid || !xd || !xn) {
        free(mid); free(xd); free(xn);
        return -1;
    }
    float scale;
    or_fp8_params(vectors, count, dim, mid, &scale);
    /* decoded codes; code distances in FP64 are exact (every product of two
     * E4M3 values is a multiple of 2^-18 below 2^18, a sum of a few
     * thousand of them fits in 53 bits) */
    for (uint64_t r = 0; r < count; r++) {
        double s = 0.0;
        for (int d = 0; d < dim; d++) {
            const double v = or_fp8_decode(or_fp8_encode(
                (vectors[r * (uint64_t)dim + d] - mid[d]) * scale));
            xd[r * (uint64_t)dim + d] = v;
            s += v * v;
        }
        xn[r] = s;
    }
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < nq; q++) {
        double *qc = (double *)malloc(sizeof(double) * (size_t)dim);
        OrHit *cand = (OrHit *)malloc(sizeof(OrHit) * ((size_t)count + 1));
        double *cd = (double *)malloc(sizeof(double) * ((size_t)count + 1));
        if (!qc || !cand || !cd) {
            free(qc); free(cand); free(cd);
            rc = -1;
            continue;
        }
        const float *qv = queries + (uint64_t)q * dim;
        double qn = 0.0;
        for (int d = 0; d < dim; d++) {
            qc[d] = or_fp8_decode(or_fp8_encode((qv[d] - mid[d]) * scale));
            qn += qc[d] * qc[d];
        }
        uint64_t n = 0;
        for (uint64_t r = 0; r < count; r++) {
            if (proots[q] >= 0 &&
                !or_pred_eval(nodes, children, invals, cols, count, proots[q], (uint32_t)r))
                continue;
            const double *x = xd + r * (uint64_t)dim;
            double dot = 0.0;
            for (int d = 0; d < dim; d++) dot += x[d] * qc[d];
            cd[n] = xn[r] + qn - 2.0 * dot;
            cand[n].node = r;
            n++;
        }
        /* top R by (code distance, row) */
        typedef struct { double d; uint64_t r; } DR;
        DR *dr = (DR *)malloc(sizeof(DR) * ((size_t)n + 1));
        if (!dr) {
            free(qc); free(cand); free(cd);
            rc = -1;
            continue;
        }
        for (uint64_t i = 0; i < n; i++) {
            dr[i].d = cd[i];
            dr[i].r = cand[i].node;
        }
        qsort(dr, n, sizeof(DR), dr_cmp);
        uint64_t R = (uint64_t)topb * over;
        if (R > 4096) R = 4096;
        if (R > n) R = n;
        for (uint64_t i = 0; i < R; i++) {
            cand[i].node = dr[i].r;
            cand[i].dist = or_l2sq(qv, vectors + dr[i].r * (uint64_t)dim, dim);
        }
        qsort(cand, R, sizeof(OrHit), hit_cmp);
        const uint32_t B = (uint32_t)(n < topb ? n : topb);
        memcpy(out + (uint64_t)q * topb, cand, B * sizeof(OrHit));
        out_cnt[q] = B;
        truncated[q] = (B == topb) ? 1 : 0;
        free(dr);
        free(qc);
        free(cand);
        free(cd);
    }
    free(mid); free(xd); free(xn);
    return rc;
}

/* ---------------------------------------------------------------------
 * Filtered HNSW level-0 search: HnswCursor::start + run_phase
 * (engine.cpp:1631-1712). cand is a min-heap on (dist, node), res a
 * max-heap on (dist, node) holding predicate-passing nodes only, capacity
 * ef. Stop when the best candidate is worse than the worst result with res
 * full, or when visited >= max_visited (exhausted), or when cand drains
 * (exhausted). Neighbours are admitted to cand when res is not full or
 * they beat the worst result; only passing ones enter res. The graph is
 * the reference's level-0 layout: l0[node * stride] = count, followed by
 * the neighbour ids.
 * ------------------------------------------------------------------- */
typedef struct { float d; uint32_t n; } Pair;

static inline int p_less(Pair a, Pair b) {
    return a.d < b.d || (a.d == b.d && a.n < b.n);
}

/* heap with comparator sign: minheap (sign=+1 -> smallest on top) */
static void ph_push(Pair *h, uint64_t *n, Pair v, int maxheap) {
    uint64_t i = (*n)++;
    h[i] = v;
    while (i) {
        uint64_t p = (i - 1) / 2;
        int better = maxheap ? p_less(h[p], h[i]) : p_less(h[i], h[p]);
        if (!better) break;
        Pair t = h[p]; h[p] = h[i]; h[i] = t;
        i = p;
    }
}

static Pair ph_pop(Pair *h, uint64_t *n, int maxheap) {
    Pair top = h[0];
    uint64_t m = --(*n);
    h[0] = h[m];
    uint64_t i = 0;
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, b = i;
        if (l < m && (maxheap ? p_less(h[b], h[l]) : p_less(h[l], h[b]))) b = l;
        if (r < m && (maxheap ? p_less(h[b], h[r]) : p_less(h[r], h[b]))) b = r;
        if (b == i) break;
        Pair t = h[i]; h[i] = h[b]; h[b] = t;
        i = b;
    }
    return top;
}

uint32_t or_hnsw_search(const float *vectors, uint64_t count, int dim,
                        const uint32_t *l0, uint32_t l0_stride,
                        const double *cols, const OrPredNode *nodes,
                        const int32_t *children, const double *invals,
                        int32_t proot, const float *query, uint32_t entry,
                        uint32_t ef, uint64_t max_visited, OrHit *out,
                        uint8_t *exhausted, uint64_t *visited_out) {
    uint8_t *seen = (uint8_t *)calloc(count, 1);
    Pair *cand = (Pair *)malloc(sizeof(Pair) * (count + 1));
    Pair *res = (Pair *)malloc(sizeof(Pair) * (ef + 2));
    uint64_t nc = 0, nr = 0, visited = 1;
    int exh = 0;

    float epd = or_l2sq(query, vectors + (uint64_t)entry * dim, dim);
    seen[entry] = 1;
    ph_push(cand, &nc, (Pair){epd, entry}, 0);
    if (proot < 0 || or_pred_eval(nodes, children, invals, cols, count, proot, entry))
        ph_push(res, &nr, (Pair){epd, entry}, 1);

    while (nc > 0) {
        Pair c = cand[0];
        if (nr >= ef && c.d > res[0].d) break;
        if (visited >= max_visited) {
            exh = 1;
            break;
        }
        ph_pop(cand, &nc, 0);
        const uint32_t *lst = l0 + (uint64_t)c.n * l0_stride;
        uint32_t cnt = lst[0];
        for (uint32_t i = 0; i < cnt; i++) {
            uint32_t n2 = lst[1 + i];
            if (seen[n2]) continue;
            seen[n2] = 1;
            visited++;
            float d = or_l2sq(query, vectors + (uint64_t)n2 * dim, dim);
            if (nr < ef || d < res[0].d) {
                ph_push(cand, &nc, (Pair){d, n2}, 0);
                if (proot < 0 ||
                    or_pred_eval(nodes, children, invals, cols, count, proot, n2)) {
                    ph_push(res, &nr, (Pair){d, n2}, 1);
                    if (nr > ef) ph_pop(res, &nr, 1);
                }
            }
        }
    }
    if (nc == 0) exh = 1;
    uint32_t k = (uint32_t)nr;
    for (uint32_t i = 0; i < k; i++) {
        out[i].node = res[i].n;
        out[i].dist = res[i].d;
    }
    or_sort_hits(out, k);
    *exhausted = (uint8_t)exh;
    if (visited_out) *visited_out = visited;
    free(seen);
    free(cand);
    free(res);
    return k;
}

/* Parallel batch over queries (one search per query). */
int or_hnsw_batch(const float *vectors, uint64_t count, int dim,
                  const uint32_t *l0, uint32_t l0_stride, const double *cols,
                  const OrPredNode *nodes, const int32_t *children,
                  const double *invals, int nq, const int32_t *proots,
                  const float *queries, const uint32_t *entries, uint32_t ef,
                  uint64_t max_visited, OrHit *out, uint32_t *out_cnt,
                  uint8_t *exhausted) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < nq; q++)
        out_cnt[q] = or_hnsw_search(vectors, count, dim, l0, l0_stride, cols,
                                    nodes, children, invals, proots[q],
                                    queries + (uint64_t)q * dim, entries[q], ef,
                                    max_visited, out + (uint64_t)q * ef,
                                    &exhausted[q], NULL);
    return 0;
}

/* ---------------------------------------------------------------------
 * Level-0 phase of the batched (device) HNSW builder. The pieces are the
 * reference builder's: the efConstruction beam (HnswBuild's search,
 * engine.cpp:760-805, whose pop/admit/stop rules are the query cursor's
 * restated in or_hnsw_search, here unfiltered), the neighbour heuristic
 * (engine.cpp:807-823: keep a candidate unless a kept node is closer to it
 * than the base node; stop at m; no pruning at all when there are no more
 * than m candidates) and the re-pruning of full lists on reverse links
 * (engine.cpp:825-845). The schedule is the batch one: every node of a
 * batch searches the level-0 graph as it stood before the batch; each
 * keeps M links; then each target takes the batch's incoming links in node
 * order — appended while its count plus the links left fit in M0, else its
 * list and the next min(left, 64 - count) links are sorted by (distance to
 * the target, node) and re-pruned to M0, until none are left.
 * l0 [count][2M+1] ([cnt, ids...]) holds the host phase's lists on entry
 * and the result on return.
 * ------------------------------------------------------------------- */
static int pair_cmp(const void *a, const void *b) {
    const Pair x = *(const Pair *)a, y = *(const Pair *)b;
    return p_less(x, y) ? -1 : (p_less(y, x) ? 1 : 0);
}

static int u64_cmp(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* the heuristic over candidates c[0..n) sorted ascending; kept ids -> keep */
static uint32_t or_heuristic(const float *X, int dim, const Pair *c, uint32_t n,
                             uint32_t m, uint32_t *keep) {
    uint32_t nk = 0;
    for (uint32_t i = 0; i < n && nk < m; i++) {
        int ok = 1;
        for (uint32_t k = 0; k < nk && ok; k++)
            if (or_l2sq(X + (uint64_t)c[i].n * dim, X + (uint64_t)keep[k] * dim, dim) < c[i].d)
                ok = 0;
        if (ok) keep[nk++] = c[i].n;
    }
    return nk;
}

int or_hnsw_batch_insert(const float *X, uint64_t count, int dim, uint32_t M, uint32_t efc,
                         uint32_t *l0, const uint32_t *rest, const uint32_t *entries,
                         const int32_t *sizes, int nbatches) {
    const uint32_t M0 = 2 * M, S = M0 + 1;
    int32_t bmax = 1;
    for (int b = 0; b < nbatches; b++) bmax = sizes[b] > bmax ? sizes[b] : bmax;
    OrHit *hits = (OrHit *)malloc(sizeof(OrHit) * (size_t)bmax * efc);
    uint32_t *cnt = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)bmax);
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)bmax * M);
    Pair *c = (Pair *)malloc(sizeof(Pair) * (efc > 64 ? efc : 64));
    uint32_t keep[64];
    if (!hits || !cnt || !keys || !c) {
        free(hits);
        free(cnt);
        free(keys);
        free(c);
        return -1;
    }
    uint64_t off = 0;
    for (int b = 0; b < nbatches; b++) {
        const int nb = sizes[b];
#pragma omp parallel for schedule(dynamic, 4)
        for (int i = 0; i < nb; i++) {
            uint8_t exh;
            const uint32_t u = rest[off + i];
            cnt[i] = or_hnsw_search(X, count, dim, l0, S, NULL, NULL, NULL, NULL, -1,
                                    X + (uint64_t)u * dim, entries[off + i], efc, count + 1,
                                    hits + (uint64_t)i * efc, &exh, NULL);
        }
        uint64_t nk = 0;
        for (int i = 0; i < nb; i++) {
            const uint32_t u = rest[off + i];
            const OrHit *h = hits + (uint64_t)i * efc;
            uint32_t n = 0;
            if (cnt[i] <= M) {
                for (uint32_t j = 0; j < cnt[i]; j++) keep[n++] = (uint32_t)h[j].node;
            } else {
                for (uint32_t j = 0; j < cnt[i]; j++) c[j] = (Pair){h[j].dist, (uint32_t)h[j].node};
                n = or_heuristic(X, dim, c, cnt[i], M, keep);
            }
            uint32_t *L = l0 + (uint64_t)u * S;
            L[0] = n;
            for (uint32_t j = 0; j < n; j++) {
                L[1 + j] = keep[j];
                keys[nk++] = ((uint64_t)keep[j] << 32) | u;
            }
        }
        qsort(keys, nk, sizeof(uint64_t), u64_cmp);
        for (uint64_t s = 0; s < nk;) {
            const uint32_t v = (uint32_t)(keys[s] >> 32);
            uint64_t e = s;
            while (e < nk && (uint32_t)(keys[e] >> 32) == v) e++;
            uint32_t *L = l0 + (uint64_t)v * S;
            const float *xv = X + (uint64_t)v * dim;
            uint32_t cc = L[0];
            uint64_t pos = s;
            while (pos < e) {
                if (cc + (e - pos) <= M0) {
                    while (pos < e) L[1 + cc++] = (uint32_t)keys[pos++];
                    break;
                }
                const uint32_t take = (uint32_t)((e - pos) < 64 - cc ? (e - pos) : 64 - cc);
                uint32_t n = 0;
                for (uint32_t j = 0; j < cc; j++)
                    c[n++] = (Pair){or_l2sq(xv, X + (uint64_t)L[1 + j] * dim, dim), L[1 + j]};
                for (uint32_t j = 0; j < take; j++) {
                    const uint32_t x = (uint32_t)keys[pos + j];
                    c[n++] = (Pair){or_l2sq(xv, X + (uint64_t)x * dim, dim), x};
                }
                pos += take;
                qsort(c, n, sizeof(Pair), pair_cmp);
                cc = or_heuristic(X, dim, c, n, M0, keep);
                for (uint32_t j = 0; j < cc; j++) L[1 + j] = keep[j];
            }
            L[0] = cc;
            s = e;
        }
        off += (uint64_t)nb;
    }
    free(hits);
    free(cnt);
    free(keys);
    free(c);
    return 0;
}

/* ---------------------------------------------------------------------
 * Upper-level greedy descent: fv_index_graph_entry (engine.cpp:1539-1559).
 * upper[upper_offs[node] + (level-1)*(M+1)] = count, then neighbour ids.
 * ------------------------------------------------------------------- */
uint32_t or_graph_entry(const float *vectors, int dim,
                        const uint64_t *upper_offs, const uint32_t *upper,
                        uint32_t M, uint64_t entry_node, int max_level,
                        const float *query) {
    uint64_t ep = entry_node;
    for (int lc = max_level; lc >= 1; lc--) {
        float d = or_l2sq(query, vectors + ep * (uint64_t)dim, dim);
        int changed = 1;
        while (changed) {
            changed = 0;
            const uint32_t *lst = upper + upper_offs[ep] + (uint64_t)(lc - 1) * (M + 1);
            uint32_t cnt = lst[0];
            for (uint32_t i = 0; i < cnt; i++) {
                float nd = or_l2sq(query, vectors + (uint64_t)lst[1 + i] * dim, dim);
                if (nd < d) {
                    d = nd;
                    ep = lst[1 + i];
                    changed = 1;
                }
            }
        }
    }
    return (uint32_t)ep;
}

/* recall@k of got against truth (|ids ∩ truth| / |truth|), the definition
 * of bench/benchmark.py:53-64 for one query. */
double or_recall(const uint64_t *got, uint32_t ngot, const uint64_t *truth,
                 uint32_t ntruth) {
    if (ntruth == 0) return 1.0;
    uint32_t hit = 0;
    for (uint32_t i = 0; i < ntruth; i++)
        for (uint32_t j = 0; j < ngot; j++)
            if (got[j] == truth[i]) {
                hit++;
                break;
            }
    return (double)hit / (double)ntruth;
}
