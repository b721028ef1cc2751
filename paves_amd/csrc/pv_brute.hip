/*
 * pv_brute.hip — batched exact filtered top-B on MI355X.
 *
 * Replaces the reference's brute path (gpu_engine.cu:348-418 span kernel,
 * 474-610 tiled kernel, 1003-1065 merge kernel, 2206-2336 router) with a
 * different pipeline built for CDNA4 (one stream per batch, no host round
 * trip inside a run):
 *
 *   1. k_phase1         predicate programs -> row bitmaps (bitmap_rows),
 *                       and in the same launch fp32 queries -> bf16 (or
 *                       int8) copy, ||q||^2, per-query error bound eps, and
 *                       for routed runs the router features
 *   2. k_scan<PREPASS>  distances of every stride-th 32-row tile; per
 *                       16-row group the minimum passing distance
 *   3. k_tau            tau[q] = B-th smallest group minimum (+2 eps): at
 *                       least B passing rows lie at or below it, so every
 *                       row of the exact top-B does too
 *   4. k_scan<MAIN>     full scan; rows with dist <= tau enter a per-query
 *                       candidate pool (LDS, flushed to HBM per WG into one
 *                       of the query's pool segments)
 *   5. k_select         B-th smallest pool key (wave-sorted bound + 64-lane
 *                       register sort for B <= 64, radix select otherwise),
 *                       collect the keys within 2 eps of it, re-rank them
 *                       with exact FP32 in the CPU cursor's summation order,
 *                       sort by (dist, row), emit FvHit + count + truncated
 *
 * Distances in the scans: BF16 MFMA (v_mfma_f32_32x32x16_bf16) computing
 * x.q with the ||x||^2 + ||q||^2 - 2 x.q epilogue, INT8 MFMA
 * (v_mfma_i32_32x32x32_i8) on x - c and q - c for integer data (exact), or
 * plain FP32 VALU sum (x - q)^2. Each approximates the reference l2sq value
 * within a per-query bound eps (0 when every value is an integer small
 * enough that all arithmetic is exact); the re-rank restores the exact FP32
 * value, so ids and distances equal the CPU cursor's (BruteCursor,
 * engine.cpp:1050-1130) whatever scan precision is used.
 *
 * Span-gather route: for batches whose driver spans (engine.cpp:464-599)
 * enumerate few rows, k_gather computes exact distances for the enumerated
 * rows only (the analogue of fv_brute_kernel) and k_select finishes.
 */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "pv_mlp.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));

namespace pv {



/* ---------------- 1. query preparation ---------------- */

/*
 * One wave per (padded) query. mode: 0 = FP32 VALU scan bound, 1 = BF16
 * bound, 3 = FP32 MFMA bound,
 * 2 = INT8 scan: also writes the query's int8 codes (dimp per query, zero
 * past dim) and their squared norm. Exact mode (i8mid == nullptr): codes
 * q - centre; the INT8 scan is exact (eps = 0) and the host takes it only
 * when every query value is an integer within [centre - 128, centre + 127]
 * (a value outside sets kStPrepFail in the query's status, which the
 * threshold stage keeps, so such a run can never pass as exact).
 * Quantized mode: codes clamp(rint((q - i8mid[d]) * i8inv), -127, 127), the
 * rows' transform. Every mode writes the status word, clearing the bit a
 * previous run may have left.
 * The scan is exact (eps = 0) when every vector and query value is an
 * integer of magnitude <= 256 (exact in BF16, products exact in FP32) and
 * every partial sum stays below 2^24: sum (x-q)^2 <= D * delta^2 with
 * delta = max(vmax - qmin, qmax - vmin); ||x||^2 + ||q||^2 <= D (V^2 + Q^2);
 * |2 x.q| <= 2 D V Q (V, Q = largest magnitudes).
 */
__device__ __forceinline__ void prep_query(
    int qi, const float *__restrict__ q, int nq, int dim, int dimb, uint16_t *__restrict__ qbf,
    float *__restrict__ qnorm, float *__restrict__ eps, float xmax_norm,
    float vmin, float vmax, int all_int, int mode, int8_t *__restrict__ qi8 = nullptr,
    float *__restrict__ qn8 = nullptr, int centre = 0, uint32_t *__restrict__ status = nullptr,
    const float *__restrict__ i8mid = nullptr, float i8inv = 1.f, int dimp = 0, int fp8 = 0) {
#pragma clang fp contract(off)
    const int l = threadIdx.x & 63;
    const float *src = q + (uint64_t)qi * dim;
    float ss = 0.f, pmax = 0.f, nmax = 0.f, s8 = 0.f;
    int nonint = 0, out8 = 0;
    for (int j = l; j < dim; j += 64) {
        float v = src[j];
        qbf[(uint64_t)qi * dimb + j] = f2bf(v);
        ss = fmaf(v, v, ss);
        pmax = fmaxf(pmax, v);
        nmax = fmaxf(nmax, -v);
        nonint |= (v != rintf(v)) ? 1 : 0;
        if (mode == 2) {
            float c;
            if (i8mid == nullptr) {
                c = v - (float)centre;
                out8 |= (c != rintf(c) || c < -128.f || c > 127.f) ? 1 : 0;
                qi8[(uint64_t)qi * dimp + j] = (int8_t)(int)fminf(fmaxf(c, -128.f), 127.f);
            } else if (fp8) {
                const uint32_t code = e4m3_encode((v - i8mid[j]) * i8inv);
                qi8[(uint64_t)qi * dimp + j] = (int8_t)code;
                c = e4m3_decode(code);
            } else {
                c = fminf(fmaxf(rintf((v - i8mid[j]) * i8inv), -127.f), 127.f);
                qi8[(uint64_t)qi * dimp + j] = (int8_t)(int)c;
            }
            s8 = fmaf(c, c, s8);
        }
    }
    if (mode == 2)
        for (int j = dim + l; j < dimp; j += 64) qi8[(uint64_t)qi * dimp + j] = 0;
    /* the BF16 operand's zero padding (the tiled copy's dims past dim) */
    for (int j = dim + l; j < dimb; j += 64) qbf[(uint64_t)qi * dimb + j] = 0;
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        ss += lane_xor(ss, o);
        pmax = fmaxf(pmax, lane_xor(pmax, o));
        nmax = fmaxf(nmax, lane_xor(nmax, o));
        nonint |= lane_xor(nonint, o);
        if (mode == 2) {
            s8 += lane_xor(s8, o);
            out8 |= lane_xor(out8, o);
        }
    }
    if (l == 0 && mode == 2) qn8[qi] = s8;
    if (l == 0 && status != nullptr && qi < nq)
        status[qi] = (mode == 2 && out8) ? kStPrepFail : 0u;
    if (l == 0) {
        qnorm[qi] = ss;
        double e = 0.0;
        const double D = (double)dim;
        const double V = fmax((double)vmax, -(double)vmin);
        const double Q = fmax((double)pmax, (double)nmax);
        const double delta = fmax((double)vmax + (double)nmax, (double)pmax - (double)vmin);
        const double lim = 16777216.0;
        const bool exact = all_int && !nonint && V <= 256.0 && Q <= 256.0 &&
                           D * delta * delta < lim && D * (V * V + Q * Q) < lim &&
                           2.0 * D * V * Q < lim;
        if (!exact && qi < nq && mode != 2) {
            const double qn = sqrt((double)ss), xn = (double)xmax_norm;
            const double g = (D + 8.0) * ldexp(1.0, -24);
            const double big = (xn + qn) * (xn + qn);
            if (mode == 1) {
                /* the MFMA k-steps run over the padded dims */
                const double u = ldexp(1.0, -8);
                const double gb = ((double)dimb + 8.0) * ldexp(1.0, -24);
                e = (4.0 * u + 2.0 * u * u) * xn * qn + 4.0 * gb * big;
            } else if (mode == 3) {
                /* FP32 MFMA x.q with the norm seed: unrounded inputs, the
                 * BF16 form's accumulation allowance */
                e = 4.0 * g * big;
            } else {
                e = 2.0 * g * big;
            }
            e = e * 1.25 + 1e-30;
        }
        eps[qi] = (float)e;
    }
}

/*
 * Phase 1 of a run, one launch: blocks [0, nprep) prepare the queries (one
 * wave per padded query, four per block; one block per query when they also
 * compute the router features), the other nbm blocks build the predicate
 * bitmaps (bitmap_rows: kBmRows rows per block, the waves split the
 * distinct programs).
 */
struct Phase1Args {
    const PvOp *ops;
    const int32_t *poff, *plen;
    const double *vals, *cols;
    uint64_t count, words;
    uint32_t *bitmap;
    uint32_t *progpart;   /* optional: passing rows per program and bitmap
                             block, [nprog][nbm] */
    const int32_t *hdr;   /* device [nprog, nops, nvals] */
    int ncols;
    uint32_t nbm;         /* bitmap blocks */
    uint32_t nprep;       /* query prep blocks (first in the grid) */
    const float *q;
    int nq, nqp, dim;
    int dimb;             /* row length of qbf (the BF16 copy's padded dims) */
    uint16_t *qbf;
    float *qnorm, *eps;
    float xmax_norm, vmin, vmax;
    int all_int, mode;
    int8_t *qi8;          /* mode 2 */
    float *qn8;
    int centre;
    const float *i8mid;   /* mode 2, quantized codes: per-dim centre, scale */
    int fp8;              /* mode 2: E4M3 codes (FP8 index) */
    float i8inv;
    int dimp;             /* mode 2: code row length */
    uint32_t *status;
    /* learned routing (dim <= 256): router features of each query except
     * its selectivity, [nqp][nfeat] — the query, then (slot dim left for the
     * selectivity) its kNear nearest-centroid distances */
    const float *cent_t;  /* centroids transposed [dim][ncent] */
    int ncent, nfeat;
    float *feat;
};

__global__ __launch_bounds__(256) void k_phase1(Phase1Args A) {
    /* the query prep blocks come first: dispatched before the bitmap
     * blocks, they never wait for a free slot behind them */
    if (blockIdx.x >= A.nprep) {
        bitmap_rows(blockIdx.x - A.nprep, A.hdr, A.ops, A.poff, A.plen, A.vals, A.cols,
                    A.count, A.ncols, A.words, A.bitmap, A.progpart, A.nbm);
        return;
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (A.feat == nullptr) {
        /* query prep only: one wave per padded query */
        const int qi = (int)blockIdx.x * 4 + w;
        if (qi >= A.nqp) return;
        prep_query(qi, A.q, A.nq, A.dim, A.dimb, A.qbf, A.qnorm, A.eps, A.xmax_norm, A.vmin, A.vmax,
                   A.all_int, A.mode, A.qi8, A.qn8, A.centre, A.status, A.i8mid, A.i8inv, A.dimp,
                   A.fp8);
        return;
    }
    /* with router features: one block per padded query; wave 0 preps it,
     * the four waves split the centroid distances (each keeps its kNear
     * nearest) and wave 0 merges the 4 x kNear — the same kNear smallest
     * distances as one wave over all centroids. The selectivity feature
     * needs the finished bitmap counts and is added by the decision. */
    const int qi = (int)blockIdx.x;
    if (qi >= A.nqp) return;
    if (w == 0)
        prep_query(qi, A.q, A.nq, A.dim, A.dimb, A.qbf, A.qnorm, A.eps, A.xmax_norm, A.vmin, A.vmax,
                   A.all_int, A.mode, A.qi8, A.qn8, A.centre, A.status, A.i8mid, A.i8inv, A.dimp,
                   A.fp8);
    if (qi >= A.nq) return; /* block-uniform */
    __shared__ float xq[256];
    __shared__ float wn[4][kNear], wm[kNear];
    const float *qv = A.q + (uint64_t)qi * A.dim;
    float *fo = A.feat + (uint64_t)qi * A.nfeat;
    for (int j = threadIdx.x; j < A.dim; j += blockDim.x) {
        xq[j] = qv[j];
        fo[j] = qv[j];
    }
    __syncthreads();
    float best[kNear];
#pragma unroll
    for (int k = 0; k < kNear; k++) best[k] = __uint_as_float(0x7f800000u);
    for (int c0 = 64 * w; c0 < A.ncent; c0 += 256) {
        const int c = min(c0 + l, A.ncent - 1);
        const float v = cent_dist(xq, A.cent_t, c, A.dim, A.ncent);
        near_insert(best, c0 + l < A.ncent ? v : __uint_as_float(0x7f800000u));
    }
    wave_nearest(best, wn[w]);
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int k = 0; k < kNear; k++) best[k] = __uint_as_float(0x7f800000u);
        near_insert(best, wn[l / kNear][l % kNear]);
        wave_nearest(best, wm);
        wave_lds_sync();
        if (l < kNear) fo[A.dim + 1 + l] = sqrtf(wm[l]);
    }
}

/* ---------------- shared scan pieces ---------------- */

/*
 * BF16 copy layout ("MFMA-tiled"): rows are grouped in tiles of 32; a tile
 * is dim/16 chunks of 1 KB, chunk s holding in 16-byte slot l the 8 bf16
 * that lane l feeds to k-step s of v_mfma_f32_32x32x16_bf16: row (l & 31),
 * dims (l >> 5) * dim/2 + 8 s .. +8. A wave therefore loads each k-step with
 * one fully coalesced 1 KB dwordx4 access and uses it as the A operand with
 * no shuffles. Written once at upload (pv_index.hip, k_derive_rows).
 */
struct ScanArgs {
    const uint16_t *vbf;
    const int8_t *vi8;        /* INT8 scan: tiled x - centre */
    const int8_t *qi8;        /* INT8 scan: [nqp][dim] q - centre */
    const float *vecs;
    const float *xnorm;
    uint64_t count, count_pad;
    int dim;
    int dimb;                 /* BF16 scan: dims of the tiled copy and of qbf
                                 rows (dim rounded up to a multiple of 128) */
    int dimp;                 /* INT8 scan: dims of the code copy and of qi8 rows */
    const uint16_t *qbf;
    const float *qf;
    const float *qnorm;
    const float *tau;
    const int32_t *pofq;
    const uint32_t *bitmap;
    uint64_t words;
    int nq;
    uint32_t ntiles;      /* 32-row tiles visited by this launch */
    uint32_t stride;      /* tile t of the launch is physical tile t * stride */
    uint64_t *pool;
    uint32_t *pcnt;
    uint64_t cap_pool;
    uint32_t *gmin;       /* [nq][ngroups] pre-pass minima (FP32 bits, >= 0) */
    uint32_t ngroups;
    const uint32_t *ones;     /* all-ones bitmap row (queries without predicate) */
    uint32_t *status;         /* per-query status bits */
    const uint8_t *skip;      /* optional [nq]: 1 = query routed elsewhere */
    const InlPred *inl;       /* inline predicates [nqp] (INL kernels) */
    const double *icol;       /* their column [count] */
    const uint32_t *irank;    /* the column's row ranks [count] (query-resident scan) */
    float xn2max;             /* max ||x||^2 over the rows (candidate-test margin) */
    uint32_t nchunk, ngroup;  /* query-major grid: row chunks, query groups */
    uint32_t qmajor;          /* grid (query tiles, row chunks): the query tiles
                                 reading one chunk are dispatched together */
    uint32_t sweep;           /* tile order: 0 = each WG streams its own contiguous
                                 chunk; 1 = all waves sweep the rows together
                                 (wave g takes tiles g, g + G, g + 2G, ...) */
};

/* row chunk (WG of the row dimension), chunk count and query-tile group of
 * a scan WG: grid (chunks, query tiles), or (query tiles, chunks) when
 * qmajor so that the query tiles sharing a chunk run together and read it
 * once from HBM (later reads hit in L2 / the last-level cache) */
/* Query-major grid (1D): workgroups are dispatched to the XCDs round-robin
 * (block b on XCD b % kXcds), so the ngroup query groups of chunk c are the
 * blocks kXcds (k ngroup + g) + c % kXcds with k = c / kXcds: all on chunk
 * c's XCD and consecutive there, so the chunk is read into that XCD's L2
 * once and reused by every query group. Blocks past the last chunk exit. */
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t scan_chunk(const ScanArgs &A) {
    return A.qmajor ? (blockIdx.x / kXcds / A.ngroup) * kXcds + blockIdx.x % kXcds : blockIdx.x;
}
__device__ __forceinline__ uint32_t scan_nchunks(const ScanArgs &A) {
    return A.qmajor ? A.nchunk : gridDim.x;
}
__device__ __forceinline__ uint32_t scan_qgroup(const ScanArgs &A) {
    return A.qmajor ? (blockIdx.x / kXcds) % A.ngroup : blockIdx.y;
}
/* a padding block of the query-major grid (uniform) */
__device__ __forceinline__ bool scan_pad(const ScanArgs &A) {
    return A.qmajor && scan_chunk(A) >= A.nchunk;
}

/* the main-scan launch of a batch's last full-scan run, kept so the scan
 * kernel alone can be re-timed (fv_gpu_batch_time_scan) */
struct ScanLaunch {
    ScanArgs M;
    bool bf, inl, i8, f32m;
    bool f8;                /* INT8 layout holding FP8 codes */
    int pt;                 /* INT8: 32-row tiles per 8 KB unit */
    int qtw, nw, nqt, gy;
    uint32_t gx;
    size_t shm;
};

/*
 * The stream of units (tile, unit within the tile) one wave of a BF16 scan
 * WG loads. Chunked order: the WG owns tiles [t0, t1); each wave starts
 * with tiles t0 + wave and t0 + NW + wave and then takes the WG's next
 * untaken tile from an LDS counter, one tile ahead of need (the counter's
 * answer is read a whole tile after the request). Waves of one WG that see
 * slower memory take fewer tiles instead of holding the WG's final flush
 * back (static strides left waves of a WG finishing up to 5 us apart).
 * Sweep order: wave g of the grid takes tiles g, g + G, ... (static).
 */
struct UnitCursor {
    uint32_t tile, c;   /* next unit to load */
    uint32_t nxt;       /* the wave's following tile */
    uint32_t pend;      /* lane 0: counter value requested for the tile after */
    uint32_t end;       /* first tile past the wave's range */
    uint32_t step;      /* sweep order: tile stride; 0 = counter */
};

template <int NW>
__device__ __forceinline__ UnitCursor cursor_init(const ScanArgs &A, int wave,
                                                  uint32_t *s_next) {
    UnitCursor u;
    const uint32_t nwg = scan_nchunks(A), wg = scan_chunk(A);
    u.c = 0;
    u.pend = 0;
    if (A.sweep) {
        u.tile = wg * (uint32_t)NW + (uint32_t)wave;
        u.step = nwg * (uint32_t)NW;
        u.nxt = u.tile + u.step;
        u.end = A.ntiles;
    } else {
        const uint32_t t0 = (uint32_t)(((uint64_t)A.ntiles * wg) / nwg);
        u.end = (uint32_t)(((uint64_t)A.ntiles * (wg + 1)) / nwg);
        u.tile = t0 + (uint32_t)wave;
        u.nxt = t0 + (uint32_t)NW + (uint32_t)wave;
        u.step = 0;
        if (threadIdx.x == 0) *s_next = t0 + 2u * NW;
    }
    return u;
}

/* first counter request of a wave (after the barrier that publishes the
 * counter's start) */
__device__ __forceinline__ void cursor_start(UnitCursor &u, uint32_t *s_next) {
    if (u.step == 0 && (threadIdx.x & 63) == 0) u.pend = atomicAdd(s_next, 1u);
}

/* move to the next unit; at a tile boundary the following tile becomes
 * current and a new one is requested */
__device__ __forceinline__ void cursor_next(UnitCursor &u, uint32_t nch, uint32_t *s_next) {
    if (++u.c < nch) return;
    u.c = 0;
    u.tile = u.nxt;
    if (u.step) {
        u.nxt = u.tile + u.step;
    } else {
        u.nxt = __builtin_amdgcn_readfirstlane(u.pend);
        if ((threadIdx.x & 63) == 0) u.pend = atomicAdd(s_next, 1u);
    }
}

/* the last tile a wave's loads may touch. The pipeline loads one unit
 * ahead, so after its last consumed unit a wave still loads one; clamped
 * here it re-reads the end of the wave's own range, which its WG has just
 * read (an L2 hit), instead of the next chunk's first tile (an HBM read
 * that chunk's WG, on another XCD, makes again). Never above ntiles - 1. */
__device__ __forceinline__ uint32_t cursor_last(const UnitCursor &u, uint32_t ntiles) {
    return max(min(u.end, ntiles), 1u) - 1u;
}

/* true when every query of the WG's 64-query tile is skipped (uniform) */
__device__ __forceinline__ bool tile_skipped(const ScanArgs &A, int qbase) {
    if (A.skip == nullptr) return false;
    for (int i = 0; i < kQT; i++) {
        const int q = qbase + i;
        if (q < A.nq && A.skip[q] == 0) return false;
    }
    return true;
}

constexpr uint32_t kSpill = 512; /* per-WG LDS candidate overflow slots */
__device__ __forceinline__ uint64_t pcnt_at(int q, uint32_t s) {
    return ((uint64_t)q * kPoolShards + s) * kPcntStride;
}
/* the segment of a query's pool (cap entries) that shard s owns */
__device__ __forceinline__ uint64_t shard_cap(uint64_t cap) { return cap / kPoolShards; }
__device__ __forceinline__ void clear_pcnt(uint32_t *pcnt, int q) {
#pragma unroll
    for (uint32_t s = 0; s < kPoolShards; s++) pcnt[pcnt_at(q, s)] = 0;
}

/* LDS candidate state of one scan WG */
struct CandLds {
    uint32_t cnt[kQT];
    uint32_t base[kQT];
    uint32_t nsp, ovf;
    uint64_t buf[kQT * kLdsCand];
    uint64_t spk[kSpill];
    uint32_t spq[kSpill];
};

/* bitmap row of a query (the all-ones row when it has no predicate) */
__device__ __forceinline__ const uint32_t *bitmap_row(const ScanArgs &A, int32_t p) {
    return p >= 0 ? A.bitmap + (uint64_t)p * A.words : A.ones;
}

/*
 * Append a candidate: per-query LDS slots first, then a WG-wide LDS spill
 * area. The scan loop touches only LDS here (no global atomics or stores,
 * whose completion would gate the next tile's prefetched loads). Spills
 * past the area set the WG's overflow flag (the host then re-runs with the
 * tightest threshold).
 */
/* Append a candidate of query qbase + qi: LDS slots first, then the WG's
 * LDS spill area, then (rare: a WG holding more than its share of one
 * query's candidates) straight into the WG's global pool segment. */
__device__ __forceinline__ void insert_candidate(const ScanArgs &A, int qbase, int qi,
                                                 uint64_t key, CandLds &C) {
    const uint32_t pos = atomicAdd(&C.cnt[qi], 1u);
    if (pos < (uint32_t)kLdsCand) {
        C.buf[qi * kLdsCand + pos] = key;
        return;
    }
    const uint32_t sp = atomicAdd(&C.nsp, 1u);
    if (sp < kSpill) {
        C.spk[sp] = key;
        C.spq[sp] = (uint32_t)qi;
        return;
    }
    const int q = qbase + qi;
    const uint32_t sh = scan_chunk(A) % kPoolShards;
    const uint64_t scap = shard_cap(A.cap_pool);
    const uint32_t gp = atomicAdd(&A.pcnt[pcnt_at(q, sh)], 1u);
    if (gp < scap)
        A.pool[(uint64_t)q * A.cap_pool + sh * scap + gp] = key;
    else
        C.ovf = 1u;
}

__device__ __forceinline__ void cand_init(CandLds &C) {
    for (int i = threadIdx.x; i < kQT; i += blockDim.x) C.cnt[i] = 0;
    if (threadIdx.x == 0) {
        C.nsp = 0;
        C.ovf = 0;
    }
}

/* row held in accumulator element i by lane l (32x32 MFMA C layout) */
__device__ __forceinline__ int acc_row(int i, int l) {
    return (i >> 2) * 8 + (l >> 5) * 4 + (i & 3);
}

/* per-tile epilogue inputs, loaded together with the tile's rows: one
 * value per lane (row lane & 31), expanded to the lane's 16 accumulator
 * rows by shuffles when the epilogue runs, so a prefetched tile costs a
 * few registers beyond its operands */
struct TileIn {
    float xr;          /* ||x||^2 of row (lane & 31) */
    uint32_t w[2];     /* predicate bitmap word of the tile, per query block */
    uint32_t valid;    /* rows < count */
    double rv;         /* inline predicates: column value of row (lane & 31) */
    uint32_t t, c;     /* logical tile of the unit and unit within the tile */
    bool live;         /* the unit is part of this wave's work */
};

/* a per-row value (held by lane row & 31) for the lane's 16 accumulator rows */
__device__ __forceinline__ void expand_norms(float xr, float (&xn)[16]) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 16; i++) xn[i] = __shfl(xr, acc_row(i, l));
}

/* expand_norms through the wave's 32-entry LDS slot: one store per row
 * lane, four 16-byte reads of rows g * 8 + (l >> 5) * 4 .. +3 (a wave's
 * LDS operations run in order) */
__device__ __forceinline__ void expand_norms_lds(float xr, float (&xn)[16], int32_t *slot) {
    const int l = threadIdx.x & 63;
    if (l < 32) slot[l] = __float_as_int(xr);
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const float4 v = *(const float4 *)(slot + g * 8 + (l >> 5) * 4);
        xn[4 * g] = v.x;
        xn[4 * g + 1] = v.y;
        xn[4 * g + 2] = v.z;
        xn[4 * g + 3] = v.w;
    }
}

/* inline predicates: the column values of the lane's 16 accumulator rows,
 * gathered from the lanes that loaded them (one value per row per lane) */
__device__ __forceinline__ void expand_rows(double rv, double (&cv)[16]) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 16; i++) cv[i] = __shfl(rv, acc_row(i, l));
}

/* inline predicates staged per WG: comparison mask (-1: no predicate) and
 * constant of query qi of query tile k at [k * 64 + qi] */
__device__ __forceinline__ void stage_inline(const ScanArgs &A, int k, int qbase, int *s_im,
                                             double *s_iv) {
    for (int i = threadIdx.x; i < kQT; i += blockDim.x) {
        const int q = qbase + i;
        int m = -1;
        double v = 0.0;
        if (q < A.nq) {
            const InlPred ip = A.inl[q];
            m = ip.col < 0 ? -1 : ip.mask;
            v = ip.value;
        }
        s_im[k * kQT + i] = m;
        s_iv[k * kQT + i] = v;
    }
}

/* the predicate word (bit = accumulator row) of the lane's query of
 * block qb in query tile k, from the expanded column values */
__device__ __forceinline__ uint32_t inline_word(const double (&cv)[16], const int *s_im,
                                                const double *s_iv, int k, int qb) {
    const int l = threadIdx.x & 63;
    const int qi = k * kQT + (l & 31) + 32 * qb;
    const int m = s_im[qi];
    if (m < 0) return 0xFFFFFFFFu;
    const double c = s_iv[qi];
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) w |= cmp_pass(cv[i], c, m) << acc_row(i, l);
    return w;
}

template <bool INL = false>
__device__ __forceinline__ void load_tile_in(const ScanArgs &A, uint32_t pt,
                                             const uint32_t *const (&wr)[2], TileIn &in) {
    const int l = threadIdx.x & 63;
    in.xr = A.xnorm[(uint64_t)pt * 32 + (l & 31)];
    if (INL) {
        in.rv = A.icol[min((uint64_t)pt * 32 + (l & 31), A.count - 1)];
    } else {
        in.w[0] = wr[0][pt];
        in.w[1] = wr[1][pt];
    }
    const uint64_t row0 = (uint64_t)pt * 32;
    in.valid = row0 + 32 <= A.count ? 0xFFFFFFFFu
               : (A.count > row0 ? (0xFFFFFFFFu >> (32 - (A.count - row0))) : 0u);
}

/*
 * Consume one 32-row x 64-query tile of FP32 distances sum (x - q)^2 held
 * in the MFMA C layout (the plain FP32 scan): lane l owns queries (l & 31)
 * + 32 qb and the 16 rows acc_row(i, l). PREPASS writes the minimum passing
 * distance of the lane's 16 rows per query (one group); MAIN appends
 * passing rows with dist <= tau to the WG's candidates.
 */
template <bool PREPASS>
__device__ __forceinline__ void tile_epilogue(const ScanArgs &A, int qbase, uint32_t t,
                                              uint32_t pt, const float (&acc)[2][16],
                                              const TileIn &in, const float (&tq)[2],
                                              CandLds &C) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int qb = 0; qb < 2; qb++) {
        const int qi = (l & 31) + 32 * qb;
        const int q = qbase + qi;
        if (q >= A.nq) continue;
        const uint32_t word = in.w[qb] & in.valid;
        if (PREPASS) {
            float m = __uint_as_float(0x7f800000u);
#pragma unroll
            for (int i = 0; i < 16; i++)
                if ((word >> acc_row(i, l)) & 1u) m = fminf(m, acc[qb][i]);
            const uint32_t grp = t * 2 + (l >> 5);
            A.gmin[(uint64_t)q * A.ngroups + grp] = __float_as_uint(m);
        } else {
            const float tau = tq[qb];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const float d = acc[qb][i];
                const int r = acc_row(i, l);
                if (((word >> r) & 1u) && d <= tau)
                    insert_candidate(A, qbase, qi,
                                     make_key(d, (uint32_t)((uint64_t)pt * 32 + r)), C);
            }
        }
    }
}

/*
 * Seeded epilogue of the MFMA scans. The accumulators of a tile start from
 * a per-element seed instead of zero, which folds in the row norm and the
 * query's filter:
 *   BF16: -||x_r||^2 / 2 where row r passes query q's filter, -inf
 *         otherwise; then acc = x.q - ||x||^2 / 2 and d = ||q||^2 - 2 acc;
 *   INT8: -floor(||x_r||^2 / 2), or kI8Off (far below any real value);
 *         then acc = x.q - floor(||x||^2 / 2) (exact, int32) and
 *         d = ||q||^2 + (||x||^2 mod 2) - 2 acc.
 * d <= tau becomes acc >= h(q), one threshold per query, so the lane's 16
 * elements of a query reduce to their maximum before the one comparison;
 * the filter test and the norm term are paid once per element, in the
 * seed (an extract of the filter bit and a bit-select).
 */
constexpr int32_t kI8Off = -(1 << 30);

/* the lane's filter word, shifted so that element i's row is at bit
 * (i >> 2) * 8 + (i & 3) */
__device__ __forceinline__ uint32_t lane_word(uint32_t w) {
    return w >> (((threadIdx.x & 63) >> 5) * 4);
}

/* all ones if element i's row passes (bit of a lane_word), else zero */
__device__ __forceinline__ uint32_t elem_mask(uint32_t wa, int i) {
    const int p = (i >> 2) * 8 + (i & 3);
    return (uint32_t)((int32_t)(wa << (31 - p)) >> 31);
}

/* BF16 seeds of a tile: sn = -||x||^2 / 2 of the lane's 16 accumulator
 * rows (expand_norms of the lane's own row's value, halved and negated
 * once per lane, both exact), w0 / w1 the query blocks' filter words with
 * the validity mask applied */
__device__ __forceinline__ void seed_bf(const float (&sn)[16], uint32_t w0, uint32_t w1,
                                        f32x16 &a0, f32x16 &a1) {
    const uint32_t v0 = lane_word(w0), v1 = lane_word(w1);
    constexpr uint32_t kNegInf = 0xff800000u;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t s = __float_as_uint(sn[i]);
        const uint32_t m0 = elem_mask(v0, i), m1 = elem_mask(v1, i);
        a0[i] = __uint_as_float((s & m0) | (kNegInf & ~m0));
        a1[i] = __uint_as_float((s & m1) | (kNegInf & ~m1));
    }
}

/* INT8 seeds from xr = ||x||^2 of the lane's own row (lane & 31, an
 * integer): -floor(xr / 2) is formed once per lane and spread to the 16
 * accumulator rows through the wave's 32-entry LDS slot (one store per
 * row lane, four 16-byte reads of rows g * 8 + (l >> 5) * 4 .. +3; a
 * wave's LDS operations run in order); par gets bit r = ||x_r||^2 is odd
 * for rows 0..31 (one ballot), read only when a candidate is emitted */
__device__ __forceinline__ void seed_i8(float xr, uint32_t w0, uint32_t w1,
                                        i32x16 &c0, i32x16 &c1, uint32_t &par,
                                        int32_t *slot) {
    const int l = threadIdx.x & 63;
    const int32_t xi = (int32_t)xr;
    const int32_t half = -(xi >> 1);
    par = (uint32_t)__ballot(xi & 1);
    const uint32_t v0 = lane_word(w0), v1 = lane_word(w1);
    if (l < 32) slot[l] = half;
    int32_t hs[16];
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const int4 v = *(const int4 *)(slot + g * 8 + (l >> 5) * 4);
        hs[4 * g] = v.x;
        hs[4 * g + 1] = v.y;
        hs[4 * g + 2] = v.z;
        hs[4 * g + 3] = v.w;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t s = (uint32_t)hs[i];
        const uint32_t m0 = elem_mask(v0, i), m1 = elem_mask(v1, i);
        c0[i] = (int32_t)((s & m0) | ((uint32_t)kI8Off & ~m0));
        c1[i] = (int32_t)((s & m1) | ((uint32_t)kI8Off & ~m1));
    }
}

/* BF16: the admission threshold of a query: acc >= h admits every row whose
 * d = ||q||^2 - 2 acc (rounded) is <= tau. The two forms round differently
 * by a few ulps of ||x||^2 + ||q||^2; h is lowered by 2^-19 (max ||x||^2 +
 * ||q||^2), far above that (a few more rows in the margin band are
 * harmless: the select stage certifies). Kept finite so that -inf seeds
 * (filtered rows) never pass, also when tau = +inf admits every row. */
__device__ __forceinline__ float bf_threshold(const ScanArgs &A, float qn, float tq) {
    const float h = fmaf(0.5f, qn - tq, -ldexpf(A.xn2max + qn, -19));
    return fmaxf(h, -__FLT_MAX__);
}

/* INT8: d <= tau  =>  2 acc >= ||q||^2 - tau + parity  =>  acc >= ceil(ceil(
 * ||q||^2 - tau) / 2): a superset (rows up to 2 units beyond tau), exact in
 * integers; clamped between the filtered seed and any real value */
__device__ __forceinline__ int32_t i8_threshold(float qn, float tq) {
    const double k = ceil((double)qn - (double)tq);
    const double tt = ceil(k * 0.5);
    return (int32_t)fmin(fmax(tt, (double)(kI8Off / 2)), (double)(1 << 30));
}

template <bool PREPASS>
__device__ __forceinline__ void epilogue_bf(const ScanArgs &A, int qbase, uint32_t t,
                                            uint32_t pt, const f32x16 &a0, const f32x16 &a1,
                                            const float (&qn)[2], const float (&tq)[2],
                                            CandLds &C) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int qb = 0; qb < 2; qb++) {
        const f32x16 &a = qb ? a1 : a0;
        const int qi = (l & 31) + 32 * qb;
        const int q = qbase + qi;
        if (q >= A.nq) continue;
        float mx = a[0];
#pragma unroll
        for (int i = 1; i < 16; i++) mx = fmaxf(mx, a[i]);
        if (PREPASS) {
            /* all rows filtered: mx = -inf, d = +inf */
            const float d = fmaxf(fmaf(-2.f, mx, qn[qb]), 0.f);
            const uint32_t grp = t * 2 + (l >> 5);
            A.gmin[(uint64_t)q * A.ngroups + grp] = __float_as_uint(d);
        } else {
            const float h = bf_threshold(A, qn[qb], tq[qb]);
            /* lane-divergent: the wave skips the insertion code when no
             * lane has a candidate among its rows */
            if (mx >= h) {
#pragma unroll
                for (int i = 0; i < 16; i++)
                    if (a[i] >= h) {
                        const float d = fmaxf(fmaf(-2.f, a[i], qn[qb]), 0.f);
                        insert_candidate(A, qbase, qi, make_key(d, (uint32_t)((uint64_t)pt * 32 +
                                                                    acc_row(i, l))), C);
                    }
            }
        }
    }
}

template <bool PREPASS>
__device__ __forceinline__ void epilogue_i8(const ScanArgs &A, int qbase, uint32_t t,
                                            uint32_t pt, const i32x16 &c0, const i32x16 &c1,
                                            uint32_t par, const float (&qn)[2],
                                            const int32_t (&th)[2], CandLds &C) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int qb = 0; qb < 2; qb++) {
        const i32x16 &a = qb ? c1 : c0;
        const int qi = (l & 31) + 32 * qb;
        const int q = qbase + qi;
        if (q >= A.nq) continue;
        int32_t mx = a[0];
#pragma unroll
        for (int i = 1; i < 16; i++) mx = max(mx, a[i]);
        const int32_t qni = (int32_t)qn[qb];
        if (PREPASS) {
            /* an upper bound of the group's minimum (parity taken as 1):
             * at least as large as a real row's distance, at most 1 more */
            const float d = mx > kI8Off / 2 ? (float)(qni + 1 - 2 * mx)
                                             : __uint_as_float(0x7f800000u);
            const uint32_t grp = t * 2 + (l >> 5);
            A.gmin[(uint64_t)q * A.ngroups + grp] = __float_as_uint(d);
        } else if (mx >= th[qb]) {
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (a[i] >= th[qb]) {
                    const float d =
                        (float)(qni + (int32_t)((par >> acc_row(i, l)) & 1u) - 2 * a[i]);
                    insert_candidate(A, qbase, qi, make_key(d, (uint32_t)((uint64_t)pt * 32 +
                                                                acc_row(i, l))), C);
                }
        }
    }
}

/* Flush the WG's LDS candidates and spills to the pool (MAIN). */
__device__ __forceinline__ void flush_candidates(const ScanArgs &A, int qbase,
                                                 CandLds &C) {
    __syncthreads();
    /* one reservation per query in this WG's pool segment, all in flight
     * together */
    const uint32_t sh = scan_chunk(A) % kPoolShards;
    const uint64_t scap = shard_cap(A.cap_pool);
    if (threadIdx.x < kQT) {
        const int qi = threadIdx.x, q = qbase + qi;
        const uint32_t n = min(C.cnt[qi], (uint32_t)kLdsCand);
        C.base[qi] = (q < A.nq && n > 0) ? atomicAdd(&A.pcnt[pcnt_at(q, sh)], n) : 0u;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < kQT * kLdsCand; idx += blockDim.x) {
        const int qi = idx / kLdsCand, j = idx % kLdsCand, q = qbase + qi;
        const uint32_t n = min(C.cnt[qi], (uint32_t)kLdsCand);
        if (q < A.nq && (uint32_t)j < n && C.base[qi] + j < scap)
            A.pool[(uint64_t)q * A.cap_pool + sh * scap + C.base[qi] + j] = C.buf[idx];
    }
    const uint32_t ns = min(C.nsp, kSpill);
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint32_t q = (uint32_t)qbase + C.spq[i];
        const uint32_t gp = atomicAdd(&A.pcnt[pcnt_at(q, sh)], 1u);
        if (gp < scap) A.pool[(uint64_t)q * A.cap_pool + sh * scap + gp] = C.spk[i];
    }
    if (C.ovf && threadIdx.x < kQT && qbase + (int)threadIdx.x < A.nq)
        atomicOr(&A.status[qbase + threadIdx.x], kStOverflow);
}

/* per-lane query state for the two column blocks */
__device__ __forceinline__ void load_query_state(const ScanArgs &A, int qbase,
                                                 float (&tq)[2], float (&qn)[2],
                                                 const uint32_t *(&wr)[2]) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int qb = 0; qb < 2; qb++) {
        const int q = qbase + (l & 31) + 32 * qb;
        const bool v = q < A.nq;
        tq[qb] = v ? A.tau[q] : 0.f;
        qn[qb] = v ? A.qnorm[q] : 0.f;
        wr[qb] = bitmap_row(A, v ? A.pofq[q] : -1);
    }
}

/* ---------------- 3/5. BF16 MFMA scan ---------------- */

/*
 * grid.x: row chunks, grid.y: query tiles of 64; NW waves per WG. A wave
 * walks its tiles (UnitCursor: taken from its WG's contiguous chunk through
 * an LDS counter, or a static sweep over all rows shared by every wave of
 * the grid) as a stream of "units" of 8 k-steps
 * (8 KB of the MFMA-tiled copy; dim/128 units per tile). Units alternate
 * between two register buffers: the loads of unit u+1 — its rows, and the
 * tile's ||x||^2 and bitmap words — are issued before the 16
 * v_mfma_f32_32x32x16_bf16 of unit u, and after the last unit of a tile its
 * epilogue runs while the next tile streams in (a third buffer, two units
 * in flight per wave, measured slower: the stream is not short of loads in
 * flight at two waves per SIMD). Loads are unconditional (indices clamped
 * to the last unit) so the pipeline is one straight line for the
 * wait-count analysis. B operands (the 64 queries) sit in LDS.
 * Runs over the tiled copy's dimb dims (dim rounded up to a multiple of
 * 128, zero-filled in rows and queries).
 */
template <bool PREPASS, int NW, bool INL = false>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2)))
void k_scan_bf16(ScanArgs A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int S = A.dimb / 16;           /* MFMA k-steps per tile */
    const uint32_t NCH = (uint32_t)(S / 8); /* units per tile */
    u32x4 *qsh = (u32x4 *)smem;          /* [2][S][64] x 16 B */
    CandLds &C = *(CandLds *)(smem + (size_t)2 * S * 64 * 16);

    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    if (scan_pad(A)) return;
    const int qbase = scan_qgroup(A) * kQT;
    if (tile_skipped(A, qbase)) return;

    __shared__ uint32_t s_next;
    UnitCursor cur = cursor_init<NW>(A, wave, &s_next);
    const u32x4 *vb = (const u32x4 *)A.vbf;
    /* units past the wave's range load a valid tile and are not consumed */
    const uint32_t tlast = cursor_last(cur, A.ntiles);
    auto phys = [&](uint32_t t) { return min(t, tlast) * A.stride; };
    auto load_rows = [&](u32x4 (&a)[8]) {
        const u32x4 *p = vb + (uint64_t)phys(cur.tile) * (4 * A.dimb) + (uint64_t)cur.c * 512 + l;
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = p[64 * j];
    };
    /* the first unit's rows are requested before the query operands are
     * staged, so their latencies overlap instead of adding up */
    u32x4 a0[8], a1[8];
    load_rows(a0);

    const int half = A.dimb / 2;
    for (int e = tid; e < 2 * S * 64; e += blockDim.x) {
        const int lane = e & 63, s = (e >> 6) % S, qb = (e >> 6) / S;
        const int q = qbase + (lane & 31) + 32 * qb;
        qsh[e] = *(const u32x4 *)(A.qbf + (uint64_t)q * A.dimb + (lane >> 5) * half + 8 * s);
    }
    __shared__ int s_im[INL ? kQT : 1];
    __shared__ double s_iv[INL ? kQT : 1];
    if (INL) stage_inline(A, 0, qbase, s_im, s_iv);
    if (!PREPASS) cand_init(C);

    float tq[2], qn[2];
    const uint32_t *wr[2];
    load_query_state(A, qbase, tq, qn, wr);
    /* the unit's epilogue inputs and identity; then step the cursor */
    auto load_in = [&](TileIn &in) {
        load_tile_in<INL>(A, phys(cur.tile), wr, in);
        in.t = cur.tile;
        in.c = cur.c;
        in.live = cur.tile < cur.end;
        cursor_next(cur, NCH, &s_next);
    };
    TileIn in0, in1;
    __syncthreads();
    cursor_start(cur, &s_next);
    load_in(in0);
    auto do_unit = [&](const u32x4 (&a)[8], const TileIn &in, f32x16 &acc0, f32x16 &acc1) {
        const uint32_t c = in.c;
        if (c == 0) {
            /* seeds: row norms and the queries' filters (epilogue_bf) */
            float sn[16];
            expand_norms(-0.5f * in.xr, sn);
            uint32_t w0 = in.w[0], w1 = in.w[1];
            if (INL) {
                double cv[16];
                expand_rows(in.rv, cv);
                w0 = inline_word(cv, s_im, s_iv, 0, 0);
                w1 = inline_word(cv, s_im, s_iv, 0, 1);
            }
            seed_bf(sn, w0 & in.valid, w1 & in.valid, acc0, acc1);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int s = (int)c * 8 + j;
            const bf16x8 av = __builtin_bit_cast(bf16x8, a[j]);
            const bf16x8 b0 = __builtin_bit_cast(bf16x8, qsh[(0 * S + s) * 64 + l]);
            const bf16x8 b1 = __builtin_bit_cast(bf16x8, qsh[(1 * S + s) * 64 + l]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b1, acc1, 0, 0, 0);
        }
        if (c == NCH - 1)
            epilogue_bf<PREPASS>(A, qbase, in.t, in.t * A.stride, acc0, acc1, qn, tq, C);
    };

    f32x16 acc0 = {}, acc1 = {};
    while (in0.live) {
        load_rows(a1);
        load_in(in1);
        do_unit(a0, in0, acc0, acc1);
        if (!in1.live) break;
        load_rows(a0);
        load_in(in0);
        do_unit(a1, in1, acc0, acc1);
    }
    if (!PREPASS) flush_candidates(A, qbase, C);
}

/*
 * Multi-query-tile variant of the BF16 scan for dimb 128 (one 8 KB unit per
 * tile): a WG serves QTW query tiles, so every data tile loaded from HBM
 * feeds QTW x 16 MFMAs and QTW epilogues instead of being re-read by QTW
 * WGs. Query operands of all QTW tiles and one candidate area per query
 * tile sit in LDS. grid = (row chunks, ceil(query tiles / QTW)).
 */
template <bool PREPASS, int QTW, int NW, bool INL>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2)))
void k_scan_bf16_mq(ScanArgs A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int S = 8;                     /* dim 128: 8 k-steps, 1 unit per tile */
    u32x4 *qsh = (u32x4 *)smem;              /* [QTW][2][S][64] x 16 B */
    CandLds *Cs = (CandLds *)(smem + (size_t)QTW * 2 * S * 64 * 16);

    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    if (scan_pad(A)) return;
    const int nqt = (A.nq + kQT - 1) / kQT;
    int qb0[QTW];
    bool qon[QTW];
    bool any = false;
#pragma unroll
    for (int k = 0; k < QTW; k++) {
        const int qt = scan_qgroup(A) * QTW + k;
        qb0[k] = qt * kQT;
        qon[k] = qt < nqt && !tile_skipped(A, qb0[k]);
        any |= qon[k];
    }
    if (!any) return;

    __shared__ uint32_t s_next;
    UnitCursor cur = cursor_init<NW>(A, wave, &s_next);
    const u32x4 *vb = (const u32x4 *)A.vbf;
    const uint32_t tlast = cursor_last(cur, A.ntiles);
    auto phys = [&](uint32_t t) { return min(t, tlast) * A.stride; };
    auto load_rows = [&](u32x4 (&a)[8]) {
        const u32x4 *p = vb + (uint64_t)phys(cur.tile) * 512 + l;
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = p[64 * j];
    };
    /* first rows in flight while the query operands are staged */
    u32x4 a0[8], a1[8];
    load_rows(a0);

    constexpr int half = 64;
    for (int e = tid; e < QTW * 2 * S * 64; e += blockDim.x) {
        const int k = e / (2 * S * 64), r = e % (2 * S * 64);
        const int lane = r & 63, st = (r >> 6) % S, qb = (r >> 6) / S;
        if (!qon[k]) continue;
        const int q = qb0[k] + (lane & 31) + 32 * qb;
        qsh[e] = *(const u32x4 *)(A.qbf + (uint64_t)q * 128 + (lane >> 5) * half + 8 * st);
    }
    __shared__ int s_im[INL ? QTW * kQT : 1];
    __shared__ double s_iv[INL ? QTW * kQT : 1];
    if (INL)
#pragma unroll
        for (int k = 0; k < QTW; k++) stage_inline(A, k, qb0[k], s_im, s_iv);
    if (!PREPASS)
#pragma unroll
        for (int k = 0; k < QTW; k++) cand_init(Cs[k]);

    float tq[QTW][2], qn[QTW][2];
    const uint32_t *wr[QTW][2];
#pragma unroll
    for (int k = 0; k < QTW; k++) load_query_state(A, qb0[k], tq[k], qn[k], wr[k]);

    auto load_in = [&](TileIn (&in)[QTW]) {
        const uint32_t pt = phys(cur.tile);
        in[0].t = cur.tile;
        in[0].live = cur.tile < cur.end;
        cursor_next(cur, 1, &s_next);
        load_tile_in<INL>(A, pt, wr[0], in[0]);
#pragma unroll
        for (int k = 1; k < QTW; k++) {
            in[k].w[0] = INL ? 0u : wr[k][0][pt];
            in[k].w[1] = INL ? 0u : wr[k][1][pt];
        }
    };
    auto do_unit = [&](const u32x4 (&a)[8], const TileIn (&in)[QTW]) {
        const uint32_t t = in[0].t;
        float sn[16];
        expand_norms(-0.5f * in[0].xr, sn);
        const uint32_t valid = in[0].valid;
        uint32_t iw[QTW][2];
        if (INL) {
            double cv[16];
            expand_rows(in[0].rv, cv);
#pragma unroll
            for (int k = 0; k < QTW; k++) {
                iw[k][0] = inline_word(cv, s_im, s_iv, k, 0);
                iw[k][1] = inline_word(cv, s_im, s_iv, k, 1);
            }
        }
#pragma unroll
        for (int k = 0; k < QTW; k++) {
            if (!qon[k]) continue;
            /* the tile's norms, validity and inline column values are shared
             * by the query tiles; the bitmap words are per query tile */
            f32x16 acc0, acc1;
            seed_bf(sn, (INL ? iw[k][0] : in[k].w[0]) & valid, (INL ? iw[k][1] : in[k].w[1]) & valid,
                    acc0, acc1);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const bf16x8 av = __builtin_bit_cast(bf16x8, a[j]);
                const bf16x8 b0 = __builtin_bit_cast(bf16x8, qsh[((k * 2 + 0) * S + j) * 64 + l]);
                const bf16x8 b1 = __builtin_bit_cast(bf16x8, qsh[((k * 2 + 1) * S + j) * 64 + l]);
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b0, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b1, acc1, 0, 0, 0);
            }
            epilogue_bf<PREPASS>(A, qb0[k], t, t * A.stride, acc0, acc1, qn[k], tq[k], Cs[k]);
        }
    };

    TileIn in0[QTW], in1[QTW];
    __syncthreads();
    cursor_start(cur, &s_next);
    load_in(in0);
    while (in0[0].live) {
        load_rows(a1);
        load_in(in1);
        do_unit(a0, in0);
        if (!in1[0].live) break;
        load_rows(a0);
        load_in(in0);
        do_unit(a1, in1);
    }
    if (!PREPASS)
#pragma unroll
        for (int k = 0; k < QTW; k++)
            if (qon[k]) flush_candidates(A, qb0[k], Cs[k]);
}

/*
 * Query-resident BF16 scan for batches of many query tiles (dimb 128, main
 * scan). A WG of kQrNW waves serves kQrNW query tiles, one per wave, whose
 * MFMA operands stay in VGPRs for the whole launch. The WG's row chunk
 * streams through a two-slot LDS ring in stages of kQrTps 32-row tiles:
 * global_load_lds_dwordx4 copies each 1 KB k-step chunk of a tile straight
 * into LDS (no VGPR round trip; wave w copies chunk w of every tile of the
 * stage), dword copies bring the stage's ||x||^2 and each wave's filter
 * words, and the copy of stage s + 1 runs while stage s is computed. Every
 * wave reads each tile once from LDS (eight 16-byte reads per lane, already
 * the MFMA operand layout) and runs its MFMAs against its resident queries,
 * so one HBM read of a tile feeds kQrNW x 17 MFMAs and no operand is
 * fetched per MFMA.
 *
 * The query operands are held scaled by -2 and the accumulators start from
 * ||x||^2: acc = ||x||^2 - 2 x.q and d = ||q||^2 + acc. Every scaling by -2
 * is exact, so d is the value k_scan_bf16 computes as ||q||^2 - 2 (x.q -
 * ||x||^2 / 2); a lane's 16 elements of a query reduce by min before one
 * comparison with acc <= tau - ||q||^2 (+ margin).
 *
 * Filters as one more MFMA: a wave's queries use at most kQrProgs distinct
 * predicate programs (slots k). B_pen[k][q] = 1 for the slot of query q's
 * program (0 elsewhere, all 0 for a query without predicate) and
 * A_pen[r][k] = 2^127 where row r fails program k (0 where it passes), so
 * one v_mfma_f32_32x32x16_bf16 started from ||x||^2 adds exactly 0 to the
 * rows a query's filter passes and 2^127 to every other row, which then
 * never meets the threshold (held below 2^126). A_pen is built from the
 * tile's kQrProgs filter words (16 VALU per tile for both query blocks)
 * instead of a per-element select in each block. A wave whose queries use
 * more programs, and the index's last (partial) tile, seed per element:
 * ||x||^2 where the row passes, +inf where it does not.
 *
 * Inline predicates (INL: one comparison per query on one column): the
 * index holds each row's rank in the column's posting order, staging turns
 * each comparison into a rank range [lo, hi), and the stage copy brings the
 * tiles' ranks: a row passes iff rank - lo < hi - lo (unsigned), one
 * comparison per element instead of a double-precision predicate.
 *
 * Candidates: each lane counts its own slots in a register (the wave owns
 * its 64 queries; lanes l and l + 32 hold the two row halves of a query and
 * own half of its kQrSlots slots each), so no LDS atomic is on the scan's
 * path; past its slots a lane spills to the WG's LDS spill area, then to
 * the query's global pool segment.
 */
constexpr int kQrNW = 8;                 /* waves = query tiles per WG */
constexpr int kQrTps = 4;                /* 32-row tiles per ring stage */
constexpr int kQrSlots = 16;             /* LDS candidate slots per query */
constexpr int kQrProgs = 16;             /* filter slots of the penalty MFMA */
constexpr uint32_t kQrSpill = 512;       /* WG-wide LDS spill slots */
constexpr uint16_t kBfBig = 0x7F00;      /* 2^127 in bf16 */
constexpr uint16_t kBfOne = 0x3F80;      /* 1.0 in bf16 */

struct QrCand {
    uint32_t ovf;
    uint32_t pad;
    uint64_t buf[kQT * kQrSlots];        /* query qi: [qi][half][kQrSlots / 2] */
};
struct QrShare {
    uint32_t nsp;
    uint32_t spq[kQrSpill];   /* query of the spilled key */
    uint64_t spk[kQrSpill];
};
/* dynamic LDS (candidates); the ring is static */
constexpr size_t kQrLds = kQrNW * sizeof(QrCand) + sizeof(QrShare);
static_assert(kQrLds + (size_t)2 * kQrTps * (8192 + 128 + 128) + (size_t)2 * kQrNW * kQrTps * kQrProgs * 4 <=
              160 * 1024, "query-resident scan LDS");

/* acc <= h admits every row with d = ||q||^2 + acc <= tau: the margin
 * 2^-18 (max ||x||^2 + ||q||^2) covers the rounding of tau - ||q||^2 and
 * of the two ways of forming d. Held below 2^126 so that filtered rows
 * (2^127 or +inf) never pass; no real row's acc comes near. */
__device__ __forceinline__ float qr_threshold(const ScanArgs &A, float qn, float tq) {
    return fminf((tq - qn) + ldexpf(A.xn2max + qn, -18), 0x1p126f);
}

/* seeds of a tile: ||x||^2 of the lane's 16 accumulator rows where the
 * query's filter word passes the row, +inf elsewhere */
__device__ __forceinline__ void seed_qr(const f32x16 &xn, uint32_t w0, uint32_t w1,
                                        f32x16 &a0, f32x16 &a1) {
    const uint32_t v0 = lane_word(w0), v1 = lane_word(w1);
    constexpr uint32_t kPosInf = 0x7f800000u;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t s = __float_as_uint(xn[i]);
        const uint32_t m0 = elem_mask(v0, i), m1 = elem_mask(v1, i);
        a0[i] = __uint_as_float((s & m0) | (kPosInf & ~m0));
        a1[i] = __uint_as_float((s & m1) | (kPosInf & ~m1));
    }
}

template <int NW, bool INL, bool PREPASS = false>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2)))
void k_scan_bf16_qr(ScanArgs A) {
    static_assert(NW == 8, "one 1 KB chunk of every stage tile per wave");
    static_assert(2 + kQrTps / 2 <= NW, "waves copying the stage's ranks");
    static_assert(kQrTps * kQrProgs == 64, "one dword copy of a wave's stage words");
    /* the ring's two slots are separate arrays, so the compiler can tell
     * that reads of one slot do not depend on the copy into the other
     * (no wait for the copy in flight) */
    __shared__ __attribute__((aligned(16))) u32x4 ring0[kQrTps * 512], ring1[kQrTps * 512];
    __shared__ __attribute__((aligned(16))) float nring0[kQrTps * 32], nring1[kQrTps * 32];
    __shared__ __attribute__((aligned(16))) uint32_t wring0[NW * kQrTps * kQrProgs],
        wring1[NW * kQrTps * kQrProgs];
    /* inline predicates: the stage's row ranks in the predicate's column */
    __shared__ __attribute__((aligned(16))) uint32_t rring0[INL ? kQrTps * 32 : 4],
        rring1[INL ? kQrTps * 32 : 4];
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    QrCand *Cs = (QrCand *)smem;                                  /* [NW] */
    QrShare &S = *(QrShare *)(smem + NW * sizeof(QrCand));

    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6, half = l >> 5;
    if (scan_pad(A)) return;
    const int nqt = (A.nq + kQT - 1) / kQT;
    bool any = false;
    for (int k = 0; k < NW; k++) {
        const int t = (int)scan_qgroup(A) * NW + k;
        any |= t < nqt && !tile_skipped(A, t * kQT);
    }
    if (!any) return;   /* uniform: no query tile of this WG has work */
    const int qbase = ((int)scan_qgroup(A) * NW + wave) * kQT;
    const bool on = qbase < nqt * kQT && !tile_skipped(A, qbase);

    const uint32_t nwg = scan_nchunks(A), wg = scan_chunk(A);
    const uint32_t c0 = (uint32_t)(((uint64_t)A.ntiles * wg) / nwg);
    const uint32_t c1 = (uint32_t)(((uint64_t)A.ntiles * (wg + 1)) / nwg);
    const uint32_t nst = (c1 - c0 + kQrTps - 1) / kQrTps;
    const u32x4 *vb = (const u32x4 *)A.vbf;

    /* the wave's queries: operands x -2 in VGPRs, norms, thresholds (-inf
     * for padding queries) and bitmap rows */
    bf16x8 bq[2][8];
    float tq[2], qn[2], hq[2];
    const uint32_t *wr[2];
    int32_t prog[2];
    uint32_t rlo[2] = {0u, 0u}, rlen[2] = {0u, 0u};   /* inline: rank range */
    load_query_state(A, qbase, tq, qn, wr);
#pragma unroll
    for (int qb = 0; qb < 2; qb++) {
        const int q = qbase + (l & 31) + 32 * qb;
        const bool v = on && q < A.nq;
        hq[qb] = v ? qr_threshold(A, qn[qb], tq[qb]) : -__builtin_inff();
        prog[qb] = (v && !INL) ? A.pofq[q] : -1;
        if (INL && v) {
            const InlPred ip = A.inl[q];
            rlo[qb] = ip.lo;
            rlen[qb] = ip.hi - ip.lo;
        }
#pragma unroll
        for (int s = 0; s < 8; s++) {
            u32x4 raw = {0u, 0u, 0u, 0u};
            if (v) raw = *(const u32x4 *)(A.qbf + (uint64_t)q * 128 + (l >> 5) * 64 + 8 * s);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float lo = -2.f * __uint_as_float(raw[e] << 16);
                const float hi = -2.f * __uint_as_float(raw[e] & 0xffff0000u);
                raw[e] = (__float_as_uint(lo) >> 16) | (__float_as_uint(hi) & 0xffff0000u);
            }
            bq[qb][s] = __builtin_bit_cast(bf16x8, raw);
        }
    }
    /* filter slots: the wave's distinct programs in order of first lane;
     * lane j records the program of slot j % kQrProgs (-1: unused) */
    int nslot = 0;
    int32_t myprog = -1;
    int slot_of[2] = {-1, -1};
    {
        int32_t rest0 = prog[0], rest1 = prog[1];
        while (nslot <= kQrProgs) {
            const uint64_t m0 = __ballot(rest0 >= 0), m1 = __ballot(rest1 >= 0);
            if ((m0 | m1) == 0) break;
            const int32_t p = m0 ? __shfl(rest0, __builtin_ctzll(m0))
                                 : __shfl(rest1, __builtin_ctzll(m1));
            if (nslot < kQrProgs) {
                if (prog[0] == p) slot_of[0] = nslot;
                if (prog[1] == p) slot_of[1] = nslot;
                if ((l & (kQrProgs - 1)) == nslot) myprog = p;
            }
            if (rest0 == p) rest0 = -1;
            if (rest1 == p) rest1 = -1;
            nslot++;
        }
    }
    const bool pen = !INL && nslot <= kQrProgs;   /* uniform */
    /* B_pen of the two query blocks: lane l holds k-slots 8 half .. +8 of
     * query column l & 31 */
    bf16x8 bpen[2];
#pragma unroll
    for (int qb = 0; qb < 2; qb++) {
        u32x4 raw = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (slot_of[qb] == 8 * half + j) raw[j >> 1] |= (uint32_t)kBfOne << (16 * (j & 1));
        bpen[qb] = __builtin_bit_cast(bf16x8, raw);
    }
    const uint32_t *prow = myprog >= 0 ? A.bitmap + (uint64_t)myprog * A.words : A.ones;

    /* copy of stage st into a ring slot: chunk `wave` of each of the
     * stage's tiles, the stage's norms (two tiles per dword copy) and the
     * wave's filter words [tile][slot] */
    auto load_stage = [&](uint32_t st, u32x4 *ring, float *nring, uint32_t *wring,
                          uint32_t *rring) {
        const uint32_t t0 = c0 + st * kQrTps;
#pragma unroll
        for (int i = 0; i < kQrTps; i++) {
            const uint32_t t = t0 + (uint32_t)i;
            if (t < c1)
                __builtin_amdgcn_global_load_lds(vb + (uint64_t)t * A.stride * 512 + wave * 64 + l,
                                                 ring + i * 512 + wave * 64, 16, 0, 0);
        }
        if (wave < kQrTps / 2) {
            const uint32_t t = min(t0 + 2u * wave + (uint32_t)(l >> 5), c1 - 1);
            __builtin_amdgcn_global_load_lds(A.xnorm + (uint64_t)t * A.stride * 32 + (l & 31),
                                             nring + 2 * wave * 32, 4, 0, 0);
        }
        if (pen && on) {
            const uint32_t t = min(t0 + (uint32_t)(l / kQrProgs), c1 - 1);
            __builtin_amdgcn_global_load_lds(prow + (uint64_t)t * A.stride,
                                             wring + wave * kQrTps * kQrProgs, 4, 0, 0);
        }
        if (INL && wave >= 2 && wave < 2 + kQrTps / 2) {
            const int h = wave - 2;
            const uint32_t t = min(t0 + 2u * h + (uint32_t)(l >> 5), c1 - 1);
            const uint64_t row = min((uint64_t)t * A.stride * 32 + (l & 31), A.count - 1);
            __builtin_amdgcn_global_load_lds(A.irank + row, rring + 2 * h * 32, 4, 0, 0);
        }
    };
    /* stage top: this wave's copies for the stage have landed; after the
     * barrier every wave's have, and every wave is done with the slot the
     * next copy overwrites */
    auto stage_sync = [&]() {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    };

    QrCand &C = Cs[wave];
    uint32_t cnt[2] = {0u, 0u};
    /* a candidate of block qb: the lane's own slots, then the WG's spill,
     * then the query's global pool segment */
    auto insert = [&](int qb, uint64_t key) {
        const int qi = (l & 31) + 32 * qb;
        if (cnt[qb] < (uint32_t)(kQrSlots / 2)) {
            C.buf[qi * kQrSlots + half * (kQrSlots / 2) + cnt[qb]] = key;
            cnt[qb]++;
            return;
        }
        const int q = qbase + qi;
        const uint32_t sp = atomicAdd(&S.nsp, 1u);
        if (sp < kQrSpill) {
            S.spk[sp] = key;
            S.spq[sp] = (uint32_t)q;
            return;
        }
        const uint32_t sh = scan_chunk(A) % kPoolShards;
        const uint64_t scap = shard_cap(A.cap_pool);
        const uint32_t gp = atomicAdd(&A.pcnt[pcnt_at(q, sh)], 1u);
        if (gp < scap)
            A.pool[(uint64_t)q * A.cap_pool + sh * scap + gp] = key;
        else
            C.ovf = 1u;
    };

    auto compute_stage = [&](uint32_t st, const u32x4 *ring, const float *nring,
                             const uint32_t *wring, const uint32_t *rring) {
#pragma unroll
        for (int i = 0; i < kQrTps; i++) {
            const uint32_t t = c0 + st * kQrTps + (uint32_t)i;
            if (t >= c1) break;
            const uint32_t pt = t * A.stride;
            bf16x8 a[8];
#pragma unroll
            for (int j = 0; j < 8; j++) a[j] = __builtin_bit_cast(bf16x8, ring[i * 512 + j * 64 + l]);
            /* rows g * 8 + (l >> 5) * 4 + 0..3 of accumulator group g */
            const float4 *nr = (const float4 *)(nring + i * 32);
            f32x16 xn;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const float4 v = nr[2 * g + half];
                xn[4 * g] = v.x;
                xn[4 * g + 1] = v.y;
                xn[4 * g + 2] = v.z;
                xn[4 * g + 3] = v.w;
            }
            const uint64_t row0 = (uint64_t)pt * 32;
            f32x16 acc0, acc1;
            if (INL) {
                /* rank range test per element: lo <= rank < lo + len */
                const u32x4 *rr = (const u32x4 *)(rring + i * 32);
                uint32_t rk[16];
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    const u32x4 v = rr[2 * g + half];
                    rk[4 * g] = v.x;
                    rk[4 * g + 1] = v.y;
                    rk[4 * g + 2] = v.z;
                    rk[4 * g + 3] = v.w;
                }
                if (row0 + 32 > A.count) {
                    /* rows past the index never pass */
#pragma unroll
                    for (int e = 0; e < 16; e++)
                        if (row0 + (uint64_t)acc_row(e, l) >= A.count) rk[e] = 0xFFFFFFFFu;
                }
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    acc0[e] = (rk[e] - rlo[0] < rlen[0]) ? xn[e] : __builtin_inff();
                    acc1[e] = (rk[e] - rlo[1] < rlen[1]) ? xn[e] : __builtin_inff();
                }
            } else if (pen && row0 + 32 <= A.count) {
                /* A_pen: row l & 31, slots 8 half .. +8: 2^127 where the
                 * slot's word fails the row */
                const u32x4 *wp = (const u32x4 *)(wring + wave * kQrTps * kQrProgs + i * kQrProgs +
                                                  8 * half);
                const u32x4 w0 = wp[0], w1 = wp[1];
                const int r = l & 31;
                u32x4 ap;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t lo = (j < 2 ? w0[2 * j] : w1[2 * j - 4]) >> r;
                    const uint32_t hi = (j < 2 ? w0[2 * j + 1] : w1[2 * j - 3]) >> r;
                    /* bit 0 of lo into bit 8 (low bf16 exponent field),
                     * bit 0 of hi into bit 24: pass -> 0, fail -> 0x7F00 */
                    const uint32_t pass = (lo & 1u) | ((hi & 1u) << 16);
                    ap[j] = ((pass ^ 0x00010001u) * 0x7F00u);
                }
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ap),
                                                               bpen[0], xn, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, ap),
                                                               bpen[1], xn, 0, 0, 0);
            } else {
                const uint32_t valid =
                    row0 + 32 <= A.count ? 0xFFFFFFFFu
                    : (A.count > row0 ? (0xFFFFFFFFu >> (32 - (A.count - row0))) : 0u);
                /* per-lane filter words, read here: the tile's words (L2
                 * hits after the first query tile of the chunk) */
                seed_qr(xn, wr[0][pt] & valid, wr[1][pt] & valid, acc0, acc1);
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], bq[0][j], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], bq[1][j], acc1, 0, 0, 0);
            }
#pragma unroll
            for (int qb = 0; qb < 2; qb++) {
                const f32x16 &c = qb ? acc1 : acc0;
                const float h = hq[qb];
                /* minima of elements 3g .. 3g + 2 (g < 5) and element 15,
                 * as three-input mins (v_min3_f32) */
                float gm[6];
#pragma unroll
                for (int g = 0; g < 5; g++) gm[g] = fminf(fminf(c[3 * g], c[3 * g + 1]), c[3 * g + 2]);
                gm[5] = c[15];
                const float mn =
                    fminf(fminf(fminf(fminf(gm[0], gm[1]), gm[2]), fminf(gm[3], gm[4])), gm[5]);
                if (PREPASS) {
                    /* the minimum over the lane's 16 rows (the tile's row
                     * group `half`) per query; +inf when every row of the
                     * group is filtered (2^127 or +inf) */
                    const int q = qbase + (l & 31) + 32 * qb;
                    if (q < A.nq) {
                        const float dmin = mn >= 0x1p126f ? __builtin_inff()
                                                          : fmaxf(qn[qb] + mn, 0.f);
                        const uint32_t grp = t * 2 + (uint32_t)half;
                        A.gmin[(uint64_t)q * A.ngroups + grp] = __float_as_uint(dmin);
                    }
                } else if (mn <= h) {
#pragma unroll
                    for (int g = 0; g < 6; g++)
                        if (gm[g] <= h) {
#pragma unroll
                            for (int j = 0; j < (g < 5 ? 3 : 1); j++) {
                                const int e = 3 * g + j;
                                if (c[e] <= h)
                                    insert(qb, make_key(fmaxf(qn[qb] + c[e], 0.f),
                                                        (uint32_t)(row0 + acc_row(e, l))));
                            }
                        }
                }
            }
        }
    };

    if (l == 0) C.ovf = 0;
    if (tid == 0) S.nsp = 0;
    if (nst > 0) load_stage(0, ring0, nring0, wring0, rring0);
    /* two stages per iteration: even stages in slot 0, odd in slot 1 */
    for (uint32_t st = 0; st < nst; st += 2) {
        stage_sync();
        if (st + 1 < nst) load_stage(st + 1, ring1, nring1, wring1, rring1);
        if (on) compute_stage(st, ring0, nring0, wring0, rring0);
        if (st + 1 >= nst) break;
        stage_sync();
        if (st + 2 < nst) load_stage(st + 2, ring0, nring0, wring0, rring0);
        if (on) compute_stage(st + 1, ring1, nring1, wring1, rring1);
    }

    if (PREPASS) return;
    /* flush: each lane its own slots of its two queries (one reservation
     * each), then the shared spill, into pool segment chunk % kPoolShards */
    __syncthreads();
    const uint32_t sh = scan_chunk(A) % kPoolShards;
    const uint64_t scap = shard_cap(A.cap_pool);
    if (on) {
#pragma unroll
        for (int qb = 0; qb < 2; qb++) {
            const int qi = (l & 31) + 32 * qb, q = qbase + qi;
            const uint32_t n = min(cnt[qb], (uint32_t)(kQrSlots / 2));
            if (q < A.nq && n > 0) {
                const uint32_t base = atomicAdd(&A.pcnt[pcnt_at(q, sh)], n);
                for (uint32_t j = 0; j < n; j++)
                    if (base + j < scap)
                        A.pool[(uint64_t)q * A.cap_pool + sh * scap + base + j] =
                            C.buf[qi * kQrSlots + half * (kQrSlots / 2) + j];
            }
        }
        if (C.ovf && qbase + l < A.nq) atomicOr(&A.status[qbase + l], kStOverflow);
    }
    const uint32_t ns = min(S.nsp, kQrSpill);
    for (uint32_t i = tid; i < ns; i += blockDim.x) {
        const uint32_t q = S.spq[i];
        const uint32_t gp = atomicAdd(&A.pcnt[pcnt_at(q, sh)], 1u);
        if (gp < scap) A.pool[(uint64_t)q * A.cap_pool + sh * scap + gp] = S.spk[i];
    }
}

/* the two 8-byte FP8 operands in a lane's 16 bytes of a k-step */
__device__ __forceinline__ long f8_lo(const u32x4 &v) {
    return (long)(((uint64_t)v.y << 32) | v.x);
}
__device__ __forceinline__ long f8_hi(const u32x4 &v) {
    return (long)(((uint64_t)v.w << 32) | v.z);
}

/*
 * INT8 scan (SQ8 indexes; code dims dimp = 64, 128 or 256). x.q of the
 * int8 codes accumulates in int32 with v_mfma_i32_32x32x32_i8 (exact), and
 * with the squared code norms (integers below 2^24) the epilogue sees the
 * exact code distance. Exact mode (integer data, codes x - centre): that is
 * the distance, eps = 0 and the select stage certifies as for integer BF16
 * data. Quantized mode: the code distance ranks the rows and the select
 * stage re-ranks the top overfetch x k exactly. The copy is half the bytes
 * of BF16 for the same rows. Layout: 32-row tiles, k-steps of 32 dims as
 * 1 KB chunks (slot l: row l & 31, dims 32 s + 16 (l >> 5) .. +16); a unit
 * is 8 chunks = 8 KB = PT tiles (PT = 256 / dimp), the cursor's logical
 * tile. Same pipeline as k_scan_bf16 otherwise.
 * F8 (FP8 index, i8_mode 3): the same layout holds E4M3 codes; each lane's
 * 16 bytes of a k-step feed two v_mfma_f32_32x32x16_fp8_fp8 (FP32
 * accumulation) and the tile takes the BF16 scan's seeds and epilogue over
 * the code distance ||xc||^2 + ||qc||^2 - 2 xc.qc.
 */
template <bool PREPASS, int NW, int PT, bool INL, bool F8 = false>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2)))
void k_scan_i8(ScanArgs A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr int KS = 8 / PT;               /* k-steps of 32 dims per tile */
    u32x4 *qsh = (u32x4 *)smem;              /* [2][KS][64] x 16 B */
    CandLds &C = *(CandLds *)(smem + (size_t)2 * KS * 64 * 16);

    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    if (scan_pad(A)) return;
    const int qbase = scan_qgroup(A) * kQT;
    if (tile_skipped(A, qbase)) return;

    __shared__ uint32_t s_next;
    __shared__ __attribute__((aligned(16))) int32_t s_half[NW][32]; /* seed spread */
    UnitCursor cur = cursor_init<NW>(A, wave, &s_next);
    const u32x4 *vb = (const u32x4 *)A.vi8;
    /* physical unit of a logical tile (units past the range load a valid one) */
    const uint32_t tlast = cursor_last(cur, A.ntiles);
    auto phys = [&](uint32_t t) { return min(t, tlast) * A.stride; };
    auto load_rows = [&](u32x4 (&a)[8]) {
        const u32x4 *p = vb + (uint64_t)phys(cur.tile) * 512 + l;
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = p[64 * j];
    };
    u32x4 a0[8], a1[8];
    load_rows(a0);

    for (int e = tid; e < 2 * KS * 64; e += blockDim.x) {
        const int lane = e & 63, s = (e >> 6) % KS, qb = (e >> 6) / KS;
        const int q = qbase + (lane & 31) + 32 * qb;
        qsh[e] = *(const u32x4 *)(A.qi8 + (uint64_t)q * A.dimp + 32 * s + 16 * (lane >> 5));
    }
    __shared__ int s_im[INL ? kQT : 1];
    __shared__ double s_iv[INL ? kQT : 1];
    if (INL) stage_inline(A, 0, qbase, s_im, s_iv);
    if (!PREPASS) cand_init(C);

    float tq[2], qn[2];
    const uint32_t *wr[2];
    load_query_state(A, qbase, tq, qn, wr);
    const int32_t th[2] = {i8_threshold(qn[0], tq[0]), i8_threshold(qn[1], tq[1])};
    auto load_in = [&](TileIn (&in)[PT]) {
        const uint32_t u = phys(cur.tile);
#pragma unroll
        for (int p = 0; p < PT; p++) {
            load_tile_in<INL>(A, u * PT + p, wr, in[p]);
            in[p].t = cur.tile;
            in[p].live = cur.tile < cur.end;
        }
        cursor_next(cur, 1, &s_next);
    };
    TileIn in0[PT], in1[PT];
    __syncthreads();
    cursor_start(cur, &s_next);
    load_in(in0);

    auto do_unit = [&](const u32x4 (&a)[8], const TileIn (&in)[PT]) {
        const uint32_t u = phys(in[0].t);
#pragma unroll
        for (int p = 0; p < PT; p++) {
            /* seeds: row norms and the queries' filters (epilogue_i8) */
            uint32_t w0 = in[p].w[0], w1 = in[p].w[1];
            if (INL) {
                double cv[16];
                expand_rows(in[p].rv, cv);
                w0 = inline_word(cv, s_im, s_iv, 0, 0);
                w1 = inline_word(cv, s_im, s_iv, 0, 1);
            }
            /* group index of the pre-pass: sampled tile x PT + p */
            const uint32_t tg = in[p].t * PT + p;
            const uint32_t pt = u * PT + p;
            if constexpr (F8) {
                /* FP8 codes: the BF16 scan's seeds and epilogue over the code
                 * distance; each lane's 16 bytes of a k-step are two 8-byte
                 * operands of v_mfma_f32_32x32x16_fp8_fp8 */
                float sn[16];
                expand_norms_lds(-0.5f * in[p].xr, sn, s_half[wave]);
                f32x16 c0, c1;
                seed_bf(sn, w0 & in[p].valid, w1 & in[p].valid, c0, c1);
#pragma unroll
                for (int k = 0; k < KS; k++) {
                    const u32x4 av = a[p * KS + k];
                    const u32x4 b0 = qsh[(0 * KS + k) * 64 + l];
                    const u32x4 b1 = qsh[(1 * KS + k) * 64 + l];
                    c0 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f8_lo(av), f8_lo(b0), c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f8_lo(av), f8_lo(b1), c1, 0, 0, 0);
                    c0 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f8_hi(av), f8_hi(b0), c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f8_hi(av), f8_hi(b1), c1, 0, 0, 0);
                }
                epilogue_bf<PREPASS>(A, qbase, tg, pt, c0, c1, qn, tq, C);
            } else {
                i32x16 c0, c1;
                uint32_t par;
                seed_i8(in[p].xr, w0 & in[p].valid, w1 & in[p].valid, c0, c1, par,
                        s_half[wave]);
#pragma unroll
                for (int k = 0; k < KS; k++) {
                    const i32x4 av = __builtin_bit_cast(i32x4, a[p * KS + k]);
                    const i32x4 b0 = __builtin_bit_cast(i32x4, qsh[(0 * KS + k) * 64 + l]);
                    const i32x4 b1 = __builtin_bit_cast(i32x4, qsh[(1 * KS + k) * 64 + l]);
                    c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, b0, c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, b1, c1, 0, 0, 0);
                }
                epilogue_i8<PREPASS>(A, qbase, tg, pt, c0, c1, par, qn, th, C);
            }
        }
    };

    while (in0[0].live) {
        load_rows(a1);
        load_in(in1);
        do_unit(a0, in0);
        if (!in1[0].live) break;
        load_rows(a0);
        load_in(in0);
        do_unit(a1, in1);
    }
    if (!PREPASS) flush_candidates(A, qbase, C);
}

/*
 * INT8 scan over code rows wider than one unit (dimp = 512, 768 or 1024):
 * a 32-row tile is dimp / 256 units of 8 KB, streamed as k_scan_bf16 streams
 * its units (loads of unit u + 1 issued before the 16 MFMAs of unit u; the
 * seeds at a tile's first unit, its epilogue after the last), with the
 * int32 accumulators, seeds and epilogue of k_scan_i8.
 */
template <bool PREPASS, int NW, bool INL, bool F8 = false>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2)))
void k_scan_i8w(ScanArgs A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int S = A.dimp / 32;              /* k-steps of 32 dims per tile */
    const uint32_t NCH = (uint32_t)(S / 8);  /* units per tile */
    u32x4 *qsh = (u32x4 *)smem;             /* [2][S][64] x 16 B */
    CandLds &C = *(CandLds *)(smem + (size_t)2 * S * 64 * 16);

    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    if (scan_pad(A)) return;
    const int qbase = scan_qgroup(A) * kQT;
    if (tile_skipped(A, qbase)) return;

    __shared__ uint32_t s_next;
    __shared__ __attribute__((aligned(16))) int32_t s_half[NW][32]; /* seed spread */
    UnitCursor cur = cursor_init<NW>(A, wave, &s_next);
    const u32x4 *vb = (const u32x4 *)A.vi8;
    const uint32_t tlast = cursor_last(cur, A.ntiles);
    auto phys = [&](uint32_t t) { return min(t, tlast) * A.stride; };
    auto load_rows = [&](u32x4 (&a)[8]) {
        const u32x4 *p = vb + (uint64_t)phys(cur.tile) * (2 * A.dimp) + (uint64_t)cur.c * 512 + l;
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = p[64 * j];
    };
    u32x4 a0[8], a1[8];
    load_rows(a0);

    for (int e = tid; e < 2 * S * 64; e += blockDim.x) {
        const int lane = e & 63, s = (e >> 6) % S, qb = (e >> 6) / S;
        const int q = qbase + (lane & 31) + 32 * qb;
        qsh[e] = *(const u32x4 *)(A.qi8 + (uint64_t)q * A.dimp + 32 * s + 16 * (lane >> 5));
    }
    __shared__ int s_im[INL ? kQT : 1];
    __shared__ double s_iv[INL ? kQT : 1];
    if (INL) stage_inline(A, 0, qbase, s_im, s_iv);
    if (!PREPASS) cand_init(C);

    float tq[2], qn[2];
    const uint32_t *wr[2];
    load_query_state(A, qbase, tq, qn, wr);
    const int32_t th[2] = {i8_threshold(qn[0], tq[0]), i8_threshold(qn[1], tq[1])};
    auto load_in = [&](TileIn &in) {
        load_tile_in<INL>(A, phys(cur.tile), wr, in);
        in.t = cur.tile;
        in.c = cur.c;
        in.live = cur.tile < cur.end;
        cursor_next(cur, NCH, &s_next);
    };
    TileIn in0, in1;
    __syncthreads();
    cursor_start(cur, &s_next);
    load_in(in0);
    /* accumulators: int32 (INT8 codes) or FP32 (FP8 codes) */
    using Acc = std::conditional_t<F8, f32x16, i32x16>;
    Acc c0 = {}, c1 = {};
    uint32_t par = 0;
    auto do_unit = [&](const u32x4 (&a)[8], const TileIn &in) {
        const uint32_t c = in.c;
        if (c == 0) {
            uint32_t w0 = in.w[0], w1 = in.w[1];
            if (INL) {
                double cv[16];
                expand_rows(in.rv, cv);
                w0 = inline_word(cv, s_im, s_iv, 0, 0);
                w1 = inline_word(cv, s_im, s_iv, 0, 1);
            }
            if constexpr (F8) {
                float sn[16];
                expand_norms_lds(-0.5f * in.xr, sn, s_half[wave]);
                seed_bf(sn, w0 & in.valid, w1 & in.valid, c0, c1);
            } else {
                seed_i8(in.xr, w0 & in.valid, w1 & in.valid, c0, c1, par, s_half[wave]);
            }
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int s = (int)c * 8 + j;
            if constexpr (F8) {
                const u32x4 av = a[j];
                const u32x4 b0 = qsh[(0 * S + s) * 64 + l];
                const u32x4 b1 = qsh[(1 * S + s) * 64 + l];
                c0 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f8_lo(av), f8_lo(b0), c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f8_lo(av), f8_lo(b1), c1, 0, 0, 0);
                c0 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f8_hi(av), f8_hi(b0), c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(f8_hi(av), f8_hi(b1), c1, 0, 0, 0);
            } else {
                const i32x4 av = __builtin_bit_cast(i32x4, a[j]);
                const i32x4 b0 = __builtin_bit_cast(i32x4, qsh[(0 * S + s) * 64 + l]);
                const i32x4 b1 = __builtin_bit_cast(i32x4, qsh[(1 * S + s) * 64 + l]);
                c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, b0, c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, b1, c1, 0, 0, 0);
            }
        }
        if (c == NCH - 1) {
            if constexpr (F8)
                epilogue_bf<PREPASS>(A, qbase, in.t, in.t * A.stride, c0, c1, qn, tq, C);
            else
                epilogue_i8<PREPASS>(A, qbase, in.t, in.t * A.stride, c0, c1, par, qn, th, C);
        }
    };
    while (in0.live) {
        load_rows(a1);
        load_in(in1);
        do_unit(a0, in0);
        if (!in1.live) break;
        load_rows(a0);
        load_in(in0);
        do_unit(a1, in1);
    }
    if (!PREPASS) flush_candidates(A, qbase, C);
}

/* ---------------- 3/5. FP32 MFMA scan ---------------- */

/*
 * The plain FP32 path: x.q in FP32 on v_mfma_f32_32x32x2_f32 (FP32 inputs
 * unrounded, FP32 accumulation) with the seeded epilogue of the BF16 scan
 * (acc = x.q - ||x||^2 / 2, filtered rows seeded -inf), and the same
 * certified select: the scan value differs from the reference's FP32 sum
 * only by accumulation rounding (eps of query-prep mode 3), and the exact
 * re-rank restores the reference's distances.
 *
 * Reads the row-major FP32 copy directly. A unit is 32 rows x 64 dims
 * (8 KB): lane l loads eight 16-byte pieces, piece j = dims 8 j + 4 (l >> 5)
 * .. +4 of row (l & 31); element e of piece j is the A operand of the
 * e-th of four MFMAs over that 8-dim chunk (lanes 0-31 give k = 0 = dim
 * 8 j + e, lanes 32-63 give k = 1 = dim 8 j + 4 + e), so four MFMAs sum
 * the chunk's eight products per (row, query). The queries sit in LDS in
 * the same lane layout. dim % 8 == 0; the last unit of a tile may cover
 * fewer than 8 chunks (its missing pieces are not loaded).
 * grid.x: row chunks, grid.y: query tiles; NW waves per WG.
 */
template <bool PREPASS, int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(2)))
void k_scan_f32m(ScanArgs A) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int D = A.dim;
    const int NC = D / 8;                       /* 8-dim chunks per row */
    const uint32_t NCH = (uint32_t)((NC + 7) / 8); /* units per tile */
    u32x4 *qsh = (u32x4 *)smem;                 /* [2][NC][64] x 16 B */
    CandLds &C = *(CandLds *)(smem + (size_t)2 * NC * 64 * 16);

    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    const int qbase = blockIdx.y * kQT;
    if (tile_skipped(A, qbase)) return;

    __shared__ uint32_t s_next;
    UnitCursor cur = cursor_init<NW>(A, wave, &s_next);
    const uint32_t tlast = cursor_last(cur, A.ntiles);
    auto phys = [&](uint32_t t) { return min(t, tlast) * A.stride; };
    const int half = l >> 5;
    auto load_rows = [&](u32x4 (&a)[8]) {
        const uint64_t row = (uint64_t)phys(cur.tile) * 32 + (l & 31);
        const float *rp = A.vecs + row * D + 64 * cur.c + 4 * half;
        const int nc = min(8, NC - 8 * (int)cur.c);
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (j < nc) a[j] = *(const u32x4 *)(rp + 8 * j);
    };
    u32x4 a0[8], a1[8];
    load_rows(a0);

    for (int e = tid; e < 2 * NC * 64; e += blockDim.x) {
        const int lane = e & 63, c = (e >> 6) % NC, qb = (e >> 6) / NC;
        const int q = qbase + (lane & 31) + 32 * qb;
        qsh[e] = *(const u32x4 *)(A.qf + (uint64_t)q * D + 8 * c + 4 * (lane >> 5));
    }
    if (!PREPASS) cand_init(C);

    float tq[2], qn[2];
    const uint32_t *wr[2];
    load_query_state(A, qbase, tq, qn, wr);
    auto load_in = [&](TileIn &in) {
        load_tile_in<false>(A, phys(cur.tile), wr, in);
        in.t = cur.tile;
        in.c = cur.c;
        in.live = cur.tile < cur.end;
        cursor_next(cur, NCH, &s_next);
    };
    TileIn in0, in1;
    __syncthreads();
    cursor_start(cur, &s_next);
    load_in(in0);
    auto do_unit = [&](const u32x4 (&a)[8], const TileIn &in, f32x16 &acc0, f32x16 &acc1) {
        const uint32_t c = in.c;
        if (c == 0) {
            float sn[16];
            expand_norms(-0.5f * in.xr, sn);
            seed_bf(sn, in.w[0] & in.valid, in.w[1] & in.valid, acc0, acc1);
        }
        const int nc = min(8, NC - 8 * (int)c);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (j >= nc) break;
            const int ch = (int)c * 8 + j;
            const u32x4 b0 = qsh[(0 * NC + ch) * 64 + l];
            const u32x4 b1 = qsh[(1 * NC + ch) * 64 + l];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float av = __uint_as_float(a[j][e]);
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, __uint_as_float(b0[e]), acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, __uint_as_float(b1[e]), acc1, 0, 0, 0);
            }
        }
        if (c == NCH - 1)
            epilogue_bf<PREPASS>(A, qbase, in.t, in.t * A.stride, acc0, acc1, qn, tq, C);
    };

    f32x16 acc0 = {}, acc1 = {};
    while (in0.live) {
        load_rows(a1);
        load_in(in1);
        do_unit(a0, in0, acc0, acc1);
        if (!in1.live) break;
        load_rows(a0);
        load_in(in0);
        do_unit(a1, in1, acc0, acc1);
    }
    if (!PREPASS) flush_candidates(A, qbase, C);
}

/* ---------------- 3/5. plain FP32 VALU scan ---------------- */

/*
 * The plain FP32 path: sum over dims of (x - q)^2 in FP32 (the arithmetic
 * of the reference tiled kernel, gpu_engine.cu:534-566) on the row-major
 * fp32 copy, producing the same 32-row x 64-query C layout as the MFMA path
 * so the epilogue is shared. Per 64-dim chunk the WG stages the query chunk
 * transposed (qT[64][64]) and each wave its 32 rows transposed
 * (xT[64][32+4]); all waves iterate in lockstep (barriers), idle waves
 * compute on a masked tile.
 */
constexpr int kXP = 36; /* padded row stride of xT */
template <bool PREPASS>
__global__ __launch_bounds__(256, 2) void k_scan_f32(ScanArgs A) {
    __shared__ float qT[64 * 64];
    __shared__ float xT[4][64 * kXP];
    __shared__ CandLds C;

    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    const int qbase = blockIdx.y * kQT;
    if (tile_skipped(A, qbase)) return;
    if (!PREPASS) cand_init(C);

    float tq[2], qn[2];
    const uint32_t *wr[2];
    load_query_state(A, qbase, tq, qn, wr);

    const uint32_t nwg = gridDim.x;
    const uint32_t t0 = (uint32_t)(((uint64_t)A.ntiles * blockIdx.x) / nwg);
    const uint32_t t1 = (uint32_t)(((uint64_t)A.ntiles * (blockIdx.x + 1)) / nwg);
    const uint32_t iters = (t1 - t0 + 3) / 4;
    float *xw = xT[wave];

    for (uint32_t it = 0; it < iters; it++) {
        const uint32_t t = t0 + it * 4 + wave;
        const bool live = t < t1;
        const uint32_t pt = (live ? t : t0) * A.stride;
        float acc[2][16];
#pragma unroll
        for (int qb = 0; qb < 2; qb++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[qb][i] = 0.f;
        for (int j0 = 0; j0 < A.dim; j0 += 64) {
            const int nd = min(64, A.dim - j0);
            __syncthreads();
            for (int e = tid; e < 64 * 64; e += blockDim.x) {
                const int qi = e & 63, j = e >> 6;
                float v = 0.f;
                if (j < nd) v = A.qf[(uint64_t)(qbase + qi) * A.dim + j0 + j];
                qT[j * 64 + qi] = v;
            }
            for (int e = l; e < 32 * 64; e += 64) {
                const int r = e >> 6, j = e & 63;
                float v = 0.f;
                if (live && j < nd) v = A.vecs[((uint64_t)pt * 32 + r) * A.dim + j0 + j];
                xw[j * kXP + r] = v;
            }
            __syncthreads();
            const int h = l >> 5, qi = l & 31;
            for (int j = 0; j < nd; j++) {
                const float q0 = qT[j * 64 + qi], q1 = qT[j * 64 + qi + 32];
#pragma unroll
                for (int g4 = 0; g4 < 4; g4++) {
                    const float4 xv = *(const float4 *)&xw[j * kXP + g4 * 8 + h * 4];
                    const float xs[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float e0 = xs[k] - q0, e1 = xs[k] - q1;
                        acc[0][g4 * 4 + k] = fmaf(e0, e0, acc[0][g4 * 4 + k]);
                        acc[1][g4 * 4 + k] = fmaf(e1, e1, acc[1][g4 * 4 + k]);
                    }
                }
            }
        }
        if (live) {
            TileIn in;
            load_tile_in(A, pt, wr, in);
            tile_epilogue<PREPASS>(A, qbase, t, pt, acc, in, tq, C);
        }
    }
    if (!PREPASS) flush_candidates(A, qbase, C);
}

/*
 * Gathers into sel[] every key that is <= an upper bound of the B-th
 * smallest (B <= 64) of n keys, over a 256-thread block. Bound: each wave
 * sorts its 64 lane minima; B lanes of a wave hold keys <= the wave's B-th
 * sorted minimum, so that entry bounds the B-th smallest key from above,
 * and the least of the four waves' entries is taken. Returns m, the number
 * of keys under the bound (m >= B). When m <= 64, sel[0..64) holds them
 * sorted ascending (padded with ~0) and sel[B-1] is the B-th smallest;
 * when 64 < m <= cap, sel[0..m) holds them unsorted; above cap sel is
 * partial.
 */
template <typename Get>
__device__ uint32_t block_collect_small(Get get, uint32_t n, uint32_t B, uint64_t *sel,
                                        uint32_t cap, uint32_t *nsel, uint64_t *wb) {
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    /* the first kHeld keys of each thread stay in registers for the second
     * pass (one global round trip instead of two) */
    constexpr int kHeld = 4;
    uint64_t held[kHeld];
    uint64_t mn = ~0ull;
#pragma unroll
    for (int j = 0; j < kHeld; j++) {
        const uint32_t i = t + (uint32_t)j * blockDim.x;
        held[j] = i < n ? get(i) : ~0ull;
        mn = held[j] < mn ? held[j] : mn;
    }
    for (uint32_t i = t + kHeld * blockDim.x; i < n; i += blockDim.x) {
        const uint64_t k = get(i);
        mn = k < mn ? k : mn;
    }
    const uint64_t srt = wave_sort64(mn);
    if (l == (int)B - 1) wb[w] = srt;
    if (t == 0) *nsel = 0;
    __syncthreads();
    uint64_t bound = wb[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) bound = wb[i] < bound ? wb[i] : bound;
    auto take = [&](uint64_t k) {
        if (k <= bound) {
            const uint32_t p = atomicAdd(nsel, 1u);
            if (p < cap) sel[p] = k;
        }
    };
#pragma unroll
    for (int j = 0; j < kHeld; j++)
        if (t + (uint32_t)j * blockDim.x < n) take(held[j]);
    for (uint32_t i = t + kHeld * blockDim.x; i < n; i += blockDim.x) take(get(i));
    __syncthreads();
    const uint32_t m = *nsel;
    if (m <= 64 && w == 0) {
        uint64_t x = (uint32_t)l < m ? sel[l] : ~0ull;
        x = wave_sort64(x);
        sel[l] = x;
    }
    __syncthreads();
    return m;
}

/* B-th smallest (B <= 64) of n keys: small-bound gather, then the radix
 * select over the gathered keys in LDS, or over all keys when too many */
template <typename Get>
__device__ uint64_t block_kth_small(Get get, uint32_t n, uint32_t B, uint64_t *sel,
                                    uint32_t cap, uint32_t *nsel, uint64_t *wb,
                                    uint32_t *hist, uint32_t *tmp, uint32_t *bc,
                                    uint32_t *m_out) {
    const uint32_t m = block_collect_small(get, n, B, sel, cap, nsel, wb);
    *m_out = m;
    if (m <= 64) return sel[B - 1];
    if (m <= cap) return block_kth([&](uint32_t i) { return sel[i]; }, m, B, hist, tmp, bc);
    return block_kth(get, n, B, hist, tmp, bc);
}

/* ---------------- 4. threshold from the pre-pass ---------------- */

/* One block per query: tau = B-th smallest group-minimum key's distance,
 * widened by 2 eps and a relative 2^-20 slack; also clears the pool
 * counter and status (a fill, so the pipeline can be graph-captured). */
struct TauArgs {
    const uint32_t *gmin;   /* FP32 bits of the group minima */
    uint32_t ngroups, B;
    uint32_t over;          /* candidates kept per query: over x k (at most kSelMax) */
    const float *eps;
    float *tau;
    uint32_t *pcnt;
    uint32_t *status;
    const uint8_t *skip;
    const uint32_t *topbq;
};

__device__ __forceinline__ void tau_block(const TauArgs &A, int q) {
    __shared__ uint32_t hist[256], tmp[8], bc[8];
    __shared__ uint64_t sel[1024], wb[4];
    __shared__ uint32_t nsel;
    const uint32_t *__restrict__ gmin = A.gmin;
    const uint32_t ngroups = A.ngroups;
    uint32_t B = A.B;
    const float *__restrict__ eps = A.eps;
    float *__restrict__ tau = A.tau;
    uint32_t *__restrict__ pcnt = A.pcnt;
    uint32_t *__restrict__ status = A.status;
    const uint8_t *__restrict__ skip = A.skip;
    const uint32_t *__restrict__ topbq = A.topbq;
    if (topbq) B = topbq[q];
    B = min(B * max(A.over, 1u), (uint32_t)kSelMax);
    if (skip && skip[q]) { /* routed elsewhere: no candidates, no outputs */
        if (threadIdx.x == 0) {
            tau[q] = -__uint_as_float(0x7f800000u);
            clear_pcnt(pcnt, q);
            status[q] = 0;
        }
        return;
    }
    const uint32_t *g = gmin + (uint64_t)q * ngroups;
    /* (distance, group) keys: distinct, ordered by distance first (the
     * minima are non-negative, so their bits order as the values) */
    auto key = [&](uint32_t i) { return ((uint64_t)g[i] << 32) | i; };
    float tv = __uint_as_float(0x7f800000u);
    if (ngroups >= B) {
        uint32_t m;
        const uint64_t kth =
            B <= 64 ? block_kth_small(key, ngroups, B, sel, 1024, &nsel, wb, hist, tmp, bc, &m)
                    : block_kth(key, ngroups, B, hist, tmp, bc);
        const float d = key_dist(kth);
        if (d < __uint_as_float(0x7f800000u)) {
            const double w = (double)d * (1.0 + ldexp(1.0, -20)) +
                             2.0 * (double)eps[q] + 1e-30;
            tv = (float)w;
            if ((double)tv < w) tv = nextafterf(tv, __uint_as_float(0x7f800000u));
        }
    }
    if (threadIdx.x == 0) {
        tau[q] = tv;
        clear_pcnt(pcnt, q);
        status[q] &= kStPrepFail;   /* written by this run's query prep */
    }
}

__global__ __launch_bounds__(256) void k_tau(TauArgs A) { tau_block(A, blockIdx.x); }

/* a routed run's threshold launch: blocks [0, nq) compute every query's
 * threshold (graph-routed ones included: the decision is made beside them,
 * and the scan and select read it), the rest are the routing decision's
 * blocks (four queries each) — one launch instead of two in the run's
 * chain */
__global__ __launch_bounds__(256) void k_tau_route(TauArgs A, LearnArgs R, int nq) {
    if ((int)blockIdx.x < nq) tau_block(A, blockIdx.x);
    else route_block(R, blockIdx.x - nq);
}

hipError_t launch_route_blocks(const LearnArgs &A, int nq, hipStream_t s) {
    hipLaunchKernelGGL(k_tau_route, dim3((nq + 3) / 4), dim3(256), 0, s, TauArgs{}, A, 0);
    return hipGetLastError();
}

/* clears pool counters and status, tau = +inf (gather route / no pre-pass) */
__global__ void k_reset(int nq, float *tau, uint32_t *pcnt, uint32_t *status) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) {
        tau[q] = __uint_as_float(0x7f800000u);
        clear_pcnt(pcnt, q);
        status[q] = 0;
    }
}

/* ---------------- span-gather route ---------------- */

/*
 * grid (nq, chunks). 16 lanes per enumerated row: enumeration position e ->
 * row through the query's driver spans (binary search over its spanoff
 * prefix, as in fv_brute_kernel, gpu_engine.cu:371-384) or e itself when
 * the query has no spans; predicate evaluated when need_eval; exact
 * CPU-order FP32 distance; key appended to the pool.
 */
__global__ __launch_bounds__(256) void k_gather(
    const float *__restrict__ vecs, const double *__restrict__ cols,
    const uint32_t *__restrict__ prows, uint64_t count, int dim,
    const float *__restrict__ qf, const FvGpuQueryDesc *__restrict__ qds,
    const FvGpuSpan *__restrict__ spans, const uint64_t *__restrict__ spanoff,
    const PvOp *__restrict__ ops, const int32_t *__restrict__ poff,
    const int32_t *__restrict__ plen, const int32_t *__restrict__ pofq,
    const double *__restrict__ vals, uint64_t *__restrict__ pool,
    uint32_t *__restrict__ pcnt, uint64_t cap_pool) {
    const int q = blockIdx.x;
    const FvGpuQueryDesc qd = qds[q];
    const int grp = threadIdx.x >> 4, l16 = threadIdx.x & 15;
    const int ngrp = blockDim.x >> 4;
    const float *qv = qf + (uint64_t)q * dim;
    const int p = pofq[q];
    const uint64_t total = qd.nspans > 0 ? qd.total : count;
    const uint64_t per = (total + gridDim.y - 1) / gridDim.y;
    const uint64_t e0 = (uint64_t)blockIdx.y * per;
    const uint64_t e1 = min(total, e0 + per);
    const uint64_t *soff = spanoff + (qd.nspans > 0 ? qd.soff0 : 0);
    for (uint64_t e = e0 + grp; e < e1; e += ngrp) {
        uint32_t row;
        if (qd.nspans == 0) {
            row = (uint32_t)e;
        } else {
            int lo = 0, hi = qd.nspans - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (soff[mid] <= e) lo = mid; else hi = mid - 1;
            }
            const FvGpuSpan s = spans[qd.span0 + lo];
            row = prows[(uint64_t)s.col * count + (s.lo + (uint32_t)(e - soff[lo]))];
        }
        bool pass = true;
        if (p >= 0 && (qd.need_eval || qd.nspans == 0))
            pass = eval_prog(ops + poff[p], plen[p], vals, cols, count, row);
        float d = 0.f;
        if (pass && row < count)
            d = cpu_order_l2(vecs + (uint64_t)row * dim, qv, dim, l16);
        if (pass && row < count && l16 == 0) {
            const uint32_t sh = blockIdx.y % kPoolShards;
            const uint64_t scap = shard_cap(cap_pool);
            const uint32_t gp = atomicAdd(&pcnt[pcnt_at(q, sh)], 1u);
            if (gp < scap) pool[(uint64_t)q * cap_pool + sh * scap + gp] = make_key(d, row);
        }
    }
}

/* ---------------- 6. select + exact re-rank ---------------- */

/*
 * One 256-thread block per query. Pool keys are unique (each row enters a
 * query's pool at most once). Finds the B-th smallest key kth (B =
 * min(topb, pool size)); collects every key whose distance is <= dist(kth)
 * + 2 eps (just the keys <= kth when eps == 0), re-ranks them with the
 * exact CPU-order FP32 distance, sorts by (dist, row) and emits the first
 * B. truncated = (B == topb) (gpu_engine.cu:1059-1064). Status bits flag a
 * pool overflow or a re-rank set beyond kSelMax; the host then re-runs.
 * over > 1 (quantized INT8 scan, keys hold code distances): the set is the
 * min(over x topb, kSelMax, pool size) smallest keys, all re-ranked exactly
 * (the reference SQ8 path's candidate re-rank, gpu_engine.cu:1911-2080).
 */
__global__ __launch_bounds__(256) void k_select(
    const uint64_t *__restrict__ pool, const uint32_t *__restrict__ pcnt,
    uint64_t cap_pool, uint32_t topb, uint32_t over, const float *__restrict__ eps,
    const float *__restrict__ vecs, const float *__restrict__ qf, int dim,
    FvHit *__restrict__ out, uint32_t out_stride, uint32_t *__restrict__ out_cnt,
    uint8_t *__restrict__ out_flag, uint32_t *__restrict__ status,
    const uint8_t *__restrict__ skip, const uint32_t *__restrict__ topbq) {
    __shared__ uint64_t sel[kSelMax], wb[4];
    __shared__ uint32_t hist[256], tmp[8], bc[8];
    __shared__ uint32_t nsel;
    extern __shared__ float qsh[];
    const int q = blockIdx.x, t = threadIdx.x;
    /* the query's per-query words and pool counts are requested together
     * (one round trip), before the skip test that needs one of them */
    const uint32_t tbq = topbq ? topbq[q] : topb; /* per-query k of an uneven batch */
    const bool skipped = skip && skip[q];         /* the graph search writes this query */
    const float e = eps ? eps[q] : 0.f;
    const uint64_t scap = shard_cap(cap_pool);
    uint32_t cnt[kPoolShards];
#pragma unroll
    for (uint32_t s = 0; s < kPoolShards; s++) cnt[s] = pcnt[pcnt_at(q, s)];
    if (skipped) return;
    topb = tbq;
    /* the query's pool segments: counts, their prefix, and an index map
     * i -> segment entry over the concatenation */
    const uint64_t *pq = pool + (uint64_t)q * cap_pool;
    uint32_t pre[kPoolShards + 1];
    uint32_t st = 0;
    pre[0] = 0;
#pragma unroll
    for (uint32_t s = 0; s < kPoolShards; s++) {
        const uint32_t c = cnt[s];
        if ((uint64_t)c > scap) st |= kStOverflow;
        pre[s + 1] = pre[s] + (uint32_t)min((uint64_t)c, scap);
    }
    const uint32_t n = pre[kPoolShards];
    auto pk_at = [&](uint32_t i) {
        uint32_t s = 0;
#pragma unroll
        for (uint32_t j = 1; j < kPoolShards; j++) s += i >= pre[j] ? 1u : 0u;
        return pq[s * scap + (i - pre[s])];
    };
    const uint32_t B = min(topb, n);
    /* keys selected for the re-rank (quantized scan: the overfetch set) */
    const uint32_t R = over > 1 ? min(min(topb * over, (uint32_t)kSelMax), n) : B;
    if (B == 0) {
        if (t == 0) {
            out_cnt[q] = 0;
            out_flag[q] = 0;
            status[q] |= st;
        }
        return;
    }
    /* eps = 0 (integer data inside the exactness bounds checked at query
     * prep): every scan form and the reference's FP32 sum are exact there,
     * so the pool keys already hold the reference's distances and the
     * re-rank (and its copy of the query) is skipped */
    const bool rerank = !(e == 0.f) || over > 1;
    if (rerank)
        for (int j = t; j < dim; j += blockDim.x) qsh[j] = qf[(uint64_t)q * dim + j];
    uint32_t m0 = kSelMax + 1;
    const uint64_t kth =
        R <= 64 ? block_kth_small(pk_at, n, R, sel, kSelMax, &nsel, wb, hist, tmp, bc, &m0)
                : block_kth(pk_at, n, R, hist, tmp, bc);
    float lim = key_dist(kth);
    if (e > 0.f) {
        double w = (double)lim + 2.0 * (double)e;
        lim = (float)w;
        if ((double)lim < w) lim = nextafterf(lim, __uint_as_float(0x7f800000u));
    }
    uint32_t m;
    if (e == 0.f && m0 <= 64) {
        /* the R smallest keys are already sorted at the front of sel */
        m = R;
    } else {
        __syncthreads();
        if (t == 0) nsel = 0;
        __syncthreads();
        for (uint32_t i = t; i < n; i += blockDim.x) {
            const uint64_t k = pk_at(i);
            const bool take = (e > 0.f) ? (key_dist(k) <= lim || k <= kth) : (k <= kth);
            if (take) {
                const uint32_t pos = atomicAdd(&nsel, 1u);
                if (pos < (uint32_t)kSelMax) sel[pos] = k;
            }
        }
        __syncthreads();
        m = nsel;
    }
    if (m > (uint32_t)kSelMax) {
        if (t == 0) {
            status[q] |= st | kStCertFail;
            out_cnt[q] = 0;
            out_flag[q] = 0;
        }
        return;
    }
    /* exact re-rank: 16 lanes per key */
    if (rerank) {
        const int grp = t >> 4, l16 = t & 15;
        for (uint32_t i = grp; i < m; i += blockDim.x / 16) {
            const uint32_t row = key_row(sel[i]);
            const float d = cpu_order_l2(vecs + (uint64_t)row * dim, qsh, dim, l16);
            if (l16 == 0) sel[i] = make_key(d, row);
        }
    }
    __syncthreads();
    if (m <= 64) {
        if (t < 64) {
            uint64_t x = (uint32_t)t < m ? sel[t] : ~0ull;
            x = wave_sort64(x);
            sel[t] = x;
        }
        __syncthreads();
    } else {
        uint32_t np2 = 1;
        while (np2 < m) np2 <<= 1;
        for (uint32_t i = m + t; i < np2; i += blockDim.x) sel[i] = ~0ull;
        __syncthreads();
        block_bitonic(sel, np2);
    }
    for (uint32_t i = t; i < B; i += blockDim.x) {
        FvHit h;
        h.node = key_row(sel[i]);
        h.dist = key_dist(sel[i]);
        out[(uint64_t)q * out_stride + i] = h;
    }
    if (t == 0) {
        out_cnt[q] = B;
        out_flag[q] = (B == topb) ? 1 : 0;
        status[q] |= st;
    }
}

/* ---------------- host side ---------------- */

static inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

void free_batch_buffers(FvGpuBatch *b) {
    b->h_stage.release();
    b->d_stage.release();
    b->d_qbf.release();
    b->d_qaux.release();
    b->d_bitmap.release();
    b->d_gmin.release();
    b->d_pool.release();
    b->d_pcnt.release();
    b->d_out.release();
    b->d_cnt.release();
    b->d_flag.release();
    b->h_out.release();
    b->d_visit.release();
    b->d_fspill.release();
    b->d_progcnt.release();
    b->d_feat.release();
    b->d_y.release();
    b->d_route.release();
    b->d_efq.release();
    b->d_qi8.release();
    b->d_qn8.release();
    b->d_vmap.release();
    if (b->aux_stream) {
        hipStreamSynchronize(b->aux_stream);
        hipStreamDestroy(b->aux_stream);
        b->aux_stream = nullptr;
    }
    if (b->ev_fork) hipEventDestroy(b->ev_fork);
    if (b->ev_join) hipEventDestroy(b->ev_join);
    b->ev_fork = b->ev_join = nullptr;
    if (b->view) {
        if (b->owner && b->owner->last_batch == b->view) b->owner->last_batch = nullptr;
        free_batch_buffers(b->view);   /* its bitmap is borrowed: not freed */
        delete b->view;
        b->view = nullptr;
    }
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->gsig.clear();
    b->warmed = false;
    delete (ScanLaunch *)b->last_scan;
    b->last_scan = nullptr;
}

int stage_batch(FvGpuIndex *g, FvGpuBatch *b, int nq, const float *queries,
                const FvGpuQueryDesc *qd, const uint8_t *dedup,
                const FvGpuPredNode *nodes, int nnodes,
                const int32_t *children, int nchildren, const double *invals,
                int ninvals, const FvGpuSpan *spans, int nspans_total,
                const uint64_t *spanoff, int nspanoff, uint32_t topb,
                char *errbuf, size_t errlen, const uint32_t *topb_q, int max_nq,
                bool keep_layout) {
    if (nq < 1 || nq > max_nq || topb == 0 || topb > FV_GPU_MAX_TOPB ||
        nspans_total > FV_GPU_MAX_SPANS * nq)
        PV_FAIL("gpu: bad brute batch (nq=%d topb=%u)", nq, topb);
    if (topb_q)
        for (int q = 0; q < nq; q++)
            if (topb_q[q] == 0 || topb_q[q] > topb)
                PV_FAIL("gpu: query %d: k %u outside 1..%u", q, topb_q[q], topb);
    if (queries == nullptr || qd == nullptr) PV_FAIL("gpu: NULL queries or descriptors");
    if (b->stream == nullptr) {
        b->stream = g->bstreams[g->next_stream];
        g->next_stream = (g->next_stream + 1) % g->nstreams;
    }
    ProgSet ps;
    if (compile_programs(nq, qd, nodes, nnodes, children, nchildren, invals,
                         ninvals, g->ncols, &ps, errbuf, errlen) != 0)
        return -1;
    /* validate spans (used by the gather route) */
    b->total.resize(nq);
    b->nspans.resize(nq);
    b->all_spans = true;
    b->any_dedup = false;
    b->sum_total = 0;
    b->max_total = 0;
    for (int q = 0; q < nq; q++) {
        const FvGpuQueryDesc &d = qd[q];
        b->total[q] = d.nspans > 0 ? d.total : g->count;
        b->nspans[q] = d.nspans;
        if (d.nspans > 0) {
            if (spans == nullptr || spanoff == nullptr || d.span0 < 0 ||
                (int64_t)d.span0 + d.nspans > nspans_total || d.soff0 < 0 ||
                (int64_t)d.soff0 + d.nspans + 1 > nspanoff)
                PV_FAIL("gpu: query %d: span range out of bounds", q);
            for (int s = 0; s < d.nspans; s++) {
                const FvGpuSpan &sp = spans[d.span0 + s];
                if (sp.col < 0 || sp.col >= g->ncols || sp.lo > sp.hi ||
                    sp.hi > g->count)
                    PV_FAIL("gpu: query %d: bad span", q);
                if (spanoff[d.soff0 + s + 1] - spanoff[d.soff0 + s] != (uint64_t)(sp.hi - sp.lo))
                    PV_FAIL("gpu: query %d: span prefix mismatch", q);
            }
            if (spanoff[d.soff0] != 0 || spanoff[d.soff0 + d.nspans] != d.total)
                PV_FAIL("gpu: query %d: span total mismatch", q);
        }
        if (d.nspans == 0 || (dedup && dedup[q]) || g->d_prows == nullptr) b->all_spans = false;
        if (dedup && dedup[q]) b->any_dedup = true;
        b->sum_total += b->total[q];
        b->max_total = std::max(b->max_total, b->total[q]);
    }

    b->owner = g;
    b->nq = nq;
    b->nqt = (nq + kQT - 1) / kQT;
    b->topb = topb;
    b->max_entry = 0;
    for (int q = 0; q < nq; q++) b->max_entry = std::max<uint64_t>(b->max_entry, qd[q].entry);
    b->nprog = (int)ps.off.size();
    b->nops = (int)ps.ops.size();
    b->nvals = (int)ps.vals.size();
    const int nqp = b->nqt * kQT;
    /* inline predicates: possible when every predicate is one comparison
     * on a single column */
    b->inl_ok = true;
    b->inl_col = -1;
    for (int q = 0; q < nq && b->inl_ok; q++) {
        const int p = ps.of_query[q];
        if (p < 0) continue;
        const PvOp &op = ps.ops[ps.off[p]];
        if (ps.len[p] != 1 || op.kind != OP_CMP || (b->inl_col >= 0 && op.col != b->inl_col))
            b->inl_ok = false;
        else
            b->inl_col = op.col;
    }
    if (!b->inl_ok) b->inl_col = -1;
    b->has_topbq = topb_q != nullptr;
    const int dim = g->dim;

    /* packed staging layout: the regions sized by the query count first,
     * then the predicate and span regions at capacities of at least twice
     * their size (powers of two); a restage keeps the batch's layout when
     * its contents fit, so the device pointers stay the same */
    const size_t n_ops = ps.ops.size(), n_vals = ps.vals.size(), n_prog = ps.off.size();
    const size_t n_sp = (size_t)std::max(nspans_total, 0), n_so = (size_t)std::max(nspanoff, 0);
    StageLayout &L = b->lay;
    const bool fits = keep_layout && L.nq == nq && L.tq == (topb_q != nullptr) &&
                      n_ops <= L.c_ops && n_vals <= L.c_vals && n_prog <= L.c_prog &&
                      n_sp <= L.c_sp && n_so <= L.c_so;
    if (!fits) {
        auto cap = [](size_t n) {
            size_t c = 16;
            while (c < 2 * n) c *= 2;
            return c;
        };
        L = StageLayout();
        L.nq = nq;
        L.tq = topb_q != nullptr;
        L.c_ops = cap(n_ops);
        L.c_vals = cap(n_vals);
        L.c_prog = cap(n_prog);
        L.c_sp = cap(n_sp);
        L.c_so = cap(n_so);
        size_t off = 0;
        L.o_q = off; off = al256(off + (size_t)nqp * dim * 4);
        L.o_qd = off; off = al256(off + (size_t)nq * sizeof(FvGpuQueryDesc));
        L.o_hdr = off; off = al256(off + 3 * sizeof(int32_t));
        L.o_pofq = off; off = al256(off + (size_t)nqp * 4);
        L.o_inl = off; off = al256(off + (size_t)nqp * sizeof(InlPred));
        L.o_tq = off; off = al256(off + (topb_q ? (size_t)nq * 4 : 0));
        L.o_ops = off; off = al256(off + L.c_ops * sizeof(PvOp));
        L.o_vals = off; off = al256(off + L.c_vals * 8);
        L.o_poff = off; off = al256(off + L.c_prog * 4);
        L.o_plen = off; off = al256(off + L.c_prog * 4);
        L.o_sp = off; off = al256(off + L.c_sp * sizeof(FvGpuSpan));
        L.o_so = off; off = al256(off + L.c_so * 8);
        L.total = off;
    }
    const size_t total_bytes = L.total;
    PV_TRY(b->h_stage.ensure(total_bytes));
    PV_TRY(b->d_stage.ensure(total_bytes));
    unsigned char *h = b->h_stage.as<unsigned char>();
    memcpy(h + L.o_q, queries, (size_t)nq * dim * 4);
    memset(h + L.o_q + (size_t)nq * dim * 4, 0, (size_t)(nqp - nq) * dim * 4);
    memcpy(h + L.o_qd, qd, (size_t)nq * sizeof(FvGpuQueryDesc));
    {
        int32_t *hdr = (int32_t *)(h + L.o_hdr);
        hdr[0] = b->nprog;
        hdr[1] = b->nops;
        hdr[2] = b->nvals;
    }
    if (n_ops) memcpy(h + L.o_ops, ps.ops.data(), n_ops * sizeof(PvOp));
    if (n_vals) memcpy(h + L.o_vals, ps.vals.data(), n_vals * 8);
    if (n_prog) memcpy(h + L.o_poff, ps.off.data(), n_prog * 4);
    if (n_prog) memcpy(h + L.o_plen, ps.len.data(), n_prog * 4);
    {
        int32_t *pofq = (int32_t *)(h + L.o_pofq);
        for (int q = 0; q < nqp; q++) pofq[q] = q < nq ? ps.of_query[q] : -1;
    }
    if (n_sp) memcpy(h + L.o_sp, spans, n_sp * sizeof(FvGpuSpan));
    if (n_so) memcpy(h + L.o_so, spanoff, n_so * 8);
    b->inl_rank = false;
    if (b->inl_ok) {
        InlPred *ip = (InlPred *)(h + L.o_inl);
        /* the rank range of a comparison: less than lb, lb .. ub equal,
         * ub .. nn greater (nn non-NULL values), NULLs from nn on */
        const std::vector<double> *sv =
            (g->d_rank != nullptr && b->inl_col >= 0 && b->inl_col < (int)g->h_sorted.size())
                ? &g->h_sorted[(size_t)b->inl_col] : nullptr;
        b->inl_rank = sv != nullptr;
        for (int q = 0; q < nqp; q++) {
            const int p = q < nq ? ps.of_query[q] : -1;
            memset(&ip[q], 0, sizeof(InlPred));
            ip[q].col = -1;
            ip[q].hi = (uint32_t)g->count;
            if (p >= 0) {
                const PvOp &op = ps.ops[ps.off[p]];
                ip[q].col = op.col;
                ip[q].mask = op.b;
                ip[q].value = op.value;
                if (sv != nullptr) {
                    const double c = op.value;
                    const uint32_t nn = (uint32_t)sv->size();
                    uint32_t lb = 0, ub = 0;
                    if (c == c) {
                        lb = (uint32_t)(std::lower_bound(sv->begin(), sv->end(), c) - sv->begin());
                        ub = (uint32_t)(std::upper_bound(sv->begin(), sv->end(), c) - sv->begin());
                    }
                    uint32_t lo = 0, hi = 0;
                    bool one = true;
                    if (c != c) {
                        /* no order against a NULL constant: only != passes
                         * (every non-NULL row) */
                        if ((op.b >> 4) & 1) hi = nn;
                    } else {
                        switch (op.b) {
                            case 0x01: hi = lb; break;                 /* <  */
                            case 0x03: hi = ub; break;                 /* <= */
                            case 0x02: lo = lb; hi = ub; break;        /* == */
                            case 0x06: lo = lb; hi = nn; break;        /* >= */
                            case 0x04: lo = ub; hi = nn; break;        /* >  */
                            default: one = false;                      /* != */
                        }
                    }
                    ip[q].lo = lo;
                    ip[q].hi = hi;
                    b->inl_rank = b->inl_rank && one;
                }
            }
        }
    }
    if (topb_q) memcpy(h + L.o_tq, topb_q, (size_t)nq * 4);
    PV_TRY(hipMemcpyAsync(b->d_stage.p, h, total_bytes, hipMemcpyHostToDevice, b->stream));
    unsigned char *d = b->d_stage.as<unsigned char>();
    b->d_q = (float *)(d + L.o_q);
    b->d_qd = (FvGpuQueryDesc *)(d + L.o_qd);
    b->d_hdr = (const int32_t *)(d + L.o_hdr);
    b->d_ops = (PvOp *)(d + L.o_ops);
    b->d_vals = (double *)(d + L.o_vals);
    b->d_poff = (int32_t *)(d + L.o_poff);
    b->d_plen = (int32_t *)(d + L.o_plen);
    b->d_pofq = (int32_t *)(d + L.o_pofq);
    b->d_spans = (FvGpuSpan *)(d + L.o_sp);
    b->d_spanoff = (uint64_t *)(d + L.o_so);
    b->d_inl_p = b->inl_ok ? (const InlPred *)(d + L.o_inl) : nullptr;
    b->d_topbq_p = topb_q ? (const uint32_t *)(d + L.o_tq) : nullptr;

    PV_TRY(b->d_qbf.ensure((size_t)nqp * qbf_dim(g) * 2));
    PV_TRY(b->d_qaux.ensure((size_t)nqp * 16));
    /* INT8 scan eligibility: every query value an integer inside the
     * index's INT8 window [centre - 128, centre + 127] */
    b->i8_ok = false;
    if (g->d_vi8 != nullptr) {
        bool ok = true;
        if (g->i8_mode == 1) {
            const float lo = (float)g->i8_center - 128.f, hi = (float)g->i8_center + 127.f;
            for (size_t i = 0, n = (size_t)nq * dim; i < n && ok; i++) {
                const float v = queries[i];
                ok = v == rintf(v) && v >= lo && v <= hi;
            }
        }
        b->i8_ok = ok;
        PV_TRY(b->d_qi8.ensure((size_t)nqp * g->dimp));
        PV_TRY(b->d_qn8.ensure((size_t)nqp * 4));
    }
    /* bitmaps and per-program counts for the layout's program capacity
     * (bitmaps: at most kBmProgs), so a restaged batch keeps the buffers */
    const size_t pcap = b->arena_bound ? (size_t)b->nprog : std::min<size_t>(L.c_prog, kBmProgs);
    if (b->nprog > 0 && b->nprog <= kBmProgs)
        PV_TRY(b->d_bitmap.ensure(std::max<size_t>(pcap, b->nprog) * (g->count_pad / 32) * 4));
    b->out_stride = topb;
    PV_TRY(b->d_out.ensure((size_t)nq * topb * sizeof(FvHit)));
    PV_TRY(b->d_cnt.ensure((size_t)nq * 4));
    PV_TRY(b->d_flag.ensure((size_t)nq));
    PV_TRY(b->d_pcnt.ensure((size_t)nqp * kPoolShards * kPcntStride * 4));
    /* per-program, per-bitmap-block passing-row counts of routed runs
     * (every run's bitmap pass writes all of them) */
    PV_TRY(b->d_progcnt.ensure(std::max<size_t>({(size_t)1, pcap, (size_t)b->nprog}) *
                               bitmap_blocks(g) * 4));
    return 0;
}

/* scan precision actually used for a request */
/* the scan a run of batch b takes: SQ8 = the exact INT8 scan (index holds
 * the INT8 copy and every query of the batch fits its window), else the
 * BF16 scan, else plain FP32 */
int scan_precision(const FvGpuIndex *g, const FvGpuBatch *b, const FvGpuBruteOpts *o) {
    const bool asked = o && o->precision >= 0;
    int p = asked ? o->precision : g->precision;
    /* an FP32 index is served by the BF16 scan unless FP32 is asked for:
     * the certified re-rank returns the same ids and distances */
    if (!asked && p == FV_GPU_PREC_FP32) p = FV_GPU_PREC_BF16;
    /* FP8 codes sit in the index's INT8-layout copy (i8_mode 3): the SQ8
     * pipeline scans them with the FP8 MFMA */
    if (p == FV_GPU_PREC_FP8) p = FV_GPU_PREC_SQ8;
    if (p == FV_GPU_PREC_SQ8 && !(g->d_vi8 != nullptr && b != nullptr && b->i8_ok))
        p = FV_GPU_PREC_BF16;
    /* a batch of kQrNW or more query tiles at 128 dims that the call does
     * not pin to a precision takes the query-resident BF16 scan also on an
     * SQ8 / FP8 index: exact results, and faster there than the code scan,
     * which re-reads the code rows once per query tile */
    if (!asked && p == FV_GPU_PREC_SQ8 && b != nullptr && b->nqt >= kQrNW &&
        g->d_vbf != nullptr && g->dimb == 128) {
        const bool inl_req = o && o->inline_pred && b->inl_ok && b->inl_col >= 0;
        if (!inl_req || (b->inl_rank && g->d_rank != nullptr)) p = FV_GPU_PREC_BF16;
    }
    if (p == FV_GPU_PREC_BF16 && (g->d_vbf == nullptr || g->dimb % 128 != 0 || g->dimb > 1024))
        p = FV_GPU_PREC_FP32;
    if (p == FV_GPU_PREC_FP32 && (g->dim % 8 != 0 || g->dim > 512)) p = FV_GPU_PREC_FP32_VALU;
    if (p != FV_GPU_PREC_FP32 && p != FV_GPU_PREC_SQ8 && p != FV_GPU_PREC_BF16)
        p = FV_GPU_PREC_FP32_VALU;
    return p;
}

/* BF16 scan launch: waves per WG (4 / 8) and inline predicates pick the
 * kernel instance */
/* grid of a scan launch: (chunks, query groups), or the 1D query-major
 * grid (scan_chunk) with its chunk and group counts set in Q */
static dim3 scan_grid(ScanArgs &Q, uint32_t gx, uint32_t groups) {
    if (!Q.qmajor) return dim3(gx, groups);
    Q.nchunk = gx;
    Q.ngroup = groups;
    return dim3(kXcds * groups * ((gx + kXcds - 1) / kXcds));
}

template <bool PREPASS>
static void launch_scan_bf16(int nw, bool inl, uint32_t gx, int nqt, size_t shm,
                             hipStream_t s, const ScanArgs &A0) {
    ScanArgs A = A0;
    const dim3 grid = scan_grid(A, gx, (uint32_t)nqt);
    if (nw == 4 && !inl) hipLaunchKernelGGL((k_scan_bf16<PREPASS, 4, false>), grid, dim3(256), shm, s, A);
    else if (nw == 4) hipLaunchKernelGGL((k_scan_bf16<PREPASS, 4, true>), grid, dim3(256), shm, s, A);
    else if (!inl) hipLaunchKernelGGL((k_scan_bf16<PREPASS, 8, false>), grid, dim3(512), shm, s, A);
    else hipLaunchKernelGGL((k_scan_bf16<PREPASS, 8, true>), grid, dim3(512), shm, s, A);
}

/* multi-query-tile BF16 scan launch (dim 128): QTW 2 -> 4 waves per WG,
 * QTW 4 -> 8 waves per WG (one WG fills a CU's LDS) */
template <bool PREPASS>
static hipError_t launch_scan_mq(int qtw, bool inl, uint32_t gx, int gy, hipStream_t s,
                                 const ScanArgs &A0) {
    ScanArgs A = A0;
    const size_t shm = (size_t)qtw * 2 * 8 * 64 * 16 + (size_t)qtw * sizeof(CandLds);
    const void *f;
    if (qtw == 2) f = inl ? (const void *)k_scan_bf16_mq<PREPASS, 2, 4, true>
                          : (const void *)k_scan_bf16_mq<PREPASS, 2, 4, false>;
    else f = inl ? (const void *)k_scan_bf16_mq<PREPASS, 4, 8, true>
                 : (const void *)k_scan_bf16_mq<PREPASS, 4, 8, false>;
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
    const dim3 grid = scan_grid(A, gx, (uint32_t)gy);
    if (qtw == 2 && inl) hipLaunchKernelGGL((k_scan_bf16_mq<PREPASS, 2, 4, true>), grid, dim3(256), shm, s, A);
    else if (qtw == 2) hipLaunchKernelGGL((k_scan_bf16_mq<PREPASS, 2, 4, false>), grid, dim3(256), shm, s, A);
    else if (inl) hipLaunchKernelGGL((k_scan_bf16_mq<PREPASS, 4, 8, true>), grid, dim3(512), shm, s, A);
    else hipLaunchKernelGGL((k_scan_bf16_mq<PREPASS, 4, 8, false>), grid, dim3(512), shm, s, A);
    return hipGetLastError();
}

/* query-resident BF16 scan launch (main scan, dim 128, bitmap predicates):
 * kQrNW query tiles per WG, gy = query groups */
template <bool PREPASS>
static hipError_t launch_scan_qr(uint32_t gx, int gy, bool inl, hipStream_t s,
                                 const ScanArgs &A0) {
    ScanArgs A = A0;
    const void *f = inl ? (const void *)k_scan_bf16_qr<kQrNW, true, PREPASS>
                        : (const void *)k_scan_bf16_qr<kQrNW, false, PREPASS>;
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kQrLds);
    if (e != hipSuccess) return e;
    const dim3 grid = scan_grid(A, gx, (uint32_t)gy);
    if (inl)
        hipLaunchKernelGGL((k_scan_bf16_qr<kQrNW, true, PREPASS>), grid, dim3(kQrNW * 64), kQrLds,
                           s, A);
    else
        hipLaunchKernelGGL((k_scan_bf16_qr<kQrNW, false, PREPASS>), grid, dim3(kQrNW * 64), kQrLds,
                           s, A);
    return hipGetLastError();
}

/* Phase 1 of a run: predicate bitmaps (when `bitmap`) and the query
 * preparation for scan precision `prec`, one launch on stream s. */
hipError_t launch_phase1(FvGpuIndex *g, FvGpuBatch *b, int prec, bool bitmap, hipStream_t s,
                         uint32_t *progpart, const PrepFeatures *pf) {
    if (bitmap && b->nprog > 0 && (b->nprog > kBmProgs || b->d_bitmap.p == nullptr))
        return hipErrorInvalidValue;
    QAux qa = qaux_of(b);
    {
        /* predicate programs -> row bitmaps, every run (the gather route
         * evaluates the predicate on the enumerated rows instead), fused
         * with the query preparation */
        Phase1Args P = {};
        P.ops = b->d_ops;
        P.poff = b->d_poff;
        P.plen = b->d_plen;
        P.vals = b->d_vals;
        P.cols = g->d_cols;
        P.count = g->count;
        P.words = g->count_pad / 32;
        P.bitmap = b->d_bitmap.as<uint32_t>();
        P.progpart = progpart;
        P.hdr = b->d_hdr;
        P.ncols = g->ncols;
        P.nbm = (bitmap && b->nprog > 0) ? bitmap_blocks(g) : 0;
        P.q = b->d_q;
        P.nq = b->nq;
        P.nqp = b->nqt * kQT;
        P.dim = g->dim;
        P.dimb = qbf_dim(g);
        P.qbf = b->d_qbf.as<uint16_t>();
        P.qnorm = qa.qnorm;
        P.eps = qa.eps;
        P.xmax_norm = g->xmax_norm;
        P.vmin = g->vmin;
        P.vmax = g->vmax;
        P.all_int = g->all_int ? 1 : 0;
        P.mode = prec == FV_GPU_PREC_SQ8 ? 2
                 : prec == FV_GPU_PREC_BF16 ? 1
                 : prec == FV_GPU_PREC_FP32 ? 3 : 0;
        P.qi8 = b->d_qi8.as<int8_t>();
        P.qn8 = b->d_qn8.as<float>();
        P.centre = g->i8_center;
        P.i8mid = g->i8_mode >= 2 ? g->d_i8mid : nullptr;
        P.fp8 = g->i8_mode == 3 ? 1 : 0;
        P.i8inv = g->i8_inv;
        P.dimp = g->dimp;
        P.status = qa.status;
        if (pf != nullptr && g->dim <= 256) {
            P.cent_t = pf->cent_t;
            P.ncent = pf->ncent;
            P.nfeat = pf->nfeat;
            P.feat = pf->feat;
        }
        /* query prep: four queries per block, or one per block with the
         * router features */
        P.nprep = P.feat ? (uint32_t)P.nqp : (uint32_t)(P.nqp + 3) / 4;
        hipLaunchKernelGGL(k_phase1, dim3(P.nbm + P.nprep), dim3(256),
                           P.nbm ? bitmap_lds_bytes(P.ncols) : 0, s, P);
        return hipGetLastError();
    }
}

/* Launch shape of the full-scan route: query tiles per WG, waves per WG,
 * WGs per CU, grid and the pre-pass sampling stride. The stride is the
 * over-fetch factor: a tau at the B-th smallest minimum of groups sampled
 * every stride-th tile admits about stride x B candidates per query. */
struct ScanShape {
    int qtw, nw, wpc, gy;
    uint32_t gx, stride;
};

/* 32-row tiles per scan unit: the INT8 scan's 8 KB unit spans 256 / dimp
 * tiles (code rows of up to 256 dims) */
static inline uint32_t scan_unit_tiles(const FvGpuIndex *g, int prec) {
    return (prec == FV_GPU_PREC_SQ8 && g->dimp <= 256) ? (uint32_t)(256 / g->dimp) : 1u;
}

/* true when another batch stream of the index still has work queued: the
 * scan then shares the chip with those batches' kernels */
static bool other_batches_in_flight(const FvGpuIndex *g, const FvGpuBatch *b) {
    if (b->capturing) return false;
    for (int i = 0; i < g->nstreams; i++)
        if (g->bstreams[i] != b->stream && hipStreamQuery(g->bstreams[i]) == hipErrorNotReady)
            return true;
    return false;
}

/* candidates per query the scan must keep: topb, or for the quantized INT8
 * scan the overfetch x topb rows that the select stage re-ranks exactly
 * (FvGpuBruteOpts.overfetch, default 4), at most kSelMax */
static uint32_t scan_overfetch(const FvGpuIndex *g, int prec, const FvGpuBruteOpts *o) {
    if (!(prec == FV_GPU_PREC_SQ8 && g->i8_mode >= 2)) return 1;
    const uint32_t f = (o && o->overfetch) ? o->overfetch : 4u;
    return std::min<uint32_t>(f, 64);
}
static uint32_t scan_candidates(const FvGpuIndex *g, const FvGpuBatch *b, int prec,
                                const FvGpuBruteOpts *o) {
    return std::min<uint32_t>(b->topb * scan_overfetch(g, prec, o), (uint32_t)kSelMax);
}

static ScanShape scan_shape(const FvGpuIndex *g, const FvGpuBatch *b, const FvGpuBruteOpts *o,
                            int prec, uint32_t stride_override, bool shared = false) {
    ScanShape sh;
    const int dim = g->dim;
    const uint32_t B = scan_candidates(g, b, prec, o);
    /* scan tiles: 32 rows, or for the INT8 scan its 8 KB units */
    const uint32_t ntiles = (uint32_t)(g->count_pad / 32) / scan_unit_tiles(g, prec);
    const bool bf = prec == FV_GPU_PREC_BF16 || prec == FV_GPU_PREC_SQ8 ||
                    prec == FV_GPU_PREC_FP32;
    /* query tiles per WG: several when many query tiles re-read the same
     * data (dim 128, BF16) */
    /* two query tiles per 4-wave WG measured fastest for multi-tile
     * batches at every WG count and tile order (1024 queries at 1M x 128:
     * 0.35-0.39 ms against 0.43-0.45 ms with four tiles per 8-wave WG) */
    /* kQrNW (8): the query-resident kernel, kQrNW query tiles per WG, for
     * batches of at least kQrNW tiles whose predicates are bitmaps or
     * inline rank ranges */
    const bool inl_req = o && o->inline_pred && b->inl_ok && b->inl_col >= 0;
    const bool qr_ok = prec == FV_GPU_PREC_BF16 && g->dimb == 128 &&
                       (!inl_req || (b->inl_rank && g->d_rank != nullptr));
    int qtw = (o && o->qtw > 0) ? o->qtw
                                : (qr_ok && b->nqt >= kQrNW ? kQrNW : (b->nqt >= 2 ? 2 : 1));
    if (prec != FV_GPU_PREC_BF16 || g->dimb != 128 || b->nqt < 2) qtw = 1;
    qtw = qtw >= kQrNW ? (qr_ok ? kQrNW : 4) : qtw >= 4 ? 4 : (qtw >= 2 ? 2 : 1);
    sh.qtw = qtw;
    sh.gy = (b->nqt + qtw - 1) / qtw;
    if (qtw == kQrNW) {
        /* one WG per CU (the LDS ring and candidate slots fill most of a
         * CU's LDS); every WG's chunk spans many ring stages */
        sh.nw = kQrNW;
        sh.wpc = 1;
        uint32_t x = (uint32_t)std::max(32, g->ncu / std::max(1, sh.gy));
        if (o && o->wg_per_cu > 0)
            x = (uint32_t)std::max(8, (std::min(o->wg_per_cu, 4) * g->ncu) / std::max(1, sh.gy));
        sh.gx = std::min<uint32_t>(x, std::max<uint32_t>(1, ntiles / (4 * kQrTps)));
        uint32_t stride = stride_override ? stride_override
                                          : (o && o->sample_stride ? o->sample_stride : 0);
        if (stride == 0) {
            /* about B x stride candidates per query: each costs the scan a
             * pass through the insertion path, each halving of the stride
             * doubles the pre-pass. sqrt(2560 / B) balances the two at 1M x
             * 128 with 1024 queries (measured optimum 16 at B = 10: 317 us
             * per batch against 365 us at 64 and 337 us at 8), within the
             * per-WG slots: about B x stride / gx candidates per query per
             * WG against the query's kQrSlots (its two lane halves split
             * the rows evenly). At B = 64 with 32 chunks that allows 6
             * (2.74M QPS on 4096-query batches at 1M x 128, against 2.47M
             * at 4 and 2.67M at 8); past the slots a WG spills, exactly */
            stride = (uint32_t)std::lround(std::sqrt(2560.0 / std::max<uint32_t>(1, B)));
            stride = std::min<uint32_t>(stride, std::max<uint32_t>(1, (uint32_t)kQrSlots * sh.gx / B));
            stride = std::min<uint32_t>(64, std::max<uint32_t>(1, stride));
            stride = std::max(stride, (ntiles + 131071) / 131072);
        }
        sh.stride = std::min(stride, ntiles);
        return sh;
    }
    /* 8 waves per WG (one WG per CU) sharing one copy of the query operands
     * and one tile counter measured best for the single-tile BF16 scan;
     * 4 waves for the two-tile variant */
    const bool forced = qtw == 1 && o && (o->waves == 4 || o->waves == 8) &&
                        !(dim > 256 && o->waves == 4);
    /* alone on the chip: 8-wave WGs, one per CU (best streaming rate).
     * Sharing it with other batches' kernels (device-resident batches on
     * several streams): 4-wave WGs, one per CU, so scans and small kernels
     * of other batches co-reside (8-wave WGs at ~180 VGPRs fill a CU's
     * register file; measured 1.24M vs 1.09M queries/s at 1M x 128) */
    const bool share = shared && qtw == 1 && dim <= 256;
    sh.nw = forced ? o->waves : ((qtw == 2 || share) ? 4 : 8);
    /* main-scan workgroups per query tile (also bounds the candidates each
     * WG collects per query: at least B / gx of them) */
    auto grid_x = [&]() {
        sh.wpc = (o && o->wg_per_cu > 0) ? std::min(o->wg_per_cu, 16)
                                         : ((bf && sh.nw == 8) || share ? 1 : 2);
        uint32_t x = (uint32_t)std::max(64, (sh.wpc * g->ncu) / std::max(1, sh.gy));
        return std::min<uint32_t>(x, std::max<uint32_t>(1, ntiles / sh.nw));
    };
    uint32_t gx = grid_x();
    /* a large B over few rows: 4-wave WGs (twice the grid) keep each WG's
     * share of a query's candidates within its LDS slots */
    if (!forced && qtw == 1 && dim <= 256 && sh.nw == 8 && (uint64_t)2 * B > (uint64_t)gx * kLdsCand) {
        sh.nw = 4;
        gx = grid_x();
    }
    sh.gx = gx;
    uint32_t stride = stride_override ? stride_override
                                      : (o && o->sample_stride ? o->sample_stride : 0);
    if (stride == 0) {
        /* the B-th smallest minimum over groups of 16 sampled rows: about
         * stride x B rows of a query fall under it; keep that near half the
         * WG's per-query LDS slots per WG */
        stride = std::min<uint32_t>(64, std::max<uint32_t>(1, (uint32_t)kLdsCand / 2 * gx / B));
        const uint32_t min_stride = (ntiles + 131071) / 131072;   /* <= 2^18 groups */
        stride = std::max(stride, min_stride);
    }
    sh.stride = std::min(stride, ntiles);
    return sh;
}

/* INT8 scan launch: waves per WG (4 / 8), tiles per unit (code dims 64:
 * 4, 128: 2, 256: 1; wider rows: the multi-unit kernel) and inline
 * predicates pick the kernel instance */
template <bool PREPASS, bool F8>
static void launch_scan_i8_t(int nw, int pt, bool inl, uint32_t gx, int nqt, size_t shm,
                             hipStream_t s, const ScanArgs &A0) {
    ScanArgs A = A0;
    const dim3 grid = scan_grid(A, gx, (uint32_t)nqt);
#define PV_I8(NW, PT, INL) \
    hipLaunchKernelGGL((k_scan_i8<PREPASS, NW, PT, INL, F8>), grid, dim3(NW * 64), shm, s, A)
#define PV_I8W(NW, INL) \
    hipLaunchKernelGGL((k_scan_i8w<PREPASS, NW, INL, F8>), grid, dim3(NW * 64), shm, s, A)
    if (A.dimp > 256) {
        if (nw == 8) { if (inl) PV_I8W(8, true); else PV_I8W(8, false); }
        else { if (inl) PV_I8W(4, true); else PV_I8W(4, false); }
    } else if (nw == 8) {
        if (pt == 4) { if (inl) PV_I8(8, 4, true); else PV_I8(8, 4, false); }
        else if (pt == 2) { if (inl) PV_I8(8, 2, true); else PV_I8(8, 2, false); }
        else { if (inl) PV_I8(8, 1, true); else PV_I8(8, 1, false); }
    } else {
        if (pt == 4) { if (inl) PV_I8(4, 4, true); else PV_I8(4, 4, false); }
        else if (pt == 2) { if (inl) PV_I8(4, 2, true); else PV_I8(4, 2, false); }
        else { if (inl) PV_I8(4, 1, true); else PV_I8(4, 1, false); }
    }
#undef PV_I8
#undef PV_I8W
}
/* INT8 or FP8 (code format of the index's INT8-layout copy) */
template <bool PREPASS>
static void launch_scan_i8(int nw, int pt, bool inl, bool f8, uint32_t gx, int nqt, size_t shm,
                           hipStream_t s, const ScanArgs &A0) {
    if (f8) launch_scan_i8_t<PREPASS, true>(nw, pt, inl, gx, nqt, shm, s, A0);
    else launch_scan_i8_t<PREPASS, false>(nw, pt, inl, gx, nqt, shm, s, A0);
}

static hipError_t launch_main_scan(const ScanLaunch &L, hipStream_t s) {
    if (L.f32m) {
        hipLaunchKernelGGL((k_scan_f32m<false, 8>), dim3(L.gx, L.nqt), dim3(512), L.shm, s, L.M);
    } else if (L.i8) {
        launch_scan_i8<false>(L.nw, L.pt, L.inl, L.f8, L.gx, L.nqt, L.shm, s, L.M);
    } else if (L.bf && L.qtw == kQrNW) {
        const hipError_t e = launch_scan_qr<false>(L.gx, L.gy, L.inl, s, L.M);
        if (e != hipSuccess) return e;
    } else if (L.bf && L.qtw > 1) {
        const hipError_t e = launch_scan_mq<false>(L.qtw, L.inl, L.gx, L.gy, s, L.M);
        if (e != hipSuccess) return e;
    } else if (L.bf) {
        launch_scan_bf16<false>(L.nw, L.inl, L.gx, L.nqt, L.shm, s, L.M);
    } else {
        hipLaunchKernelGGL(k_scan_f32<false>, dim3(L.gx, L.nqt), dim3(256), 0, s, L.M);
    }
    return hipGetLastError();
}

/* the options a NULL pointer stands for: index precision, automatic route
 * and shape (a zero-filled struct would ask for the FP32 scan) */
static inline FvGpuBruteOpts default_opts() {
    FvGpuBruteOpts o = {};
    o.precision = -1;
    return o;
}

/* Enqueue the brute pipeline. stride_override > 0 forces the pre-pass
 * sample stride (1 = every row: the tightest threshold). */
int run_brute_ex(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *o,
                 uint32_t stride_override, const uint8_t *skip, bool have_phase1,
                 char *errbuf, size_t errlen, const LearnArgs *decide) {
    hipStream_t s = b->stream;
    g->ev_stream = s;
    b->out_stride = b->topb; /* a graph run may have widened it; d_out holds nq x topb */
    const int prec = scan_precision(g, b, o);
    const int nq = b->nq, nqp = b->nqt * kQT, dim = g->dim;
    const uint32_t B = b->topb;
    /* quantized INT8 scan: the scan keeps the top over x k by code distance */
    const uint32_t over = scan_overfetch(g, prec, o);
    const uint32_t Bc = scan_candidates(g, b, prec, o);
    QAux qa = qaux_of(b);
    FvGpuStats &st = g->last;
    memset(&st, 0, sizeof(st));
    st.precision = (prec == FV_GPU_PREC_SQ8 && g->i8_mode == 3) ? FV_GPU_PREC_FP8 : prec;
    st.nprog = (uint32_t)b->nprog;
    g->last_batch = b;

    int route = o ? o->route : 0;
    if (route == 0) {
        /* gather when every query enumerates a short span list whose
         * random-row reads cost less than one shared scan pass */
        const double row_bytes =
            (double)dim * (prec == FV_GPU_PREC_BF16 ? 2.0 : (prec == FV_GPU_PREC_SQ8 ? 1.0 : 4.0));
        const double scan_bytes = (double)g->count * row_bytes;
        const double gather_bytes = (double)b->sum_total * dim * 4.0 * 2.0;
        route = (b->all_spans && b->max_total <= 65536 && gather_bytes < scan_bytes) ? 2 : 1;
    }
    b->last_route = route;
    st.route = route;

    /* inline predicates: one comparison per query evaluated in the scan
     * epilogue, no bitmap pass (BF16 scan, all predicates on one column) */
    const bool inl = (prec == FV_GPU_PREC_BF16 || prec == FV_GPU_PREC_SQ8) && o &&
                     o->inline_pred && b->inl_ok &&
                     b->inl_col >= 0 && !have_phase1;
    if (route == 1 && !inl && require_bitmaps(b, errbuf, errlen) != 0) return -1;
    b->ran_bitmap = route == 1 && !inl;

    ev_record(g, 0);
    if (!have_phase1) PV_TRY(launch_phase1(g, b, prec, route == 1 && !inl, s, nullptr));
    ev_record(g, 6);
    ev_record(g, 1);

    if (route == 2) {
        if (b->any_dedup) PV_FAIL("gpu: gather route cannot deduplicate spans");
        if (g->d_prows == nullptr)
            for (int q = 0; q < nq; q++)
                if (b->nspans[q] > 0) PV_FAIL("gpu: gather route needs posting rows");
        if (b->max_total > (uint64_t)(1u << 22))
            PV_FAIL("gpu: gather route: enumeration of %llu rows too large",
                    (unsigned long long)b->max_total);
        const uint32_t chunks = (uint32_t)std::min<uint64_t>(
            64, std::max<uint64_t>(1, (b->max_total + 4095) / 4096));
        /* gather chunk y appends to pool segment y % kPoolShards: a segment
         * holds at most ceil(chunks / shards) chunks of ceil(total / chunks) */
        const uint64_t per_chunk = (b->max_total + chunks - 1) / chunks;
        const uint64_t seg = (uint64_t)((chunks + kPoolShards - 1) / kPoolShards) * per_chunk;
        b->cap_pool = kPoolShards * std::max<uint64_t>(128, seg);
        PV_TRY(b->d_pool.ensure((size_t)nq * b->cap_pool * 8));
        hipLaunchKernelGGL(k_reset, dim3((nq + 255) / 256), dim3(256), 0, s, nq,
                           qa.tau, b->d_pcnt.as<uint32_t>(), qa.status);
        hipLaunchKernelGGL(k_gather, dim3(nq, chunks), dim3(256), 0, s, g->d_vecs,
                           g->d_cols, g->d_prows, g->count, dim, b->d_q, b->d_qd,
                           b->d_spans, b->d_spanoff, b->d_ops, b->d_poff, b->d_plen,
                           b->d_pofq, b->d_vals, b->d_pool.as<uint64_t>(),
                           b->d_pcnt.as<uint32_t>(), b->cap_pool);
        PV_TRY(hipGetLastError());
        if (decide) {
            /* no threshold launch on this route: the decision blocks alone */
            hipLaunchKernelGGL(k_tau_route, dim3((nq + 3) / 4), dim3(256), 0, s, TauArgs{},
                               *decide, 0);
            PV_TRY(hipGetLastError());
        }
        ev_record(g, 4);
        /* gathered distances are already exact: eps = 0 */
        hipLaunchKernelGGL(k_select, dim3(nq), dim3(256), dim * 4, s,
                           b->d_pool.as<uint64_t>(), b->d_pcnt.as<uint32_t>(),
                           b->cap_pool, B, 1u, (const float *)nullptr,
                           g->d_vecs, b->d_q, dim, b->d_out.as<FvHit>(),
                           b->out_stride, b->d_cnt.as<uint32_t>(),
                           b->d_flag.as<uint8_t>(), qa.status, skip,
                           b->d_topbq_p);
        PV_TRY(hipGetLastError());
        ev_record(g, 5);
        st.scan_rows = b->sum_total;
        st.scan_bytes = b->sum_total * (uint64_t)dim * 4;
        return 0;
    }

    /* ---- full-scan route ---- */
    /* pre-pass samples every stride-th tile (INT8: 8 KB unit of PT tiles) */
    const bool i8 = prec == FV_GPU_PREC_SQ8;
    const bool f8 = i8 && g->i8_mode == 3;
    const int pt = (int)scan_unit_tiles(g, prec);
    const uint32_t ntiles = (uint32_t)(g->count_pad / 32) / (uint32_t)pt;
    const bool bf = prec == FV_GPU_PREC_BF16;
    const bool f32m = prec == FV_GPU_PREC_FP32;   /* FP32 MFMA scan (8-wave WGs) */
    const bool auto_shape = !(o && (o->waves || o->wg_per_cu));
    const ScanShape sh = scan_shape(g, b, o, prec, stride_override,
                                    (bf || i8) && auto_shape && other_batches_in_flight(g, b));
    const int qtw = sh.qtw, gy = sh.gy, nw = f32m ? 8 : sh.nw;
    const uint32_t gx = sh.gx, stride = sh.stride;
    const uint32_t stiles = (ntiles + stride - 1) / stride;
    b->stride = stride;
    b->ngroups = stiles * 2 * (uint32_t)pt;
    b->cap_pool = std::max<uint64_t>(16384, (uint64_t)8 * Bc * stride);
    if (stride_override == 1) b->cap_pool = std::max<uint64_t>(b->cap_pool, (uint64_t)64 * Bc + 4096);
    PV_TRY(b->d_gmin.ensure((size_t)nqp * b->ngroups * 4));
    PV_TRY(b->d_pool.ensure((size_t)nq * b->cap_pool * 8));

    ScanArgs A = {};
    A.vbf = g->d_vbf;
    A.vecs = g->d_vecs;
    A.xnorm = g->d_norm;
    A.count = g->count;
    A.count_pad = g->count_pad;
    A.dim = dim;
    A.dimb = qbf_dim(g);
    A.qbf = b->d_qbf.as<uint16_t>();
    A.qf = b->d_q;
    A.qnorm = qa.qnorm;
    A.tau = qa.tau;
    A.pofq = b->d_pofq;
    A.bitmap = b->d_bitmap.as<uint32_t>();
    A.words = g->count_pad / 32;
    A.nq = nq;
    A.pool = b->d_pool.as<uint64_t>();
    A.pcnt = b->d_pcnt.as<uint32_t>();
    A.cap_pool = b->cap_pool;
    A.gmin = b->d_gmin.as<uint32_t>();
    A.ngroups = b->ngroups;
    A.ones = g->d_ones;
    A.status = qa.status;
    A.skip = skip;
    A.inl = inl ? b->d_inl_p : nullptr;
    A.icol = inl ? g->d_cols + (uint64_t)b->inl_col * g->count : nullptr;
    A.irank = (inl && g->d_rank != nullptr) ? g->d_rank + (uint64_t)b->inl_col * g->count : nullptr;
    A.sweep = (o && o->order == 2) ? 1u : 0u;
    /* query-major grid: asked for, or by default for rows above 256 dims
     * (several query tiles re-read each chunk: running them together lets
     * the re-reads hit in cache) */
    const int order = o ? o->order : 0;
    A.qmajor = ((bf || i8) && (order == 3 || (order == 0 && dim > 256 && b->nqt >= 2))) ? 1u : 0u;
    A.xn2max = g->xmax_norm * g->xmax_norm;
    if (i8) {
        A.vi8 = g->d_vi8;
        A.qi8 = b->d_qi8.as<int8_t>();
        A.xnorm = g->d_norm8;
        A.qnorm = b->d_qn8.as<float>();
        /* code norms stay below dimp x (largest code)^2 */
        A.xn2max = (float)g->dimp * (f8 ? 448.f * 448.f : 16384.f);
    }
    A.dimp = g->dimp;
    /* INT8: the queries' code operands, dimp x 64 bytes */
    const size_t shm_i8 = (size_t)2 * (g->dimp / 32) * 64 * 16 + sizeof(CandLds);
    if (i8 && shm_i8 > 64 * 1024)
        for (const void *f : {(const void *)k_scan_i8w<true, 8, false>,
                              (const void *)k_scan_i8w<false, 8, false>,
                              (const void *)k_scan_i8w<true, 8, true>,
                              (const void *)k_scan_i8w<false, 8, true>,
                              (const void *)k_scan_i8w<true, 4, false>,
                              (const void *)k_scan_i8w<false, 4, false>,
                              (const void *)k_scan_i8w<true, 4, true>,
                              (const void *)k_scan_i8w<false, 4, true>,
                              (const void *)k_scan_i8w<true, 8, false, true>,
                              (const void *)k_scan_i8w<false, 8, false, true>,
                              (const void *)k_scan_i8w<true, 8, true, true>,
                              (const void *)k_scan_i8w<false, 8, true, true>,
                              (const void *)k_scan_i8w<true, 4, false, true>,
                              (const void *)k_scan_i8w<false, 4, false, true>,
                              (const void *)k_scan_i8w<true, 4, true, true>,
                              (const void *)k_scan_i8w<false, 4, true, true>})
            PV_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm_i8));

    /* 4 waves per WG up to 256 dims (2 WGs per CU), 8 waves above (the
     * query operands then take most of the LDS) */
    const size_t shm_bf = (size_t)2 * (A.dimb / 16) * 64 * 16 + sizeof(CandLds);
    if (bf && shm_bf > 64 * 1024) {
        for (const void *f : {(const void *)k_scan_bf16<true, 8, false>,
                              (const void *)k_scan_bf16<false, 8, false>,
                              (const void *)k_scan_bf16<true, 8, true>,
                              (const void *)k_scan_bf16<false, 8, true>})
            PV_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm_bf));
    }
    if (bf && shm_bf > 160 * 1024) PV_FAIL("gpu: dim %d too large for the BF16 scan", dim);
    /* FP32 MFMA scan: the 64 queries' FP32 operands (dim x 256 bytes) in LDS */
    const size_t shm_f32m = (size_t)2 * (dim / 8) * 64 * 16 + sizeof(CandLds);
    if (f32m && shm_f32m > 160 * 1024) PV_FAIL("gpu: dim %d too large for the FP32 MFMA scan", dim);
    if (f32m && shm_f32m > 64 * 1024)
        for (const void *f : {(const void *)k_scan_f32m<true, 8>, (const void *)k_scan_f32m<false, 8>})
            PV_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm_f32m));
    /* pre-pass over sampled tiles (with a decision made in this run's
     * threshold launch, the route is not known yet: no skipping) */
    {
        ScanArgs P = A;
        if (decide) P.skip = nullptr;
        P.ntiles = stiles;
        P.stride = stride;
        if (f32m) {
            const uint32_t gxp = std::min<uint32_t>(gx, std::max<uint32_t>(1, (stiles + nw - 1) / nw));
            hipLaunchKernelGGL((k_scan_f32m<true, 8>), dim3(gxp, b->nqt), dim3(512), shm_f32m, s, P);
        } else if (bf || i8) {
            /* each pre-pass WG walks several sampled tiles (at most the
             * main scan's grid): staging 64 query operands per WG costs as
             * much as a few tiles */
            const uint32_t gxp = std::min<uint32_t>(gx, std::max<uint32_t>(1, (stiles + nw - 1) / nw));
            if (i8) launch_scan_i8<true>(nw, pt, inl, f8, gxp, b->nqt, shm_i8, s, P);
            else if (qtw == kQrNW) {
                /* chunks of at least a few ring stages of sampled tiles */
                const uint32_t gq = std::min<uint32_t>(
                    gx, std::max<uint32_t>(1, stiles / (4 * kQrTps)));
                PV_TRY(launch_scan_qr<true>(gq, gy, inl, s, P));
            }
            else if (qtw > 1) PV_TRY(launch_scan_mq<true>(qtw, inl, gxp, gy, s, P));
            else launch_scan_bf16<true>(nw, inl, gxp, b->nqt, shm_bf, s, P);
        } else {
            const uint32_t gx = std::max<uint32_t>(1, (stiles + 3) / 4);
            hipLaunchKernelGGL(k_scan_f32<true>, dim3(gx, b->nqt), dim3(256), 0, s, P);
        }
        PV_TRY(hipGetLastError());
    }
    ev_record(g, 2);
    {
        TauArgs T;
        T.gmin = b->d_gmin.as<uint32_t>();
        T.ngroups = b->ngroups;
        T.B = B;
        T.over = over;
        T.eps = qa.eps;
        T.tau = qa.tau;
        T.pcnt = b->d_pcnt.as<uint32_t>();
        T.status = qa.status;
        T.skip = decide ? nullptr : skip;
        T.topbq = b->d_topbq_p;
        if (decide)
            hipLaunchKernelGGL(k_tau_route, dim3(nq + (nq + 3) / 4), dim3(256), 0, s, T, *decide,
                               nq);
        else
            hipLaunchKernelGGL(k_tau, dim3(nq), dim3(256), 0, s, T);
    }
    PV_TRY(hipGetLastError());
    ev_record(g, 3);
    {
        ScanLaunch L;
        L.M = A;
        L.M.ntiles = ntiles;
        L.M.stride = 1;
        L.bf = bf;
        L.f32m = f32m;
        L.inl = inl;
        L.i8 = i8;
        L.f8 = f8;
        L.pt = pt;
        L.qtw = qtw;
        L.nw = nw;
        L.nqt = b->nqt;
        L.gy = gy;
        L.gx = gx;
        L.shm = i8 ? shm_i8 : (f32m ? shm_f32m : shm_bf);
        PV_TRY(launch_main_scan(L, s));
        if (!b->arena_bound) {
            if (b->last_scan == nullptr) b->last_scan = new ScanLaunch();
            *(ScanLaunch *)b->last_scan = L;
        }
    }
    ev_record(g, 4);
    hipLaunchKernelGGL(k_select, dim3(nq), dim3(256), dim * 4, s,
                       b->d_pool.as<uint64_t>(), b->d_pcnt.as<uint32_t>(),
                       b->cap_pool, B, over, qa.eps, g->d_vecs, b->d_q, dim,
                       b->d_out.as<FvHit>(), b->out_stride,
                       b->d_cnt.as<uint32_t>(), b->d_flag.as<uint8_t>(), qa.status, skip,
                           b->d_topbq_p);
    PV_TRY(hipGetLastError());
    ev_record(g, 5);
    st.scan_rows = g->count;
    st.scan_bytes = g->count * (bf ? (uint64_t)A.dimb * 2 : (uint64_t)dim * (i8 ? 1 : 4)) *
                    (uint64_t)b->nqt;
    return 0;
}

static int run_brute(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *o,
                     uint32_t stride_override, char *errbuf, size_t errlen) {
    return run_brute_ex(g, b, o, stride_override, nullptr, false, errbuf, errlen);
}

/* copy outputs and per-query status to pinned host memory (synchronous) */
static int fetch_raw(FvGpuIndex *g, FvGpuBatch *b, char *errbuf, size_t errlen) {
    const int nq = b->nq, nqp = b->nqt * kQT;
    const size_t o_out = 0;
    const size_t o_cnt = al256((size_t)nq * b->out_stride * sizeof(FvHit));
    const size_t o_flag = al256(o_cnt + (size_t)nq * 4);
    const size_t o_st = al256(o_flag + (size_t)nq);
    const size_t tot = o_st + (size_t)nqp * 4;
    PV_TRY(b->h_out.ensure(tot));
    unsigned char *h = b->h_out.as<unsigned char>();
    QAux qa = qaux_of(b);
    PV_TRY(hipMemcpyAsync(h + o_out, b->d_out.p, (size_t)nq * b->out_stride * sizeof(FvHit),
                          hipMemcpyDeviceToHost, b->stream));
    PV_TRY(hipMemcpyAsync(h + o_cnt, b->d_cnt.p, (size_t)nq * 4, hipMemcpyDeviceToHost, b->stream));
    PV_TRY(hipMemcpyAsync(h + o_flag, b->d_flag.p, (size_t)nq, hipMemcpyDeviceToHost, b->stream));
    PV_TRY(hipMemcpyAsync(h + o_st, qa.status, (size_t)nq * 4, hipMemcpyDeviceToHost, b->stream));
    PV_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

static void fetch_views(FvGpuBatch *b, const FvHit **out, const uint32_t **cnt,
                        const uint8_t **flag, const uint32_t **status) {
    const int nq = b->nq;
    const size_t o_cnt = al256((size_t)nq * b->out_stride * sizeof(FvHit));
    const size_t o_flag = al256(o_cnt + (size_t)nq * 4);
    const size_t o_st = al256(o_flag + (size_t)nq);
    unsigned char *h = b->h_out.as<unsigned char>();
    *out = (const FvHit *)h;
    *cnt = (const uint32_t *)(h + o_cnt);
    *flag = (const uint8_t *)(h + o_flag);
    *status = (const uint32_t *)(h + o_st);
}

static void fill_stats_times(FvGpuIndex *g) {
    if (!g->timing) return;
    FvGpuStats &st = g->last;
    st.t_bitmap_us = ev_elapsed_us(g, 0, 6);
    st.t_prepass_us = ev_elapsed_us(g, 1, 2);
    st.t_scan_us = st.route == 2 ? ev_elapsed_us(g, 1, 4) : ev_elapsed_us(g, 3, 4);
    if (st.route == 2) st.t_gather_us = st.t_scan_us;
    st.t_merge_us = ev_elapsed_us(g, 2, 3);
    st.t_rerank_us = ev_elapsed_us(g, 4, 5);
    st.t_total_us = ev_elapsed_us(g, 0, 5);
    /* candidate-pool sizes of the last batch (timing mode only: a copy) */
    FvGpuBatch *b = g->last_batch;
    if (b && b->nq > 0 && b->d_pcnt.p) {
        std::vector<uint32_t> pc((size_t)b->nq * kPoolShards * kPcntStride);
        if (hipMemcpy(pc.data(), b->d_pcnt.p, pc.size() * 4, hipMemcpyDeviceToHost) ==
            hipSuccess) {
            uint64_t s = 0;
            for (int q = 0; q < b->nq; q++)
                for (uint32_t h = 0; h < kPoolShards; h++)
                    s += pc[((size_t)q * kPoolShards + h) * kPcntStride];
            st.candidates = s;
        }
    }
}

} // namespace pv

using namespace pv;

namespace pv {

/* popcount of each program's bitmap row (pv_learn.hip) */
__global__ void k_prog_count(const uint32_t *__restrict__ bitmap, uint64_t words,
                             uint32_t *__restrict__ cnt);

void feedback_begin(FvGpuIndex *g, hipStream_t s) {
    if (g->fb_ev[0] == nullptr &&
        (hipEventCreate(&g->fb_ev[0]) != hipSuccess || hipEventCreate(&g->fb_ev[1]) != hipSuccess)) {
        (void)hipGetLastError();
        return;
    }
    (void)hipEventRecord(g->fb_ev[0], s);
}

void feedback_end(FvGpuIndex *g, FvGpuBatch *b, hipStream_t s, bool bitmaps) {
    if (g->fb_ev[1] == nullptr) return;
    (void)hipEventRecord(g->fb_ev[1], s);
    if (bitmaps && b->nprog > 0 &&
        b->d_order.ensure(((size_t)b->nq + (size_t)b->nprog) * 4) == hipSuccess) {
        hipLaunchKernelGGL(k_prog_count, dim3((unsigned)b->nprog), dim3(256), 0, s,
                           b->d_bitmap.as<uint32_t>(), g->count_pad / 32,
                           b->d_order.as<uint32_t>() + b->nq);
        (void)hipGetLastError();
    }
}

void feedback_record(FvGpuIndex *g, FvGpuBatch *b, int strategy, const FvGpuQueryDesc *qd,
                     uint32_t want, bool bitmaps) {
    if (g->fb_ev[0] == nullptr || g->count == 0) return;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g->fb_ev[0], g->fb_ev[1]) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    const int nq = b->nq;
    std::vector<uint32_t> pc;
    if (bitmaps && b->nprog > 0) {
        pc.resize((size_t)b->nprog);
        if (hipMemcpy(pc.data(), b->d_order.as<uint32_t>() + nq, pc.size() * 4,
                      hipMemcpyDeviceToHost) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
    }
    /* program of each query, from the host copy of the staging area */
    const int32_t *pofq = (const int32_t *)(b->h_stage.as<unsigned char>() + b->lay.o_pofq);
    std::vector<int32_t> sel((size_t)nq);
    std::vector<uint32_t> wq((size_t)nq, want);
    const double n = (double)g->count;
    for (int q = 0; q < nq; q++) {
        const int p = pofq[q];
        double pass = n;
        if (p >= 0)
            pass = !pc.empty() ? (double)pc[(size_t)p]
                               : (double)std::min<uint64_t>(qd[q].total, g->count);
        sel[(size_t)q] = (int32_t)std::lround(std::min(pass, n) * 1000.0 / n);
    }
    std::lock_guard<std::mutex> lk(g->fb_mu);
    fv_gpu_feedback_update(&g->fb, strategy, nq, sel.data(), wq.data(), (double)ms * 1000.0);
}

} // namespace pv

static int brute_sync(FvGpuIndex *g, int nq, const float *queries,
                      const FvGpuQueryDesc *qd, const uint8_t *dedup,
                      const FvGpuPredNode *nodes, int nnodes,
                      const int32_t *children, int nchildren,
                      const double *invals, int ninvals, const FvGpuSpan *spans,
                      int nspans_total, const uint64_t *spanoff, int nspanoff,
                      uint32_t topb, const FvGpuBruteOpts *opts, FvHit *out,
                      uint32_t *out_cnt, uint8_t *truncated, char *errbuf,
                      size_t errlen) {
    if (g == nullptr) PV_FAIL("gpu: NULL index");
    /* argument checks in the order of gpu_engine.cu:2217-2222 */
    if (nq < 1 || nq > FV_GPU_MAX_BATCH || topb == 0 || topb > FV_GPU_MAX_TOPB ||
        nspans_total > FV_GPU_MAX_SPANS * nq)
        PV_FAIL("gpu: bad brute batch (nq=%d topb=%u)", nq, topb);
    PV_TRY(hipSetDevice(g->device));
    if (out == nullptr || out_cnt == nullptr || truncated == nullptr)
        PV_FAIL("gpu: NULL output arrays");
    if (g->sync_batch == nullptr) g->sync_batch = new FvGpuBatch();
    FvGpuBatch *b = g->sync_batch;
    if (stage_batch(g, b, nq, queries, qd, dedup, nodes, nnodes, children,
                    nchildren, invals, ninvals, spans, nspans_total, spanoff,
                    nspanoff, topb, errbuf, errlen) != 0)
        return -1;
    feedback_begin(g, b->stream);
    if (run_brute(g, b, opts, 0, errbuf, errlen) != 0) return -1;
    feedback_end(g, b, b->stream, b->ran_bitmap);
    if (fetch_raw(g, b, errbuf, errlen) != 0) return -1;
    fill_stats_times(g);
    const FvHit *ho;
    const uint32_t *hc, *hs;
    const uint8_t *hf;
    fetch_views(b, &ho, &hc, &hf, &hs);
    bool bad = false;
    for (int q = 0; q < nq; q++) bad |= hs[q] != 0;
    if (bad) {
        /* pool overflow or oversized re-rank set: retry with the tightest
         * threshold (pre-pass over every row) */
        FvGpuBruteOpts o2 = opts ? *opts : default_opts();
        if (o2.route == 0) o2.route = 1;
        g->last.certified_fail = 0;
        for (int q = 0; q < nq; q++) g->last.certified_fail += hs[q] != 0;
        const uint32_t nfail = g->last.certified_fail;
        if (run_brute(g, b, &o2, 1, errbuf, errlen) != 0) return -1;
        feedback_end(g, b, b->stream, b->ran_bitmap);
        if (fetch_raw(g, b, errbuf, errlen) != 0) return -1;
        fill_stats_times(g);
        g->last.certified_fail = nfail;
        fetch_views(b, &ho, &hc, &hf, &hs);
        for (int q = 0; q < nq; q++)
            if (hs[q] != 0)
                PV_FAIL("gpu: query %d: candidate set exceeds capacity (status %u)",
                        q, hs[q]);
    }
    for (int q = 0; q < nq; q++) {
        const uint32_t c = hc[q];
        memcpy(out + (uint64_t)q * topb, ho + (uint64_t)q * b->out_stride,
               (size_t)c * sizeof(FvHit));
        out_cnt[q] = c;
        truncated[q] = hf[q];
    }
    feedback_record(g, b, 0, qd, topb, b->ran_bitmap);
    return 0;
}

extern "C" int fv_gpu_brute_batch(FvGpuIndex *g, int nq, const float *queries,
                                  const FvGpuQueryDesc *qd, const uint8_t *dedup,
                                  const FvGpuPredNode *nodes, int nnodes,
                                  const int32_t *children, int nchildren,
                                  const double *invals, int ninvals,
                                  const FvGpuSpan *spans, int nspans_total,
                                  const uint64_t *spanoff, int nspanoff,
                                  uint32_t topb, FvHit *out, uint32_t *out_cnt,
                                  uint8_t *truncated, char *errbuf,
                                  size_t errlen) {
    try {
        return brute_sync(g, nq, queries, qd, dedup, nodes, nnodes, children,
                          nchildren, invals, ninvals, spans, nspans_total,
                          spanoff, nspanoff, topb, nullptr, out, out_cnt,
                          truncated, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

extern "C" int fv_gpu_brute_batch_ex(FvGpuIndex *g, int nq, const float *queries,
                                     const FvGpuQueryDesc *qd, const uint8_t *dedup,
                                     const FvGpuPredNode *nodes, int nnodes,
                                     const int32_t *children, int nchildren,
                                     const double *invals, int ninvals,
                                     const FvGpuSpan *spans, int nspans_total,
                                     const uint64_t *spanoff, int nspanoff,
                                     uint32_t topb, const FvGpuBruteOpts *opts,
                                     FvHit *out, uint32_t *out_cnt,
                                     uint8_t *truncated, char *errbuf,
                                     size_t errlen) {
    try {
        return brute_sync(g, nq, queries, qd, dedup, nodes, nnodes, children,
                          nchildren, invals, ninvals, spans, nspans_total,
                          spanoff, nspanoff, topb, opts, out, out_cnt,
                          truncated, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

extern "C" FvGpuBatch *fv_gpu_batch_stage(
    FvGpuIndex *g, int nq, const float *queries, const FvGpuQueryDesc *qd,
    const uint8_t *dedup, const FvGpuPredNode *nodes, int nnodes,
    const int32_t *children, int nchildren, const double *invals, int ninvals,
    const FvGpuSpan *spans, int nspans_total, const uint64_t *spanoff,
    int nspanoff, uint32_t topb, char *errbuf, size_t errlen) {
    if (g == nullptr) {
        snprintf(errbuf, errlen, "gpu: NULL index");
        return nullptr;
    }
    try {
        if (hipSetDevice(g->device) != hipSuccess) {
            snprintf(errbuf, errlen, "gpu: hipSetDevice failed");
            return nullptr;
        }
        FvGpuBatch *b = new FvGpuBatch();
        /* device-resident batches take up to kMaxBatchEx queries (the
         * synchronous entry points keep the reference's FV_GPU_MAX_BATCH) */
        if (stage_batch(g, b, nq, queries, qd, dedup, nodes, nnodes, children,
                        nchildren, invals, ninvals, spans, nspans_total, spanoff,
                        nspanoff, topb, errbuf, errlen, nullptr, kMaxBatchEx) != 0 ||
            hipStreamSynchronize(b->stream) != hipSuccess) {
            free_batch_buffers(b);
            delete b;
            return nullptr;
        }
        return b;
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return nullptr;
    }
}

extern "C" FvGpuBatch *fv_gpu_batch_stage_ex(
    FvGpuIndex *g, int nq, const float *queries, const FvGpuQueryDesc *qd,
    const uint8_t *dedup, const FvGpuPredNode *nodes, int nnodes,
    const int32_t *children, int nchildren, const double *invals, int ninvals,
    const FvGpuSpan *spans, int nspans_total, const uint64_t *spanoff,
    int nspanoff, const uint32_t *topb_q, uint32_t topb, char *errbuf, size_t errlen) {
    if (g == nullptr) {
        snprintf(errbuf, errlen, "gpu: NULL index");
        return nullptr;
    }
    try {
        if (hipSetDevice(g->device) != hipSuccess) {
            snprintf(errbuf, errlen, "gpu: hipSetDevice failed");
            return nullptr;
        }
        FvGpuBatch *b = new FvGpuBatch();
        if (stage_batch(g, b, nq, queries, qd, dedup, nodes, nnodes, children,
                        nchildren, invals, ninvals, spans, nspans_total, spanoff,
                        nspanoff, topb, errbuf, errlen, topb_q, kMaxBatchEx) != 0 ||
            hipStreamSynchronize(b->stream) != hipSuccess) {
            free_batch_buffers(b);
            delete b;
            return nullptr;
        }
        return b;
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return nullptr;
    }
}

/* What a captured run bakes into its kernel arguments: every device buffer
 * of the batch and the sizes that shape the launches. A graph search or a
 * routed run on the same batch may grow (reallocate) the output buffer or
 * another one; a replay after that would write through a freed pointer,
 * so a changed signature discards the capture. The program counts are not
 * part of it (the kernels read them from the staging area), nor are the
 * spans unless the captured run took the gather route, whose grid and pool
 * follow the longest enumeration. */
static std::vector<uint64_t> capture_signature(const FvGpuBatch *b) {
    const bool gather = b->last_route == 2;
    return {(uint64_t)(uintptr_t)b->d_stage.p, (uint64_t)(uintptr_t)b->d_qbf.p,
            (uint64_t)(uintptr_t)b->d_qaux.p, (uint64_t)(uintptr_t)b->d_bitmap.p,
            (uint64_t)(uintptr_t)b->d_gmin.p, (uint64_t)(uintptr_t)b->d_pool.p,
            (uint64_t)(uintptr_t)b->d_pcnt.p, (uint64_t)(uintptr_t)b->d_out.p,
            (uint64_t)(uintptr_t)b->d_cnt.p, (uint64_t)(uintptr_t)b->d_flag.p,
            (uint64_t)(uintptr_t)b->d_qi8.p, (uint64_t)(uintptr_t)b->d_qn8.p,
            (uint64_t)(uintptr_t)b->d_progcnt.p, (uint64_t)(uintptr_t)b->d_hdr,
            (uint64_t)(uintptr_t)b->d_ops, (uint64_t)(uintptr_t)b->d_vals,
            (uint64_t)(uintptr_t)b->d_poff, (uint64_t)(uintptr_t)b->d_plen,
            (uint64_t)(uintptr_t)b->d_pofq, (uint64_t)(uintptr_t)b->d_spans,
            (uint64_t)(uintptr_t)b->d_spanoff, (uint64_t)(uintptr_t)b->d_q,
            (uint64_t)(uintptr_t)b->d_qd,
            (uint64_t)b->nq, (uint64_t)b->topb, (uint64_t)(b->nprog > 0),
            b->cap_pool, (uint64_t)b->ngroups, (uint64_t)b->stride,
            (uint64_t)b->inl_ok, (uint64_t)(int64_t)b->inl_col, (uint64_t)b->inl_rank,
            (uint64_t)b->i8_ok,
            (uint64_t)(uintptr_t)b->d_inl_p, (uint64_t)(uintptr_t)b->d_topbq_p,
            (uint64_t)b->last_route, gather ? b->max_total : 0,
            (uint64_t)(gather && b->all_spans), (uint64_t)(gather && b->any_dedup)};
}

static void drop_capture(FvGpuBatch *b) {
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->gsig.clear();
}

/* the capture still fits the batch as staged now (same buffers and launch
 * sizes; bitmaps, if the captured run builds them, hold every program) */
static bool capture_valid(const FvGpuBatch *b) {
    if (b->gexec == nullptr || b->gsig != capture_signature(b)) return false;
    if (b->graph_bitmap && b->nprog > kBmProgs) return false;
    return true;
}

/* run through a hipGraph captured for this batch and these options: the
 * first call runs eagerly (allocating every buffer), the second captures,
 * later calls replay — also after fv_gpu_batch_restage wrote a new batch
 * into the same buffers. Timing events are not recorded in graph mode. */
static int run_brute_graph(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *o,
                           char *errbuf, size_t errlen) {
    if (b->gexec && !capture_valid(b)) {
        /* a buffer the capture points at was reallocated (or the batch
         * changed shape): run eagerly once more, then capture again */
        drop_capture(b);
        b->warmed = false;
    }
    if (b->gexec && memcmp(&b->gopts, o, sizeof(*o)) != 0) {
        drop_capture(b);
        b->warmed = false;
    }
    if (b->gexec == nullptr) {
        const bool timing = g->timing;
        if (!b->warmed || timing) {
            b->warmed = true;
            return run_brute(g, b, o, 0, errbuf, errlen);
        }
        PV_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        b->capturing = true;
        const int rc = run_brute(g, b, o, 0, errbuf, errlen);
        b->capturing = false;
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &gr);
        if (rc != 0) {
            if (gr) hipGraphDestroy(gr);
            return -1;
        }
        PV_TRY(e);
        b->graph = gr;
        PV_TRY(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->gopts = *o;
        b->gsig = capture_signature(b);
        b->graph_bitmap = b->ran_bitmap;
    }
    /* the replay's select writes topb hits per query (a graph search in
     * between may have left the graph's stride) */
    b->out_stride = b->topb;
    PV_TRY(hipGraphLaunch(b->gexec, b->stream));
    return 0;
}

/* Restage: a new set of queries, predicates and spans written into an
 * existing batch. Waits for the batch's pending work, then stages through
 * its own host and device buffers, keeping the staging layout when the new
 * contents fit it. Returns 1 when the batch's captured run stays valid (a
 * graph run replays it as is), 0 when none is captured or the next graph
 * run captures again, -1 on error (the batch is then unusable until staged
 * again successfully). */
extern "C" int fv_gpu_batch_restage(
    FvGpuIndex *g, FvGpuBatch *b, int nq, const float *queries, const FvGpuQueryDesc *qd,
    const uint8_t *dedup, const FvGpuPredNode *nodes, int nnodes, const int32_t *children,
    int nchildren, const double *invals, int ninvals, const FvGpuSpan *spans,
    int nspans_total, const uint64_t *spanoff, int nspanoff, const uint32_t *topb_q,
    uint32_t topb, char *errbuf, size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g) PV_FAIL("gpu: bad batch");
    try {
        PV_TRY(hipSetDevice(g->device));
        PV_TRY(hipStreamSynchronize(b->stream));
        if (b->aux_stream) PV_TRY(hipStreamSynchronize(b->aux_stream));
        if (stage_batch(g, b, nq, queries, qd, dedup, nodes, nnodes, children, nchildren, invals,
                        ninvals, spans, nspans_total, spanoff, nspanoff, topb, errbuf, errlen,
                        topb_q, kMaxBatchEx, true) != 0) {
            drop_capture(b);
            b->warmed = false;
            return -1;
        }
        PV_TRY(hipStreamSynchronize(b->stream));
        return capture_valid(b) ? 1 : 0;
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

extern "C" int fv_gpu_batch_run_brute(FvGpuIndex *g, FvGpuBatch *b,
                                      const FvGpuBruteOpts *opts, char *errbuf,
                                      size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g) PV_FAIL("gpu: bad batch");
    try {
        if (opts && opts->graph && !g->timing) return run_brute_graph(g, b, opts, errbuf, errlen);
        return run_brute(g, b, opts, 0, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

/* Autotuner: times every candidate launch shape of the full-scan route on
 * this batch (query tiles per WG {1, 2, 4} where the multi-tile scan
 * applies, waves per single-tile WG {4, 8} up to 256 dims, WGs per CU
 * {1, 2, 3, 4}, tile order {chunked, sweep, query-major grid}, pre-pass
 * stride = over-fetch factor {auto/2, auto, 2 auto}), keeps those whose
 * runs certify every query (no
 * pool overflow, no oversized re-rank set) and returns the fastest. All
 * candidates compute the same exact result; only the time differs. */
static int tune_batch(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *base, int reps,
                      FvGpuBruteOpts *best, double *best_us, int *ntried, char *errbuf,
                      size_t errlen) {
    FvGpuBruteOpts o0 = base ? *base : default_opts();
    o0.graph = 0;
    o0.qtw = 0;
    o0.wg_per_cu = 0;
    o0.sample_stride = 0;
    o0.order = 0;
    o0.waves = 0;
    const int prec = scan_precision(g, b, &o0);
    reps = std::max(1, std::min(reps, 64));
    std::vector<uint32_t> st((size_t)b->nq);
    QAux qa = qaux_of(b);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PV_TRY(hipEventCreate(&e0));
    PV_TRY(hipEventCreate(&e1));
    struct Ev {
        hipEvent_t a, b;
        ~Ev() { hipEventDestroy(a); hipEventDestroy(b); }
    } guard{e0, e1};

    /* one timed candidate: a warm run, then reps runs between events;
     * +inf when a query is not certified */
    auto trial = [&](const FvGpuBruteOpts &o, double *us) -> int {
        if (run_brute(g, b, &o, 0, errbuf, errlen) != 0) return -1;
        PV_TRY(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps; r++)
            if (run_brute(g, b, &o, 0, errbuf, errlen) != 0) return -1;
        PV_TRY(hipEventRecord(e1, b->stream));
        PV_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        PV_TRY(hipEventElapsedTime(&ms, e0, e1));
        PV_TRY(hipMemcpy(st.data(), qa.status, st.size() * 4, hipMemcpyDeviceToHost));
        bool ok = true;
        for (uint32_t v : st) ok &= v == 0;
        *us = ok ? 1e3 * ms / reps : HUGE_VAL;
        return 0;
    };

    double t0;
    if (trial(o0, &t0) != 0) return -1;
    *best = o0;
    *best_us = t0;
    *ntried = 1;
    if (b->last_route != 1) return 0;   /* gather route: no scan shape to tune */
    const bool mq = prec == FV_GPU_PREC_BF16 && g->dimb == 128 && b->nqt >= 2;
    /* the query-resident kernel: bitmap predicates, at least kQrNW tiles */
    const bool qr = mq && b->nqt >= kQrNW &&
                    (!(o0.inline_pred && b->inl_ok && b->inl_col >= 0) ||
                     (b->inl_rank && g->d_rank != nullptr));
    const int qtws[4] = {1, 2, 4, kQrNW};
    const bool wv = (prec == FV_GPU_PREC_BF16 || prec == FV_GPU_PREC_SQ8) && g->dim <= 256;
    for (int qi = 0; qi < (mq ? (qr ? 4 : 3) : 1); qi++)
        for (int waves = 4; waves <= ((qi == 0 && wv) ? 8 : 4); waves += 4)
        for (int wpc = 1; wpc <= (qi == 3 ? 2 : 4); wpc++)
        for (int order = 1; order <= (b->nqt >= 2 ? 3 : 2); order++) {
            if (qi == 3 && order == 2) continue;   /* no sweep order: same as chunked */
            FvGpuBruteOpts o = o0;
            o.qtw = qtws[qi];
            o.waves = qi == 0 && wv ? waves : 0;
            o.wg_per_cu = wpc;
            o.order = order;
            const uint32_t sa = scan_shape(g, b, &o, prec, 0).stride;
            const uint32_t cand[3] = {std::max<uint32_t>(1, sa / 2), sa,
                                      std::min<uint32_t>(64, sa * 2)};
            for (int si = 0; si < 3; si++) {
                if (si > 0 && cand[si] == cand[si - 1]) continue;
                o.sample_stride = cand[si];
                double t;
                if (trial(o, &t) != 0) return -1;
                ++*ntried;
                if (t < *best_us) {
                    *best_us = t;
                    *best = o;
                }
            }
        }
    if (!std::isfinite(*best_us))
        PV_FAIL("gpu: no launch shape certifies every query of this batch");
    return 0;
}

extern "C" int fv_gpu_batch_tune(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *base,
                                 int reps, FvGpuBruteOpts *best, double *best_us,
                                 int *ntried, char *errbuf, size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g || best == nullptr)
        PV_FAIL("gpu: bad batch");
    try {
        PV_TRY(hipSetDevice(g->device));
        const bool timing = g->timing;
        g->timing = false;
        double us = 0.0;
        int n = 0;
        const int rc = tune_batch(g, b, base, reps, best, &us, &n, errbuf, errlen);
        g->timing = timing;
        if (best_us) *best_us = us;
        if (ntried) *ntried = n;
        return rc;
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

/* Main-scan kernel time of a batch: one full run (so tau and the pool
 * are set up), then the same scan kernel launched reps times back to back
 * between two events on the batch's stream, then one more full run so the
 * batch's outputs and pool state are those of an ordinary run; returns
 * when that run has finished. */
extern "C" int fv_gpu_batch_time_scan(FvGpuIndex *g, FvGpuBatch *b, const FvGpuBruteOpts *opts,
                                      int reps, double *us, char *errbuf, size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g || us == nullptr || reps < 1)
        PV_FAIL("gpu: bad batch");
    try {
        PV_TRY(hipSetDevice(g->device));
        const bool timing = g->timing;
        g->timing = false;
        struct Restore {
            FvGpuIndex *g;
            bool t;
            ~Restore() { g->timing = t; }
        } restore{g, timing};
        if (run_brute(g, b, opts, 0, errbuf, errlen) != 0) return -1;
        if (b->last_route != 1 || b->last_scan == nullptr)
            PV_FAIL("gpu: time_scan: the batch did not take the full-scan route");
        const ScanLaunch L = *(const ScanLaunch *)b->last_scan;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        PV_TRY(hipEventCreate(&e0));
        struct Ev {
            hipEvent_t *a, *b;
            ~Ev() {
                if (*a) hipEventDestroy(*a);
                if (*b) hipEventDestroy(*b);
            }
        } guard{&e0, &e1};
        PV_TRY(hipEventCreate(&e1));
        PV_TRY(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps; r++) PV_TRY(launch_main_scan(L, b->stream));
        PV_TRY(hipEventRecord(e1, b->stream));
        PV_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        PV_TRY(hipEventElapsedTime(&ms, e0, e1));
        *us = 1e3 * (double)ms / reps;
        /* restore an ordinary run's outputs and wait for it: a run left in
         * flight would make the next batch's scan take the shared-chip
         * shape */
        if (run_brute(g, b, opts, 0, errbuf, errlen) != 0) return -1;
        PV_TRY(hipStreamSynchronize(b->stream));
        return 0;
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

extern "C" int fv_gpu_sync(FvGpuIndex *g, char *errbuf, size_t errlen) {
    if (g == nullptr) PV_FAIL("gpu: NULL index");
    PV_TRY(hipStreamSynchronize(g->stream));
    for (int i = 0; i < g->nstreams; i++) PV_TRY(hipStreamSynchronize(g->bstreams[i]));
    fill_stats_times(g);
    return 0;
}

extern "C" int fv_gpu_batch_fetch(FvGpuIndex *g, FvGpuBatch *b, FvHit *out,
                                  uint32_t *out_cnt, uint8_t *flags,
                                  char *errbuf, size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g) PV_FAIL("gpu: bad batch");
    if (fetch_raw(g, b, errbuf, errlen) != 0) return -1;
    const FvHit *ho;
    const uint32_t *hc, *hs;
    const uint8_t *hf;
    fetch_views(b, &ho, &hc, &hf, &hs);
    int rc = 0;
    for (int q = 0; q < b->nq; q++) {
        const uint32_t c = hc[q];
        if (out) memcpy(out + (uint64_t)q * b->out_stride, ho + (uint64_t)q * b->out_stride,
                        (size_t)c * sizeof(FvHit));
        if (out_cnt) out_cnt[q] = c;
        if (flags) flags[q] = hf[q];
        if (hs[q] != 0) rc = -1;
    }
    if (rc) snprintf(errbuf, errlen, "gpu: some queries need a re-run (candidate capacity)");
    return rc;
}

extern "C" void fv_gpu_batch_free(FvGpuBatch *b) {
    if (!b) return;
    if (b->owner) {
        hipSetDevice(b->owner->device);
        if (b->stream) hipStreamSynchronize(b->stream);
        if (b->owner->last_batch == b) b->owner->last_batch = nullptr;
    }
    free_batch_buffers(b);
    delete b;
}
