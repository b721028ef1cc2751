/*
 * pv_device.h — device helpers shared by the HIP kernels.
 */
#ifndef PV_DEVICE_H
#define PV_DEVICE_H

#include "pv_internal.h"

namespace pv {

/* (dist, row) packed so that unsigned key order equals the lexicographic
 * (dist, row) order for non-negative distances (gpu_engine.cu:435-438). */
__device__ __forceinline__ uint64_t make_key(float d, uint32_t row) {
    return ((uint64_t)__float_as_uint(d) << 32) | row;
}
__device__ __forceinline__ float key_dist(uint64_t k) {
    return __uint_as_float((uint32_t)(k >> 32));
}
__device__ __forceinline__ uint32_t key_row(uint64_t k) { return (uint32_t)k; }

/* float -> bf16 bits, round to nearest even (NaN stays NaN). */
__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

/*
 * Scalar tail of the reference l2sq (dims past the last 16-dim block) as
 * the reference compiler emits `sum += d * d` (-O3): with >= 8 elements
 * left, one 8-wide in-order block with separately rounded products, then a
 * fused multiply-add per remaining element.
 */
__device__ __forceinline__ float tail_l2(const float *__restrict__ x,
                                         const float *__restrict__ q, int i,
                                         int dim, float t) {
    if (dim - i >= 8) {
        for (int j = 0; j < 8; j++, i++) {
            const float d = __fsub_rn(x[i], q[i]);
            t = __fadd_rn(t, __fmul_rn(d, d));
        }
    }
    for (; i < dim; i++) {
        const float d = __fsub_rn(x[i], q[i]);
        t = __fmaf_rn(d, d, t);
    }
    return t;
}

/*
 * Exact FP32 squared distance in the summation order of the reference CPU
 * cursor's AVX-512 l2sq (engine.cpp:43-63): 16 accumulator lanes, each over
 * dims i+j of 32-dim blocks (acc0) and i+16+j (acc1) with fused
 * multiply-add, one trailing 16-dim block into acc0, then the reduction
 * tree ((t[j] + t[j+8]) ... ) of _mm512_reduce_add_ps and a scalar tail.
 * Called by 16 consecutive lanes (lane16 = lane & 15) that hold the same
 * row/query; the result is valid in every lane of the group. No
 * contraction: every operation is an explicit round-to-nearest intrinsic.
 */
__device__ __forceinline__ float cpu_order_l2(const float *__restrict__ x,
                                              const float *__restrict__ q,
                                              int dim, int lane16) {
    float acc0 = 0.f, acc1 = 0.f;
    int i = 0;
    for (; i + 32 <= dim; i += 32) {
        float d0 = __fsub_rn(x[i + lane16], q[i + lane16]);
        float d1 = __fsub_rn(x[i + 16 + lane16], q[i + 16 + lane16]);
        acc0 = __fmaf_rn(d0, d0, acc0);
        acc1 = __fmaf_rn(d1, d1, acc1);
    }
    for (; i + 16 <= dim; i += 16) {
        float d = __fsub_rn(x[i + lane16], q[i + lane16]);
        acc0 = __fmaf_rn(d, d, acc0);
    }
    float t = __fadd_rn(acc0, acc1);
    t = __fadd_rn(t, __shfl_xor(t, 8, 16));
    t = __fadd_rn(t, __shfl_xor(t, 4, 16));
    t = __fadd_rn(t, __shfl_xor(t, 2, 16));
    t = __fadd_rn(t, __shfl_xor(t, 1, 16));
    t = tail_l2(x, q, i, dim, t);
    return t;
}

/*
 * Postfix predicate program evaluation for one row. Semantics of
 * FvPred::eval_node (engine.cpp:397-430): comparisons in double, NULL
 * stored as NaN (all comparisons false), NE = v == v && v != c, IN = any
 * equal. The stack is a 64-bit shift register (depth checked on the host).
 * colv(c) returns the row's value in column c.
 */
/* comparison through the op's outcome mask (kCmpPassMask), branch-free */
__device__ __forceinline__ uint32_t cmp_pass(double v, double c, int32_t mask) {
    const int o = (v != v) ? 3 : (v < c) ? 0 : (v == c) ? 1 : (v > c) ? 2 : 4;
    return ((uint32_t)mask >> o) & 1u;
}

template <typename ColFn>
__device__ __forceinline__ bool eval_prog_fn(const PvOp *__restrict__ ops, int n,
                                             const double *__restrict__ vals, ColFn colv) {
    uint64_t st = 0;
    for (int i = 0; i < n; i++) {
        const PvOp op = ops[i];
        switch (op.kind) {
            case OP_CMP: {
                st = (st << 1) | (uint64_t)cmp_pass(colv(op.col), op.value, op.b);
                break;
            }
            case OP_IN: {
                const double v = colv(op.col);
                bool r = false;
                for (int j = 0; j < op.b; j++) r |= (v == vals[op.a + j]);
                st = (st << 1) | (uint64_t)r;
                break;
            }
            case OP_AND2: {
                uint64_t t = st & (st >> 1) & 1u;
                st = ((st >> 2) << 1) | t;
                break;
            }
            case OP_OR2: {
                uint64_t t = (st | (st >> 1)) & 1u;
                st = ((st >> 2) << 1) | t;
                break;
            }
            case OP_NOT: st ^= 1u; break;
            case OP_TRUE: st = (st << 1) | 1u; break;
            default: st = (st << 1); break; /* OP_FALSE */
        }
    }
    return (st & 1u) != 0;
}

__device__ __forceinline__ bool eval_prog(const PvOp *__restrict__ ops, int n,
                                          const double *__restrict__ vals,
                                          const double *__restrict__ cols,
                                          uint64_t count, uint32_t row) {
    return eval_prog_fn(ops, n, vals,
                        [&](int c) { return cols[(uint64_t)c * count + row]; });
}

/*
 * Predicate bitmaps of kBmRows consecutive rows per 256-thread block, row
 * r = block * kBmRows + j * 256 + threadIdx.x for j < kBmRowsPerThread.
 * Every distinct program is evaluated for the block's rows with each op
 * decoded once for all of a thread's rows; the rows' first four column
 * values stay in registers across programs. Program offsets, a window of
 * kBmOps ops and (when they fit) the IN values are staged in LDS so that op
 * decoding costs LDS reads instead of chains of dependent scalar-cache
 * misses. Result words of kBmGroup programs are collected in LDS and
 * written as whole 128-byte runs (one run per program per block). When
 * progcnt is given, each program's passing-row count is added to it (one
 * atomic per program per block; the caller zeroes it beforehand).
 * Called by every thread of the block (it synchronises).
 */
constexpr int kBmRowsPerThread = 8;
constexpr int kBmRows = 256 * kBmRowsPerThread;
constexpr int kBmOps = 512;      /* LDS op window */
constexpr int kBmVals = 256;     /* IN values staged in LDS when they fit */
constexpr int kBmProgs = 1024;   /* max distinct programs (one per query) */
constexpr int kBmGroup = 16;     /* programs per coalesced write-out */

__device__ __forceinline__ void bitmap_rows(uint64_t block, int nprog, int nops,
                                            const PvOp *__restrict__ ops,
                                            const int32_t *__restrict__ poff,
                                            const int32_t *__restrict__ plen,
                                            const double *__restrict__ vals, int nvals,
                                            const double *__restrict__ cols, uint64_t count,
                                            int ncols, uint64_t words,
                                            uint32_t *__restrict__ bitmap,
                                            uint32_t *__restrict__ progcnt) {
    constexpr int R = kBmRowsPerThread;
    __shared__ PvOp s_ops[kBmOps];
    __shared__ int32_t s_off[kBmProgs], s_len[kBmProgs];
    __shared__ double s_vals[kBmVals];
    __shared__ uint64_t s_bm[kBmGroup][4 * R];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int i = t; i < nprog; i += blockDim.x) {
        s_off[i] = poff[i];
        s_len[i] = plen[i];
    }
    const bool vals_lds = nvals <= kBmVals;
    if (vals_lds)
        for (int i = t; i < nvals; i += blockDim.x) s_vals[i] = vals[i];

    uint64_t row[R];
    bool live[R];
    double cv[R][4];
#pragma unroll
    for (int j = 0; j < R; j++) {
        row[j] = block * kBmRows + (uint64_t)j * 256 + t;
        live[j] = row[j] < count;
#pragma unroll
        for (int c = 0; c < 4; c++)
            cv[j][c] = (c < ncols && live[j]) ? cols[(uint64_t)c * count + row[j]] : 0.0;
    }
    /* one column's values for all of the thread's rows (c is uniform) */
    auto colv = [&](int c, double (&v)[R]) {
        switch (c) {
            case 0:
#pragma unroll
                for (int j = 0; j < R; j++) v[j] = cv[j][0];
                break;
            case 1:
#pragma unroll
                for (int j = 0; j < R; j++) v[j] = cv[j][1];
                break;
            case 2:
#pragma unroll
                for (int j = 0; j < R; j++) v[j] = cv[j][2];
                break;
            case 3:
#pragma unroll
                for (int j = 0; j < R; j++) v[j] = cv[j][3];
                break;
            default:
#pragma unroll
                for (int j = 0; j < R; j++)
                    v[j] = live[j] ? cols[(uint64_t)c * count + row[j]] : 0.0;
        }
    };
    int win_lo = 0, win_hi = 0; /* ops[win_lo, win_hi) are in s_ops */
    __syncthreads();
    for (int p = 0; p < nprog; p++) {
        /* block-uniform: every thread walks the same programs and ops */
        const int o = __builtin_amdgcn_readfirstlane(s_off[p]);
        const int n = __builtin_amdgcn_readfirstlane(s_len[p]);
        uint64_t st[R];
#pragma unroll
        for (int j = 0; j < R; j++) st[j] = 0;
        for (int i = 0; i < n; i++) {
            const int oi = o + i;
            if (oi >= win_hi) {
                __syncthreads();
                win_lo = oi;
                win_hi = min(oi + kBmOps, nops);
                for (int k = t; k < win_hi - win_lo; k += blockDim.x) s_ops[k] = ops[win_lo + k];
                __syncthreads();
            }
            /* op fields into scalar registers: every dispatch below is a
             * scalar branch taken once per op for all of the thread's rows */
            const int32_t *ow = (const int32_t *)&s_ops[oi - win_lo];
            const int kind = __builtin_amdgcn_readfirstlane(ow[0]);
            const int col = __builtin_amdgcn_readfirstlane(ow[1]);
            const int oa = __builtin_amdgcn_readfirstlane(ow[2]);
            const int ob = __builtin_amdgcn_readfirstlane(ow[3]);
            const double c = __hiloint2double(__builtin_amdgcn_readfirstlane(ow[5]),
                                              __builtin_amdgcn_readfirstlane(ow[4]));
            uint32_t r[R];
            if (kind == OP_CMP) {
                double v[R];
                colv(col, v);
                switch (oa) {
                    case FV_CMP_LT:
#pragma unroll
                        for (int j = 0; j < R; j++) r[j] = v[j] < c;
                        break;
                    case FV_CMP_LE:
#pragma unroll
                        for (int j = 0; j < R; j++) r[j] = v[j] <= c;
                        break;
                    case FV_CMP_EQ:
#pragma unroll
                        for (int j = 0; j < R; j++) r[j] = v[j] == c;
                        break;
                    case FV_CMP_GE:
#pragma unroll
                        for (int j = 0; j < R; j++) r[j] = v[j] >= c;
                        break;
                    case FV_CMP_GT:
#pragma unroll
                        for (int j = 0; j < R; j++) r[j] = v[j] > c;
                        break;
                    default: /* NE: non-NULL and different */
#pragma unroll
                        for (int j = 0; j < R; j++) r[j] = (v[j] == v[j]) && (v[j] != c);
                }
#pragma unroll
                for (int j = 0; j < R; j++) st[j] = (st[j] << 1) | r[j];
            } else if (kind == OP_IN) {
                double v[R];
                colv(col, v);
#pragma unroll
                for (int j = 0; j < R; j++) r[j] = 0;
                if (vals_lds) {
                    for (int k = 0; k < ob; k++) {
                        const double x = s_vals[oa + k];
#pragma unroll
                        for (int j = 0; j < R; j++) r[j] |= (v[j] == x) ? 1u : 0u;
                    }
                } else {
                    for (int k = 0; k < ob; k++) {
                        const double x = vals[oa + k];
#pragma unroll
                        for (int j = 0; j < R; j++) r[j] |= (v[j] == x) ? 1u : 0u;
                    }
                }
#pragma unroll
                for (int j = 0; j < R; j++) st[j] = (st[j] << 1) | r[j];
            } else if (kind == OP_AND2) {
#pragma unroll
                for (int j = 0; j < R; j++)
                    st[j] = ((st[j] >> 2) << 1) | (st[j] & (st[j] >> 1) & 1u);
            } else if (kind == OP_OR2) {
#pragma unroll
                for (int j = 0; j < R; j++)
                    st[j] = ((st[j] >> 2) << 1) | ((st[j] | (st[j] >> 1)) & 1u);
            } else if (kind == OP_NOT) {
#pragma unroll
                for (int j = 0; j < R; j++) st[j] ^= 1u;
            } else if (kind == OP_TRUE) {
#pragma unroll
                for (int j = 0; j < R; j++) st[j] = (st[j] << 1) | 1u;
            } else { /* OP_FALSE */
#pragma unroll
                for (int j = 0; j < R; j++) st[j] <<= 1;
            }
        }
        const int slot = p % kBmGroup;
#pragma unroll
        for (int j = 0; j < R; j++) {
            /* rows j*256 + wv*64 .. +63 of the block: 64-bit word 4j + wv */
            const uint64_t m = __ballot(live[j] && (st[j] & 1u));
            if (lane == 0) s_bm[slot][4 * j + wv] = m;
        }
        if (slot == kBmGroup - 1 || p == nprog - 1) {
            __syncthreads();
            const int pb = p - slot;
            for (int i = t; i < (slot + 1) * 2 * 4 * R; i += blockDim.x) {
                const int pp = i / (8 * R), wd = i % (8 * R);
                const uint64_t w = block * (kBmRows / 32) + wd;
                if (w < words)
                    bitmap[(uint64_t)(pb + pp) * words + w] =
                        (uint32_t)(s_bm[pp][wd >> 1] >> (32 * (wd & 1)));
            }
            if (progcnt != nullptr && t <= slot) {
                /* passing rows of program pb + t in this block */
                uint32_t c = 0;
                for (int k = 0; k < 4 * R; k++) c += (uint32_t)__popcll(s_bm[t][k]);
                if (c) atomicAdd(&progcnt[pb + t], c);
            }
            __syncthreads();
        }
    }
}

/* 64-bit lane exchange (two 32-bit bpermutes) */
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int m) {
    const int lo = __shfl_xor((int)(uint32_t)x, m);
    const int hi = __shfl_xor((int)(uint32_t)(x >> 32), m);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

/* ascending bitonic sort of one key per lane across a 64-lane wave
 * (register-resident, no LDS) */
__device__ __forceinline__ uint64_t wave_sort64(uint64_t x) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t y = shfl_xor_u64(x, j);
            const bool keep_min = ((l & j) == 0) == ((l & k) == 0);
            const uint64_t lo = x < y ? x : y, hi = x < y ? y : x;
            x = keep_min ? lo : hi;
        }
    }
    return x;
}

/* ---- learned-router features (shared by the query prep of a run and by
 * the feature / decision kernels, so every path computes the same values) */

constexpr int kNear = 16;   /* nearest-centroid distance features */

/* make a wave's LDS writes visible to its own lanes */
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* squared distance of the query in xs (LDS) to centroid c of the
 * transposed centroid array [dim][ncent]: 32 coordinates per step, all
 * loads issued before the arithmetic (L2-latency bound), four partial sums.
 * The one summation order of every feature path (training and inference). */
__device__ __forceinline__ float cent_dist(const float *xs, const float *__restrict__ cent_t,
                                           int c, int dim, int ncent) {
    const float *ck = cent_t + c;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int k = 0;
    for (; k + 32 <= dim; k += 32) {
        float v[32];
#pragma unroll
        for (int j = 0; j < 32; j++) v[j] = ck[(uint64_t)(k + j) * ncent];
#pragma unroll
        for (int j = 0; j < 32; j += 4) {
            const float d0 = xs[k + j] - v[j], d1 = xs[k + j + 1] - v[j + 1];
            const float d2 = xs[k + j + 2] - v[j + 2], d3 = xs[k + j + 3] - v[j + 3];
            s0 = fmaf(d0, d0, s0);
            s1 = fmaf(d1, d1, s1);
            s2 = fmaf(d2, d2, s2);
            s3 = fmaf(d3, d3, s3);
        }
    }
    for (; k < dim; k++) {
        const float d = xs[k] - ck[(uint64_t)k * ncent];
        s0 = fmaf(d, d, s0);
    }
    return (s0 + s1) + (s2 + s3);
}

/* lane-sorted insertion of v into the lane's kNear smallest */
__device__ __forceinline__ void near_insert(float (&best)[kNear], float v) {
#pragma unroll
    for (int k = 0; k < kNear; k++) {
        const float lo = fminf(best[k], v), hi = fmaxf(best[k], v);
        best[k] = lo;
        v = hi;
    }
}

/* the wave's kNear smallest of all lanes' lists, ascending, to out[0..kNear) */
__device__ __forceinline__ void wave_nearest(const float (&best)[kNear], float *out) {
    const int l = threadIdx.x & 63;
    int head = 0;
    for (int r = 0; r < kNear; r++) {
        float h = __uint_as_float(0x7f800000u);
#pragma unroll
        for (int k = 0; k < kNear; k++)
            if (k == head) h = best[k];
        float m = h;
        for (int off = 32; off; off >>= 1) m = fminf(m, __shfl_xor(m, off));
        const uint64_t who = __ballot(h == m);
        const int lane = __ffsll((unsigned long long)who) - 1;
        if (l == lane) head++;
        if (l == 0) out[r] = m;
    }
}

} // namespace pv

#endif
