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

/* FP8 E4M3 (OCP "e4m3fn": 1 sign, 4 exponent bits with bias 7, 3 mantissa
 * bits; largest finite 448, subnormals m x 2^-9, no infinity, 0x7F / 0xFF
 * NaN): the code format of FP8 indexes, the operand format of
 * v_mfma_f32_32x32x16_fp8_fp8. Encoding rounds to nearest, ties to even,
 * saturates at +-448 and never writes -0 or NaN. Written out (not the
 * conversion instruction) so that the codes are a function stated here. */
__device__ __forceinline__ uint8_t e4m3_encode(float v) {
    const float a = fabsf(v);
    const uint32_t s = v < 0.f ? 0x80u : 0u;
    if (!(a < 448.f)) return (uint8_t)(s | 0x7Eu);   /* saturate (NaN too) */
    if (a < 0x1p-6f) {
        /* subnormals: multiples of 2^-9 (8 x 2^-9 is the smallest normal,
         * whose code is 0x08) */
        const uint32_t m = (uint32_t)rintf(a * 512.f);
        return (uint8_t)(m == 0 ? 0u : (s | m));
    }
    int e;
    (void)frexpf(a, &e);                       /* a in [2^(e-1), 2^e) */
    const float frac = ldexpf(a, 1 - e) - 1.f;  /* [0, 1), exact */
    uint32_t m = (uint32_t)rintf(frac * 8.f);
    int E = e - 1 + 7;
    if (m == 8u) {
        m = 0u;
        E++;
    }
    return (uint8_t)(s | ((uint32_t)E << 3) | m);
}
__device__ __forceinline__ float e4m3_decode(uint32_t c) {
    const int E = (int)((c >> 3) & 15u), m = (int)(c & 7u);
    const float a = E == 0 ? ldexpf((float)m, -9) : ldexpf(1.f + (float)m * 0.125f, E - 7);
    return (c & 0x80u) ? -a : a;
}

/* float -> bf16 bits, round to nearest even (NaN stays NaN). */
__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

/*
 * Lane exchange for the butterflies of the sorts and reductions:
 * lane_xor(x, j) is lane (l ^ j)'s x, for j in {1, 2, 4, 8, 16, 32}.
 *   j = 1, 2:      DPP quad_perm on a VALU move (no LDS round trip)
 *   j = 4, 8, 16:  ds_swizzle in bitmask mode (and 0x1f, xor j) inside each
 *                  32-lane half (no address register, unlike ds_bpermute)
 *   j = 32:        v_permlane32_swap (VALU): swapping lanes 32..63 of one
 *                  copy with lanes 0..31 of another leaves the upper half's
 *                  partners in the first result and the lower half's in
 *                  the second
 * j must fold to a constant once the caller's loop is unrolled (any other
 * value takes ds_bpermute), and every lane of the aligned group of 2j
 * lanes around the caller must be active: the partner's register is read
 * directly. Partners are exactly l ^ j, so a butterfly built on it
 * associates its sums exactly as one built on __shfl_xor.
 */
__device__ __forceinline__ uint32_t lane_xor(uint32_t x, int j) {
    switch (j) {
    case 1: return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);
    case 2: return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);
    case 4: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (4 << 10));
    case 8: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (8 << 10));
    case 16: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (16 << 10));
    case 32: {
        const auto p = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return (__lane_id() & 32) ? p[0] : p[1];
    }
    default: return (uint32_t)__shfl_xor((int)x, j);
    }
}
__device__ __forceinline__ float lane_xor(float x, int j) {
    return __uint_as_float(lane_xor(__float_as_uint(x), j));
}
__device__ __forceinline__ int lane_xor(int x, int j) {
    return (int)lane_xor((uint32_t)x, j);
}
__device__ __forceinline__ uint64_t lane_xor(uint64_t x, int j) {
    const uint32_t lo = lane_xor((uint32_t)x, j), hi = lane_xor((uint32_t)(x >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}

/*
 * Inclusive prefix sum across a 64-lane wave on DPP alone: Hillis-Steele
 * inside each 16-lane row (row_shr 1, 2, 4, 8; lanes shifted in from
 * outside the row read 0), then row 15's total into rows 1 and 3
 * (row_bcast:15, rows 0 and 2 masked off) and lane 31's into rows 2 and 3
 * (row_bcast:31). Every lane must be active.
 */
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return v;
}

/*
 * Scalar tail of the reference l2sq (dims past the last 16-dim block) as
 * the reference compiler emits `sum += d * d` (-O3): with >= 8 elements
 * left, one 8-wide in-order block with separately rounded products, then a
 * fused multiply-add per remaining element. Contraction is off here: the
 * library builds with -ffp-contract=fast, which would otherwise fuse the
 * 8-wide block's separately rounded products into the sums.
 */
__device__ __forceinline__ float tail_l2(const float *__restrict__ x,
                                         const float *__restrict__ q, int i,
                                         int dim, float t) {
#pragma clang fp contract(off)
    if (dim - i >= 8) {
        for (int j = 0; j < 8; j++, i++) {
            /* plain operators: this block's contract(off) applies to them
             * (the __f*_rn helpers are compiled with the file's setting) */
            const float d = x[i] - q[i];
            const float p = d * d;
            t = t + p;
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
#pragma clang fp contract(off)
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
    t = __fadd_rn(t, lane_xor(t, 8));
    t = __fadd_rn(t, lane_xor(t, 4));
    t = __fadd_rn(t, lane_xor(t, 2));
    t = __fadd_rn(t, lane_xor(t, 1));
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

/* make a wave's LDS writes visible to its own lanes */
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/*
 * Predicate bitmaps of kBmRows consecutive rows per 256-thread block.
 *
 * The block first stages its rows' values of the first kBmLdsCols columns
 * in LDS (dynamic shared memory of bitmap_lds_bytes(ncols) bytes), with the
 * program table, the first kBmOps ops and (when they fit) the IN values.
 * Then the four waves split the programs — wave w evaluates programs w,
 * w + 4, ... over all of the block's rows — so the serial chain of a wave
 * (decode, compare, ballot: latency-bound) is a quarter of the programs
 * rather than all of them. A wave covers the rows in kBmSlots slots of 64
 * (lane l of slot j: row j * 64 + l of the block), kBmChunk slots at a time.
 *
 * Single-op programs (one comparison, or one IN list: most filters) are
 * evaluated straight into lane masks: the ballot of one f64 compare per
 * slot is the result. Longer programs run the postfix interpreter on a
 * 64-bit per-lane stack and ballot its top. Each program's masks (its
 * kBmRows / 32 bitmap words) are staged in the wave's LDS slot and written
 * every kBmGroup programs as 16-byte runs (stores inside the program loop
 * would make later reuse of their registers wait for them). When progpart is
 * given, the block stores each program's passing-row count in it, layout
 * [nprog][bitmap blocks] (plain stores, zeros included: no clearing, no
 * atomics — per-block atomics on a few adjacent counters serialise in one
 * L2 channel).
 * Called by every thread of the block (it synchronises).
 */
constexpr int kBmRows = 1024;              /* rows per block */
constexpr int kBmSlots = kBmRows / 64;     /* 64-row slots per wave */
constexpr int kBmChunk = 8;                /* slots evaluated together */
constexpr int kBmLdsCols = 4;              /* columns staged in LDS */
constexpr int kBmOps = 128;                /* ops staged in LDS */
constexpr int kBmVals = 256;               /* IN values staged in LDS when they fit */
constexpr int kBmProgs = 1024;             /* max distinct programs (one per query) */
constexpr int kBmProgLds = 256;            /* program table entries staged in LDS */
constexpr int kBmGroup = 8;                /* programs per wave between write-outs */

/* Predicate bitmaps exist for batches of at most kBmProgs distinct
 * programs (one bitmap row each is allocated at staging). Every entry point
 * that builds or reads them checks this first; larger batches (up to
 * kMaxBatchEx queries) run only the brute path with inline predicates. */
inline int require_bitmaps(const FvGpuBatch *b, char *errbuf, size_t errlen) {
    if (b->nprog > kBmProgs)
        PV_FAIL("gpu: %d distinct predicates: predicate bitmaps hold at most %d", b->nprog,
                kBmProgs);
    return 0;
}

/* blocks of a bitmap pass over an index */
inline uint32_t bitmap_blocks(const FvGpuIndex *g) {
    return (uint32_t)((g->count_pad + kBmRows - 1) / kBmRows);
}

/* dynamic LDS of a bitmap block */
__host__ __device__ inline size_t bitmap_lds_bytes(int ncols) {
    return (size_t)(ncols < kBmLdsCols ? ncols : kBmLdsCols) * kBmRows * sizeof(double);
}

/* result masks of kBmSlots slots into a wave's staging slot (lane 0 writes;
 * the masks are wave-uniform); rows at or beyond `count` are cleared;
 * returns the number of passing rows */
__device__ __forceinline__ uint32_t bm_stage(uint64_t (&m)[kBmSlots], uint4 *stg, int lane,
                                             bool full, uint64_t r0, uint64_t count,
                                             bool counting) {
    if (!full) {
#pragma unroll
        for (int j = 0; j < kBmSlots; j++)
            m[j] &= __ballot(r0 + (uint64_t)j * 64 + lane < count);
    }
    uint32_t pc = 0;
    if (counting) {
#pragma unroll
        for (int j = 0; j < kBmSlots; j++) pc += (uint32_t)__popcll(m[j]);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kBmSlots / 2; k++)
            stg[k] = make_uint4((uint32_t)m[2 * k], (uint32_t)(m[2 * k] >> 32),
                                (uint32_t)m[2 * k + 1], (uint32_t)(m[2 * k + 1] >> 32));
    }
    return pc;
}

/* a single-comparison program over the block's rows staged in LDS (colp:
 * the column's kBmRows values): one compare per slot, its ballot is the
 * result; straight-line code per comparison kind */
template <int OP>
__device__ __forceinline__ uint32_t bm_cmp(const double *colp, double c, uint4 *stg, int lane,
                                           bool full, uint64_t r0, uint64_t count,
                                           bool counting) {
    double v[kBmSlots];
#pragma unroll
    for (int j = 0; j < kBmSlots; j++) v[j] = colp[j * 64 + lane];
    uint64_t m[kBmSlots];
#pragma unroll
    for (int j = 0; j < kBmSlots; j++) {
        bool r;
        if (OP == FV_CMP_LT) r = v[j] < c;
        else if (OP == FV_CMP_LE) r = v[j] <= c;
        else if (OP == FV_CMP_EQ) r = v[j] == c;
        else if (OP == FV_CMP_GE) r = v[j] >= c;
        else if (OP == FV_CMP_GT) r = v[j] > c;
        else r = (v[j] == v[j]) && (v[j] != c); /* NE: non-NULL and different */
        m[j] = __ballot(r);
    }
    return bm_stage(m, stg, lane, full, r0, count, counting);
}

/* a single-IN program over the block's rows staged in LDS */
__device__ __forceinline__ uint32_t bm_in(const double *colp, const double *inv, int nv,
                                          uint4 *stg, int lane, bool full, uint64_t r0,
                                          uint64_t count, bool counting) {
    double v[kBmSlots];
#pragma unroll
    for (int j = 0; j < kBmSlots; j++) v[j] = colp[j * 64 + lane];
    uint64_t m[kBmSlots];
#pragma unroll
    for (int j = 0; j < kBmSlots; j++) m[j] = 0;
    for (int k = 0; k < nv; k++) {
        const double x = inv[k];
#pragma unroll
        for (int j = 0; j < kBmSlots; j++) m[j] |= __ballot(v[j] == x);
    }
    return bm_stage(m, stg, lane, full, r0, count, counting);
}

/* The program, op and IN-value counts come from device memory (hdr =
 * [nprog, nops, nvals] in the batch's staging area), so a captured launch
 * serves any batch restaged into the same buffers. */
__device__ __forceinline__ void bitmap_rows(uint64_t block, const int32_t *__restrict__ hdr,
                                            const PvOp *__restrict__ ops,
                                            const int32_t *__restrict__ poff,
                                            const int32_t *__restrict__ plen,
                                            const double *__restrict__ vals,
                                            const double *__restrict__ cols, uint64_t count,
                                            int ncols, uint64_t words,
                                            uint32_t *__restrict__ bitmap,
                                            uint32_t *__restrict__ progpart, uint32_t nblocks) {
    constexpr int CH = kBmChunk;
    extern __shared__ double s_col[];          /* [min(ncols, kBmLdsCols)][kBmRows] */
    __shared__ PvOp s_ops[kBmOps];
    __shared__ int2 s_prog[kBmProgLds];
    __shared__ double s_vals[kBmVals];
    __shared__ uint4 s_stage[4][kBmGroup][kBmSlots / 2];
    __shared__ uint32_t s_pc[4][kBmGroup];
    const int t = threadIdx.x, lane = t & 63;
    const int nprog = __builtin_amdgcn_readfirstlane(hdr[0]);
    const int nops = __builtin_amdgcn_readfirstlane(hdr[1]);
    const int nvals = __builtin_amdgcn_readfirstlane(hdr[2]);
    /* wave index as a scalar: the program loop below is wave-uniform */
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint64_t r0 = block * kBmRows;
    const int nc = ncols < kBmLdsCols ? ncols : kBmLdsCols;
    const int nwin = nops < kBmOps ? nops : kBmOps;
    const bool vals_lds = nvals <= kBmVals;
    /* every global load of the staging is issued before any LDS write, so
     * the prologue costs one memory round trip, not one per table */
    static_assert(kBmOps <= 256 && kBmProgLds <= 256 && kBmVals <= 256, "one entry per thread");
    double cvals[kBmLdsCols][kBmRows / 256];
#pragma unroll
    for (int c = 0; c < kBmLdsCols; c++)
#pragma unroll
        for (int k = 0; k < kBmRows / 256; k++) {
            const uint64_t row = r0 + (uint64_t)k * 256 + t;
            cvals[c][k] = (c < nc && row < count) ? cols[(uint64_t)c * count + row] : 0.0;
        }
    const bool has_prog = t < nprog && t < kBmProgLds;
    const int2 prog = has_prog ? make_int2(poff[t], plen[t]) : make_int2(0, 0);
    const bool has_op = t < nwin;
    const PvOp op = has_op ? ops[t] : PvOp{};
    const bool has_val = vals_lds && t < nvals;
    const double val = has_val ? vals[t] : 0.0;
#pragma unroll
    for (int c = 0; c < kBmLdsCols; c++)
        if (c < nc) {
#pragma unroll
            for (int k = 0; k < kBmRows / 256; k++) s_col[c * kBmRows + k * 256 + t] = cvals[c][k];
        }
    if (has_prog) s_prog[t] = prog;
    if (has_op) s_ops[t] = op;
    if (has_val) s_vals[t] = val;
    __syncthreads();

    const bool full = r0 + kBmRows <= count;
    /* a column's values for slots [j0, j0 + CH) of the lane (c is uniform) */
    auto colv = [&](int c, int j0, double (&v)[CH]) {
        if (c < nc) {
#pragma unroll
            for (int j = 0; j < CH; j++) v[j] = s_col[c * kBmRows + (j0 + j) * 64 + lane];
        } else {
#pragma unroll
            for (int j = 0; j < CH; j++) {
                const uint64_t row = r0 + (uint64_t)(j0 + j) * 64 + lane;
                v[j] = row < count ? cols[(uint64_t)c * count + row] : 0.0;
            }
        }
    };
    /* write out the wave's staged programs: ordinals q0 .. q0 + ng - 1 of
     * the wave are programs 4 q + wv; lane l stores words 4 (l & 7) .. +3
     * of staged program l >> 3 */
    auto flush = [&](int q0, int ng) {
        wave_lds_sync();
        const int gi = lane >> 3, part = lane & 7;
        if (gi < ng) {
            const uint4 v = s_stage[wv][gi][part];
            const uint64_t w = block * (kBmRows / 32) + 4 * part;
            uint2 *dst = (uint2 *)(bitmap + (uint64_t)(4 * (q0 + gi) + wv) * words + w);
            if (w < words) dst[0] = make_uint2(v.x, v.y);
            if (w + 2 < words) dst[1] = make_uint2(v.z, v.w);
        }
        if (progpart != nullptr && lane < ng)
            progpart[(uint64_t)(4 * (q0 + lane) + wv) * nblocks + block] = s_pc[wv][lane];
        wave_lds_sync();
    };

    for (int p = wv; p < nprog; p += 4) {
        const int2 pr = p < kBmProgLds ? s_prog[p] : make_int2(poff[p], plen[p]);
        const int o = __builtin_amdgcn_readfirstlane(pr.x);
        const int n = __builtin_amdgcn_readfirstlane(pr.y);
        const int q = p >> 2, gs = q % kBmGroup;
        uint32_t pc = 0;
        /* single-op programs on a staged column: straight-line code */
        const bool cnt = progpart != nullptr;
        bool done = false;
        if (n == 1 && o < nwin) {
            const int32_t *ow = (const int32_t *)&s_ops[o];
            const int kind = __builtin_amdgcn_readfirstlane(ow[0]);
            const int col = __builtin_amdgcn_readfirstlane(ow[1]);
            const int oa = __builtin_amdgcn_readfirstlane(ow[2]);
            const int ob = __builtin_amdgcn_readfirstlane(ow[3]);
            const double c = __hiloint2double(__builtin_amdgcn_readfirstlane(ow[5]),
                                              __builtin_amdgcn_readfirstlane(ow[4]));
            if (col < nc && (kind == OP_CMP || (kind == OP_IN && vals_lds))) {
                const double *colp = s_col + col * kBmRows;
                uint4 *stg = s_stage[wv][gs];
                if (kind == OP_IN) pc = bm_in(colp, s_vals + oa, ob, stg, lane, full, r0, count, cnt);
                else if (oa == FV_CMP_LT) pc = bm_cmp<FV_CMP_LT>(colp, c, stg, lane, full, r0, count, cnt);
                else if (oa == FV_CMP_LE) pc = bm_cmp<FV_CMP_LE>(colp, c, stg, lane, full, r0, count, cnt);
                else if (oa == FV_CMP_EQ) pc = bm_cmp<FV_CMP_EQ>(colp, c, stg, lane, full, r0, count, cnt);
                else if (oa == FV_CMP_GE) pc = bm_cmp<FV_CMP_GE>(colp, c, stg, lane, full, r0, count, cnt);
                else if (oa == FV_CMP_GT) pc = bm_cmp<FV_CMP_GT>(colp, c, stg, lane, full, r0, count, cnt);
                else pc = bm_cmp<FV_CMP_NE>(colp, c, stg, lane, full, r0, count, cnt);
                done = true;
            }
        }
        for (int j0 = 0; j0 < (done ? 0 : kBmSlots); j0 += CH) {
            uint64_t m[CH];
            uint64_t st[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) st[j] = 0;
            for (int i = 0; i < n; i++) {
                /* op fields into scalar registers: every dispatch below is
                 * a scalar branch taken once per op for all of the slots */
                const int oi = o + i;
                int32_t f[6];
                if (oi < nwin) {
                    const int32_t *ow = (const int32_t *)&s_ops[oi];
#pragma unroll
                    for (int k = 0; k < 6; k++) f[k] = ow[k];
                } else {
                    const int32_t *ow = (const int32_t *)&ops[oi];
#pragma unroll
                    for (int k = 0; k < 6; k++) f[k] = ow[k];
                }
                const int kind = __builtin_amdgcn_readfirstlane(f[0]);
                const int col = __builtin_amdgcn_readfirstlane(f[1]);
                const int oa = __builtin_amdgcn_readfirstlane(f[2]);
                const int ob = __builtin_amdgcn_readfirstlane(f[3]);
                const double c = __hiloint2double(__builtin_amdgcn_readfirstlane(f[5]),
                                                  __builtin_amdgcn_readfirstlane(f[4]));
                if (n == 1 && kind == OP_CMP) {
                    /* the whole program is one comparison: its ballots are
                     * the result */
                    double v[CH];
                    colv(col, j0, v);
                    switch (oa) {
                        case FV_CMP_LT:
#pragma unroll
                            for (int j = 0; j < CH; j++) m[j] = __ballot(v[j] < c);
                            break;
                        case FV_CMP_LE:
#pragma unroll
                            for (int j = 0; j < CH; j++) m[j] = __ballot(v[j] <= c);
                            break;
                        case FV_CMP_EQ:
#pragma unroll
                            for (int j = 0; j < CH; j++) m[j] = __ballot(v[j] == c);
                            break;
                        case FV_CMP_GE:
#pragma unroll
                            for (int j = 0; j < CH; j++) m[j] = __ballot(v[j] >= c);
                            break;
                        case FV_CMP_GT:
#pragma unroll
                            for (int j = 0; j < CH; j++) m[j] = __ballot(v[j] > c);
                            break;
                        default: /* NE: non-NULL and different */
#pragma unroll
                            for (int j = 0; j < CH; j++)
                                m[j] = __ballot((v[j] == v[j]) && (v[j] != c));
                    }
                    break;
                }
                if (n == 1 && kind == OP_IN) {
                    double v[CH];
                    colv(col, j0, v);
#pragma unroll
                    for (int j = 0; j < CH; j++) m[j] = 0;
                    for (int k = 0; k < ob; k++) {
                        const double x = vals_lds ? s_vals[oa + k] : vals[oa + k];
#pragma unroll
                        for (int j = 0; j < CH; j++) m[j] |= __ballot(v[j] == x);
                    }
                    break;
                }
                uint32_t r[CH];
                if (kind == OP_CMP) {
                    double v[CH];
                    colv(col, j0, v);
                    switch (oa) {
                        case FV_CMP_LT:
#pragma unroll
                            for (int j = 0; j < CH; j++) r[j] = v[j] < c;
                            break;
                        case FV_CMP_LE:
#pragma unroll
                            for (int j = 0; j < CH; j++) r[j] = v[j] <= c;
                            break;
                        case FV_CMP_EQ:
#pragma unroll
                            for (int j = 0; j < CH; j++) r[j] = v[j] == c;
                            break;
                        case FV_CMP_GE:
#pragma unroll
                            for (int j = 0; j < CH; j++) r[j] = v[j] >= c;
                            break;
                        case FV_CMP_GT:
#pragma unroll
                            for (int j = 0; j < CH; j++) r[j] = v[j] > c;
                            break;
                        default: /* NE: non-NULL and different */
#pragma unroll
                            for (int j = 0; j < CH; j++) r[j] = (v[j] == v[j]) && (v[j] != c);
                    }
#pragma unroll
                    for (int j = 0; j < CH; j++) st[j] = (st[j] << 1) | r[j];
                } else if (kind == OP_IN) {
                    double v[CH];
                    colv(col, j0, v);
#pragma unroll
                    for (int j = 0; j < CH; j++) r[j] = 0;
                    for (int k = 0; k < ob; k++) {
                        const double x = vals_lds ? s_vals[oa + k] : vals[oa + k];
#pragma unroll
                        for (int j = 0; j < CH; j++) r[j] |= (v[j] == x) ? 1u : 0u;
                    }
#pragma unroll
                    for (int j = 0; j < CH; j++) st[j] = (st[j] << 1) | r[j];
                } else if (kind == OP_AND2) {
#pragma unroll
                    for (int j = 0; j < CH; j++)
                        st[j] = ((st[j] >> 2) << 1) | (st[j] & (st[j] >> 1) & 1u);
                } else if (kind == OP_OR2) {
#pragma unroll
                    for (int j = 0; j < CH; j++)
                        st[j] = ((st[j] >> 2) << 1) | ((st[j] | (st[j] >> 1)) & 1u);
                } else if (kind == OP_NOT) {
#pragma unroll
                    for (int j = 0; j < CH; j++) st[j] ^= 1u;
                } else if (kind == OP_TRUE) {
#pragma unroll
                    for (int j = 0; j < CH; j++) st[j] = (st[j] << 1) | 1u;
                } else { /* OP_FALSE */
#pragma unroll
                    for (int j = 0; j < CH; j++) st[j] <<= 1;
                }
                if (i == n - 1) {
#pragma unroll
                    for (int j = 0; j < CH; j++) m[j] = __ballot((st[j] & 1u) != 0);
                }
            }
            if (n == 0) {
#pragma unroll
                for (int j = 0; j < CH; j++) m[j] = 0;
            }
            if (!full) {
#pragma unroll
                for (int j = 0; j < CH; j++)
                    m[j] &= __ballot(r0 + (uint64_t)(j0 + j) * 64 + lane < count);
            }
            if (cnt) {
#pragma unroll
                for (int j = 0; j < CH; j++) pc += (uint32_t)__popcll(m[j]);
            }
            if (lane == 0) {
#pragma unroll
                for (int k = 0; k < CH / 2; k++)
                    s_stage[wv][gs][j0 / 2 + k] =
                        make_uint4((uint32_t)m[2 * k], (uint32_t)(m[2 * k] >> 32),
                                   (uint32_t)m[2 * k + 1], (uint32_t)(m[2 * k + 1] >> 32));
            }
        }
        if (lane == 0) s_pc[wv][gs] = pc;
        if (gs == kBmGroup - 1 || p + 4 >= nprog) flush(q - gs, gs + 1);
    }
}

/* ---- block-level select and sort (256-thread blocks) ---- */

/* inclusive prefix sum over a 256-thread block */
__device__ __forceinline__ uint32_t block_scan_incl(uint32_t v, uint32_t *tmp) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_scan_incl(v);
    if (l == 63) tmp[w] = v;
    __syncthreads();
    uint32_t add = 0;
    for (int i = 0; i < w; i++) add += tmp[i];
    __syncthreads();
    return v + add;
}

/*
 * k-th smallest (1-based) of n unique u64 keys, 8-bit MSD radix select
 * over a 256-thread block. Keys are read through `get`.
 */
template <typename Get>
__device__ uint64_t block_kth(Get get, uint32_t n, uint32_t k, uint32_t *hist,
                              uint32_t *tmp, uint32_t *bc) {
    uint64_t prefix = 0;
    const int t = threadIdx.x;
    for (int shift = 56; shift >= 0; shift -= 8) {
        hist[t] = 0;
        __syncthreads();
        for (uint32_t i = t; i < n; i += blockDim.x) {
            const uint64_t key = get(i);
            if (shift == 56 || (key >> (shift + 8)) == prefix)
                atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        const uint32_t c = hist[t];
        const uint32_t incl = block_scan_incl(c, tmp);
        const uint32_t excl = incl - c;
        if (excl < k && k <= incl) {
            bc[0] = t;
            bc[1] = excl;
            bc[2] = c;
        }
        __syncthreads();
        const uint32_t dg = bc[0], ex = bc[1], cnt = bc[2];
        prefix = (prefix << 8) | dg;
        k -= ex;
        __syncthreads();
        if (cnt == 1 && shift > 0) {
            /* unique key under this prefix: fetch it */
            for (uint32_t i = t; i < n; i += blockDim.x) {
                const uint64_t key = get(i);
                if ((key >> shift) == prefix) {
                    bc[4] = (uint32_t)key;
                    bc[5] = (uint32_t)(key >> 32);
                }
            }
            __syncthreads();
            const uint64_t r = ((uint64_t)bc[5] << 32) | bc[4];
            __syncthreads();
            return r;
        }
    }
    return prefix;
}

/* ascending bitonic sort of n (power of two) keys in LDS */
__device__ inline void block_bitonic(uint64_t *a, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
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

/* ascending bitonic sort of one key per lane across a 64-lane wave
 * (register-resident: DPP, swizzle and permlane exchanges, no LDS data) */
__device__ __forceinline__ uint64_t wave_sort64(uint64_t x) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t y = lane_xor(x, j);
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
#pragma unroll
        for (int off = 32; off; off >>= 1) m = fminf(m, lane_xor(m, off));
        const uint64_t who = __ballot(h == m);
        const int lane = __ffsll((unsigned long long)who) - 1;
        if (l == lane) head++;
        if (l == 0) out[r] = m;
    }
}

} // namespace pv

#endif
