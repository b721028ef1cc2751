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
 */
__device__ __forceinline__ bool eval_prog(const PvOp *__restrict__ ops, int n,
                                          const double *__restrict__ vals,
                                          const double *__restrict__ cols,
                                          uint64_t count, uint32_t row) {
    uint64_t st = 0;
    for (int i = 0; i < n; i++) {
        const PvOp op = ops[i];
        switch (op.kind) {
            case OP_CMP: {
                double v = cols[(uint64_t)op.col * count + row];
                bool r;
                switch (op.a) {
                    case FV_CMP_LT: r = v < op.value; break;
                    case FV_CMP_LE: r = v <= op.value; break;
                    case FV_CMP_EQ: r = v == op.value; break;
                    case FV_CMP_GE: r = v >= op.value; break;
                    case FV_CMP_GT: r = v > op.value; break;
                    case FV_CMP_NE: r = (v == v) && (v != op.value); break;
                    default: r = false;
                }
                st = (st << 1) | (uint64_t)r;
                break;
            }
            case OP_IN: {
                double v = cols[(uint64_t)op.col * count + row];
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

} // namespace pv

#endif
