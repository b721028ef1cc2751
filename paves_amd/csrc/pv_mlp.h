/*
 * pv_mlp.h — inference side of the learned router, shared by the decision
 * kernels (pv_learn.hip) and the threshold launch that runs the decision
 * blocks beside the threshold blocks (pv_brute.hip).
 *
 * MLP: x = [ q (dim) | log10 selectivity | distances to the kNear nearest
 * centroids ] (dim + 17 inputs), two hidden layers of 64 (ReLU), two
 * outputs (log ef, brute-vs-graph logit); BF16 inference weights with FP32
 * accumulation.
 */
#pragma once

#include "pv_device.h"

namespace pv {

constexpr int kHid = 64;
constexpr int kMaxFeat = 1024 + 17;

__host__ __device__ inline int model_nfeat(int dim) { return dim + 1 + kNear; }
__host__ __device__ inline int model_nparams(int F) { return kHid * F + kHid + kHid * kHid + kHid + 2 * kHid + 2; }

struct MlpLayout {
    int F;
    int w1, b1, w2, b2, w3, b3, n;
};
__host__ __device__ inline MlpLayout mlp_layout(int F) {
    MlpLayout L;
    L.F = F;
    L.w1 = 0;
    L.b1 = L.w1 + kHid * F;
    L.w2 = L.b1 + kHid;
    L.b2 = L.w2 + kHid * kHid;
    L.w3 = L.b2 + kHid;
    L.b3 = L.w3 + 2 * kHid;
    L.n = L.b3 + 2;
    return L;
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

/* the BF16 inference parameters staged in LDS when they fit (dim <= 156) */
constexpr int kLdsParams = 16384;

/* all threads of the block copy n BF16 parameters (n padded to 8) to LDS */
__device__ __forceinline__ void stage_params(const uint16_t *__restrict__ wt, int n,
                                             uint16_t *ws) {
    const uint4 *src = (const uint4 *)wt;
    uint4 *dst = (uint4 *)ws;
    for (int i = threadIdx.x; i < (n + 7) / 8; i += blockDim.x) dst[i] = src[i];
}

/*
 * MLP forward for one query by one wave (lane j = hidden unit j) with the
 * inference copy of the parameters (BF16; W1 and W2 transposed, W1t[F][64]
 * and W2t[64][64], so step i reads 64 consecutive values). xs holds the raw
 * features (standardised here into xh), hs is 64 floats of LDS.
 */
__device__ inline void wave_mlp(const float *xs, int F, const uint16_t *wt,
                         const float *__restrict__ mean, const float *__restrict__ istd,
                         float *xh, float *hs, float &y0, float &y1) {
    const int l = threadIdx.x & 63;
    const MlpLayout L = mlp_layout(F);
    for (int j = l; j < F; j += 64) xh[j] = (xs[j] - mean[j]) * istd[j];
    wave_lds_sync();
    float a = bf2f(wt[L.b1 + l]);
#pragma unroll 8
    for (int i = 0; i < F; i++) a = fmaf(bf2f(wt[L.w1 + i * kHid + l]), xh[i], a);
    hs[l] = fmaxf(a, 0.f);
    wave_lds_sync();
    float b = bf2f(wt[L.b2 + l]);
#pragma unroll 8
    for (int i = 0; i < kHid; i++) b = fmaf(bf2f(wt[L.w2 + i * kHid + l]), hs[i], b);
    b = fmaxf(b, 0.f);
    y0 = wave_sum(bf2f(wt[L.w3 + l]) * b) + bf2f(wt[L.b3 + 0]);
    y1 = wave_sum(bf2f(wt[L.w3 + kHid + l]) * b) + bf2f(wt[L.b3 + 1]);
    wave_lds_sync();
}

/* arguments of a routed run's decision: features, model, policy, and the
 * per-query route / ef outputs */
struct LearnArgs {
    const float *q;
    int nq, dim, F;
    const int32_t *pofq;
    const uint32_t *progpart;  /* [nprog][nbm] passing rows per bitmap block */
    uint32_t nbm;
    uint64_t count;
    const float *cent;
    int ncent;
    const uint16_t *wt;
    const float *mean, *istd;
    int mode;
    uint32_t ef_fixed, ef_min, ef_max;
    float scale, bias;
    uint8_t *route;
    uint32_t *efq;
    float *feat_out, *y_out;
};

/* passing rows of query qi's predicate (all rows without one): the wave
 * sums its program's per-block counts of the bitmap pass */
__device__ __forceinline__ double wave_pass_count(const LearnArgs &A, int qi) {
    const int p = A.pofq[qi];
    if (p < 0) return (double)A.count;
    const uint32_t *part = A.progpart + (uint64_t)p * A.nbm;
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x & 63; i < A.nbm; i += 64) c += part[i];
    for (int o = 32; o; o >>= 1) c += __shfl_xor(c, o);
    return (double)c;
}

/* route/ef from the model outputs (mode 0/1) or the fixed policy */
__device__ __forceinline__ void route_decision(const LearnArgs &A, int qi, float y0, float y1) {
    uint32_t ef = A.ef_fixed;
    uint8_t r = 1;
    if (A.mode == 3) {
        r = 0;
    } else if (A.mode == 0 || A.mode == 1) {
        const float e = expf(fminf(y0, 20.f)) * A.scale;
        ef = (uint32_t)fminf(ceilf(e), (float)A.ef_max);
        r = (A.mode == 1 || y1 > A.bias) ? 1 : 0;
    }
    A.efq[qi] = min(max(ef, A.ef_min), A.ef_max);
    A.route[qi] = r;
}

/*
 * Decision kernel when the query prep of phase 1 already wrote each query's
 * features except its selectivity into feat_out (dim <= 256): one wave per
 * query, four per block — the selectivity from the bitmap pass's counts,
 * then the MLP and route/ef.
 */
__device__ __forceinline__ void route_block(const LearnArgs &A, int blk) {
    __shared__ __attribute__((aligned(16))) uint16_t ws[kLdsParams];
    __shared__ __attribute__((aligned(16))) float xs[4][kMaxFeat];
    __shared__ float xh[4][kMaxFeat];
    __shared__ float hs[4][kHid];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int qi = blk * 4 + w;
    const int np = mlp_layout(A.F).n;
    const bool lds = np <= kLdsParams;
    if (lds) stage_params(A.wt, np, ws);
    const bool live = qi < A.nq;
    const double pass = live ? wave_pass_count(A, qi) : 0.0;
    float *fq = A.feat_out + (uint64_t)(live ? qi : 0) * A.F;
    if (live)
        for (int j = l; j < A.F; j += 64) xs[w][j] = fq[j];
    __syncthreads(); /* every wave reaches it (no early return above) */
    if (!live) return;
    if (l == 0) {
        const float sel = (float)log10(fmax(pass, 1.0) / (double)A.count);
        xs[w][A.dim] = sel;
        fq[A.dim] = sel;
    }
    wave_lds_sync();
    float y0, y1;
    wave_mlp(xs[w], A.F, lds ? ws : A.wt, A.mean, A.istd, xh[w], hs[w], y0, y1);
    if (l == 0) {
        if (A.y_out) {
            A.y_out[(uint64_t)qi * 2 + 0] = y0;
            A.y_out[(uint64_t)qi * 2 + 1] = y1;
        }
        route_decision(A, qi, y0, y1);
    }
}


} // namespace pv
