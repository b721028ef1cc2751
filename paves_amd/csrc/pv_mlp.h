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
#pragma unroll
    for (int o = 32; o; o >>= 1) v += lane_xor(v, o);
    return v;
}

/* the BF16 inference parameters staged in LDS when they fit (dim <= 156) */
constexpr int kLdsParams = 16384;

/* the same staging split in two for a 256-thread block and n <= kLdsParams
 * (at most kStagePer 16-byte pieces per thread): loads into registers,
 * then the LDS stores, so other loads can be issued in between */
constexpr int kStagePer = 8;
static_assert(kLdsParams <= kStagePer * 8 * 256, "one pass of a 256-thread block");
__device__ __forceinline__ void stage_load(const uint16_t *__restrict__ wt, int n,
                                           uint4 (&v)[kStagePer]) {
    const uint4 *src = (const uint4 *)wt;
    const int n16 = (n + 7) / 8;
#pragma unroll
    for (int k = 0; k < kStagePer; k++) {
        const int i = k * 256 + threadIdx.x;
        if (i < n16) v[k] = src[i];
    }
}
__device__ __forceinline__ void stage_store(int n, const uint4 (&v)[kStagePer], uint16_t *ws) {
    uint4 *dst = (uint4 *)ws;
    const int n16 = (n + 7) / 8;
#pragma unroll
    for (int k = 0; k < kStagePer; k++) {
        const int i = k * 256 + threadIdx.x;
        if (i < n16) dst[i] = v[k];
    }
}

/* all threads of the block copy n BF16 parameters (n padded to 8) to LDS;
 * a thread's loads are all issued before its LDS writes (one memory round
 * trip for a 256-thread block) */
__device__ __forceinline__ void stage_params(const uint16_t *__restrict__ wt, int n,
                                             uint16_t *ws) {
    constexpr int kPer = 8;
    const uint4 *src = (const uint4 *)wt;
    uint4 *dst = (uint4 *)ws;
    const int n16 = (n + 7) / 8;
    for (int base = 0; base < n16; base += kPer * blockDim.x) {
        uint4 v[kPer];
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int i = base + k * blockDim.x + threadIdx.x;
            if (i < n16) v[k] = src[i];
        }
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int i = base + k * blockDim.x + threadIdx.x;
            if (i < n16) dst[i] = v[k];
        }
    }
}

/*
 * MLP forward for one query by one wave (lane j = hidden unit j) with the
 * inference copy of the parameters (BF16; W1 and W2 transposed, W1t[F][64]
 * and W2t[64][64], so step i reads 64 consecutive values). xs holds the raw
 * features (standardised here into xh), hs is 64 floats of LDS.
 */
/* the MLP on standardised inputs xh (LDS) */
__device__ inline void wave_mlp_std(const float *xh, int F, const uint16_t *wt, float *hs,
                                    float &y0, float &y1) {
    const int l = threadIdx.x & 63;
    const MlpLayout L = mlp_layout(F);
    /* four partial sums over i mod 4 (independent FMA chains), summed as
     * ((s0 + s1) + (s2 + s3)) + bias: the one order of every inference path */
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int i = 0;
#pragma unroll 4
    for (; i + 4 <= F; i += 4) {
        s0 = fmaf(bf2f(wt[L.w1 + (i + 0) * kHid + l]), xh[i + 0], s0);
        s1 = fmaf(bf2f(wt[L.w1 + (i + 1) * kHid + l]), xh[i + 1], s1);
        s2 = fmaf(bf2f(wt[L.w1 + (i + 2) * kHid + l]), xh[i + 2], s2);
        s3 = fmaf(bf2f(wt[L.w1 + (i + 3) * kHid + l]), xh[i + 3], s3);
    }
    for (; i < F; i++) s0 = fmaf(bf2f(wt[L.w1 + i * kHid + l]), xh[i], s0);
    const float a = ((s0 + s1) + (s2 + s3)) + bf2f(wt[L.b1 + l]);
    hs[l] = fmaxf(a, 0.f);
    wave_lds_sync();
    float t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
#pragma unroll 4
    for (int k = 0; k < kHid; k += 4) {
        t0 = fmaf(bf2f(wt[L.w2 + (k + 0) * kHid + l]), hs[k + 0], t0);
        t1 = fmaf(bf2f(wt[L.w2 + (k + 1) * kHid + l]), hs[k + 1], t1);
        t2 = fmaf(bf2f(wt[L.w2 + (k + 2) * kHid + l]), hs[k + 2], t2);
        t3 = fmaf(bf2f(wt[L.w2 + (k + 3) * kHid + l]), hs[k + 3], t3);
    }
    float b = ((t0 + t1) + (t2 + t3)) + bf2f(wt[L.b2 + l]);
    b = fmaxf(b, 0.f);
    y0 = wave_sum(bf2f(wt[L.w3 + l]) * b) + bf2f(wt[L.b3 + 0]);
    y1 = wave_sum(bf2f(wt[L.w3 + kHid + l]) * b) + bf2f(wt[L.b3 + 1]);
    wave_lds_sync();
}

__device__ inline void wave_mlp(const float *xs, int F, const uint16_t *wt,
                                const float *__restrict__ mean, const float *__restrict__ istd,
                                float *xh, float *hs, float &y0, float &y1) {
    const int l = threadIdx.x & 63;
    for (int j = l; j < F; j += 64) xh[j] = (xs[j] - mean[j]) * istd[j];
    wave_lds_sync();
    wave_mlp_std(xh, F, wt, hs, y0, y1);
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

/* passing rows of a query whose program index is p (all rows without a
 * predicate, p < 0): the wave sums the program's per-block counts of the
 * bitmap pass. The caller reads p (A.pofq) early, so that its round trip
 * overlaps its other loads. */
__device__ __forceinline__ double wave_pass_count(const LearnArgs &A, int p) {
    if (p < 0) return (double)A.count;
    const uint32_t *part = A.progpart + (uint64_t)p * A.nbm;
    const uint32_t l = threadIdx.x & 63;
    uint32_t c = 0;
    /* 16 loads in flight per lane per step (one round trip per 1M rows) */
    for (uint32_t base = 0; base < A.nbm; base += 64 * 16) {
        uint32_t v[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t i = base + 64 * k + l;
            v[k] = i < A.nbm ? part[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < 16; k++) c += v[k];
    }
#pragma unroll
    for (int o = 32; o; o >>= 1) c += lane_xor(c, o);
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
    constexpr int kFeat = 256 + 1 + kNear; /* F <= dim + 17 with dim <= 256 */
    constexpr int kPerLane = (kFeat + 63) / 64;
    __shared__ __attribute__((aligned(16))) uint16_t ws[kLdsParams];
    __shared__ float xh[4][kFeat];
    __shared__ float hs[4][kHid];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int qi = blk * 4 + w;
    const int np = mlp_layout(A.F).n;
    const bool lds = np <= kLdsParams;
    /* every global read of the block is issued before the first wait:
     * features and standardisation into registers, the parameters into
     * LDS, the per-block counts (after the query's program index) */
    const bool live = qi < A.nq;
    const int prog = live ? A.pofq[qi] : -1; /* the pass count's loads depend on it */
    float *fq = A.feat_out + (uint64_t)(live ? qi : 0) * A.F;
    float fv[kPerLane], mu[kPerLane], is[kPerLane];
#pragma unroll
    for (int k = 0; k < kPerLane; k++) {
        const int j = l + 64 * k;
        const bool in = live && j < A.F;
        fv[k] = in ? fq[j] : 0.f;
        mu[k] = in ? A.mean[j] : 0.f;
        is[k] = in ? A.istd[j] : 0.f;
    }
    /* the parameters go to LDS after the pass count's loads are issued:
     * one round trip for the program index, one for everything else */
    uint4 sv[kStagePer];
    if (lds) stage_load(A.wt, np, sv);
    const double pass = live ? wave_pass_count(A, prog) : 0.0;
    if (lds) stage_store(np, sv, ws);
    __syncthreads(); /* every wave reaches it (no early return above) */
    if (!live) return;
    const float sel = (float)log10(fmax(pass, 1.0) / (double)A.count);
#pragma unroll
    for (int k = 0; k < kPerLane; k++) {
        const int j = l + 64 * k;
        if (j < A.F) xh[w][j] = ((j == A.dim ? sel : fv[k]) - mu[k]) * is[k];
    }
    if (l == 0) fq[A.dim] = sel;
    wave_lds_sync();
    float y0, y1;
    wave_mlp_std(xh[w], A.F, lds ? ws : A.wt, hs[w], y0, y1);
    if (l == 0) {
        if (A.y_out) {
            A.y_out[(uint64_t)qi * 2 + 0] = y0;
            A.y_out[(uint64_t)qi * 2 + 1] = y1;
        }
        route_decision(A, qi, y0, y1);
    }
}


} // namespace pv
