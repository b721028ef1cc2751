/*
 * pv_learn.hip — the learned per-query beam width and brute-vs-graph router.
 *
 * The reference picks ef from a fixed policy (ef = max(k, ef_search x scale)
 * with scale in {1, 0.8, 0.4} by selectivity, exec.c:422-424) and routes by
 * a cost race (planner.c:185-346). Here a small MLP predicts, per query,
 * log(ef) and a brute-vs-graph logit from
 *   x = [ q (dim) | log10 selectivity | distances to the 16 nearest centroids ]
 * (dim + 17 inputs; 145 at dim 128), two hidden layers of 64 (ReLU), two
 * outputs. Inputs are standardised with the training set's mean and std.
 *
 * Kernels:
 *   k_prog_count   passing rows of each predicate program (bitmap popcount)
 *   k_features     one wave per query: the feature vector above
 *   k_mlp_infer    one wave per query, lane j = hidden unit j, BF16 weights
 *                  with FP32 accumulation
 *   k_train_fwd    one wave per training sample: forward pass, loss and the
 *                  back-propagated deltas of each layer (FP32)
 *   k_train_wgrad  one thread per parameter: sum over the minibatch of
 *                  delta x activation (fixed order, deterministic)
 *   k_adam         Adam update
 * Parameter layout (FP32, nparams = 64(F+1) + 64*65 + 2*65):
 *   W1[64][F] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
 * Loss per sample: (y0 - log ef)^2 * valid + lambda * BCE(y1, route), mean
 * over the minibatch.
 */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "pv_mlp.h"

namespace pv {


/* ---------------- features ---------------- */

/* one 256-thread block per program: popcount of its bitmap row */
__global__ __launch_bounds__(256) void k_prog_count(const uint32_t *__restrict__ bitmap,
                                                    uint64_t words, uint32_t *__restrict__ cnt) {
    __shared__ uint32_t part[4];
    const uint32_t *row = bitmap + (uint64_t)blockIdx.x * words;
    uint32_t c = 0;
    for (uint64_t w = threadIdx.x; w < words; w += blockDim.x) c += __popc(row[w]);
#pragma unroll
    for (int o = 32; o; o >>= 1) c += lane_xor(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

/*
 * Feature vector of one query, computed by one wave into xs[0..F):
 *   [q | log10(max(pass, 1) / count) | sqrt of the 16 smallest squared
 *    distances to the centroids, ascending].
 * Lane l owns centroids l, l+64, ...: each row is read with independent
 * 16-byte loads (many in flight per lane), the query from LDS by
 * broadcast; each lane keeps its 16 best, then 16 rounds of wave minimum
 * merge the lanes' lists. Wave-local (no block barrier).
 */
/*
 * Features of one query by one wave into xs[0..F): the query, log10 of its
 * predicate's selectivity, and the distances to its kNear nearest
 * centroids, ascending (centroids transposed, [dim][ncent]).
 */
__device__ void wave_features(const float *__restrict__ qv, int dim, double pass, uint64_t count,
                              const float *__restrict__ cent_t, int ncent, float *xs) {
    const int l = threadIdx.x & 63;
    for (int j = l; j < dim; j += 64) xs[j] = qv[j];
    if (l == 0) xs[dim] = (float)log10(fmax(pass, 1.0) / (double)count);
    wave_lds_sync();
    float best[kNear];
#pragma unroll
    for (int k = 0; k < kNear; k++) best[k] = __uint_as_float(0x7f800000u);
    for (int c0 = 0; c0 < ncent; c0 += 64) {
        const int c = min(c0 + l, ncent - 1);
        const float v = cent_dist(xs, cent_t, c, dim, ncent);
        near_insert(best, c0 + l < ncent ? v : __uint_as_float(0x7f800000u));
    }
    __shared__ float wn[4][kNear];
    wave_nearest(best, wn[(threadIdx.x >> 6) & 3]);
    wave_lds_sync();
    if (l < kNear) xs[dim + 1 + l] = sqrtf(wn[(threadIdx.x >> 6) & 3][l]);
    wave_lds_sync();
}

/* one wave per query (4 per block): feature vectors to out [nq][F] */
__global__ __launch_bounds__(256) void k_features(const float *__restrict__ q, int nq, int dim,
                                                  const int32_t *__restrict__ pofq,
                                                  const uint32_t *__restrict__ pcnt,
                                                  uint64_t count, const float *__restrict__ cent,
                                                  int ncent, float *__restrict__ out) {
    __shared__ __attribute__((aligned(16))) float xs[4][kMaxFeat];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int qi = blockIdx.x * 4 + w;
    if (qi >= nq) return; /* no block barrier below */
    const int F = model_nfeat(dim);
    const int p = pofq[qi];
    wave_features(q + (uint64_t)qi * dim, dim, p < 0 ? (double)count : (double)pcnt[p], count,
                  cent, ncent, xs[w]);
    for (int j = l; j < F; j += 64) out[(uint64_t)qi * F + j] = xs[w][j];
}

/* ---------------- inference ---------------- */

/* one wave per query (4 per block): model outputs y [n][2] */
__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ x, int n, int F,
                                                   const uint16_t *__restrict__ wt,
                                                   const float *__restrict__ mean,
                                                   const float *__restrict__ istd,
                                                   float *__restrict__ y) {
    __shared__ __attribute__((aligned(16))) uint16_t ws[kLdsParams];
    __shared__ float xs[4][kMaxFeat], xh[4][kMaxFeat];
    __shared__ float hs[4][kHid];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int s = blockIdx.x * 4 + w;
    const int np = mlp_layout(F).n;
    const bool lds = np <= kLdsParams;
    if (lds) stage_params(wt, np, ws);
    __syncthreads(); /* every wave reaches it (no early return above) */
    if (s >= n) return;
    for (int j = l; j < F; j += 64) xs[w][j] = x[(uint64_t)s * F + j];
    wave_lds_sync();
    float y0, y1;
    wave_mlp(xs[w], F, lds ? ws : wt, mean, istd, xh[w], hs[w], y0, y1);
    if (l == 0) {
        y[(uint64_t)s * 2 + 0] = y0;
        y[(uint64_t)s * 2 + 1] = y1;
    }
}

/* ---------------- training ---------------- */

struct TrainArgs {
    const float *x;       /* [n][F] raw features */
    const float *t;       /* [n][3]: log ef, route (0/1), ef valid (0/1) */
    const int32_t *idx;   /* minibatch sample ids [B] */
    int B, F;
    const float *w;       /* FP32 params */
    const float *mean, *istd;
    float lambda;         /* weight of the routing loss */
    float *xh;            /* [B][F] standardised inputs */
    float *a1, *a2;       /* [B][64] post-ReLU activations */
    float *d1, *d2;       /* [B][64] deltas at the pre-activations */
    float *d3;            /* [B][2] output deltas */
    float *loss;          /* [B] per-sample loss */
};

/* one wave per sample of the minibatch */
__global__ __launch_bounds__(256) void k_train_fwd(TrainArgs A) {
    __shared__ float xs[4][kMaxFeat];
    __shared__ float h1s[4][kHid], d2s[4][kHid];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + w;
    const bool act = b < A.B; /* every wave reaches every barrier */
    const int F = A.F;
    const MlpLayout L = mlp_layout(F);
    const int s = act ? A.idx[b] : 0;
    for (int j = l; j < F; j += 64) {
        const float v = act ? (A.x[(uint64_t)s * F + j] - A.mean[j]) * A.istd[j] : 0.f;
        xs[w][j] = v;
        if (act) A.xh[(uint64_t)b * F + j] = v;
    }
    __syncthreads();
    float z1 = A.w[L.b1 + l];
    for (int i = 0; i < F; i++) z1 = fmaf(A.w[L.w1 + l * F + i], xs[w][i], z1);
    const float h1 = fmaxf(z1, 0.f);
    h1s[w][l] = h1;
    __syncthreads();
    float z2 = A.w[L.b2 + l];
    for (int i = 0; i < kHid; i++) z2 = fmaf(A.w[L.w2 + l * kHid + i], h1s[w][i], z2);
    const float h2 = fmaxf(z2, 0.f);
    const float y0 = wave_sum(A.w[L.w3 + l] * h2) + A.w[L.b3 + 0];
    const float y1 = wave_sum(A.w[L.w3 + kHid + l] * h2) + A.w[L.b3 + 1];
    const float t0 = act ? A.t[(uint64_t)s * 3 + 0] : 0.f;
    const float t1 = act ? A.t[(uint64_t)s * 3 + 1] : 0.f;
    const float valid = act ? A.t[(uint64_t)s * 3 + 2] : 0.f;
    /* d loss / d y: squared error on log ef, logistic loss on the route */
    const float e0 = y0 - t0;
    const float sig = 1.f / (1.f + expf(-y1));
    const float g0 = 2.f * e0 * valid;
    const float g1 = A.lambda * (sig - t1);
    const float d2 = (A.w[L.w3 + l] * g0 + A.w[L.w3 + kHid + l] * g1) * (z2 > 0.f ? 1.f : 0.f);
    d2s[w][l] = d2;
    __syncthreads();
    float bd = 0.f;
    for (int i = 0; i < kHid; i++) bd = fmaf(A.w[L.w2 + i * kHid + l], d2s[w][i], bd);
    const float d1 = bd * (z1 > 0.f ? 1.f : 0.f);
    if (!act) return;
    A.a1[(uint64_t)b * kHid + l] = h1;
    A.a2[(uint64_t)b * kHid + l] = h2;
    A.d1[(uint64_t)b * kHid + l] = d1;
    A.d2[(uint64_t)b * kHid + l] = d2;
    if (l == 0) {
        A.d3[(uint64_t)b * 2 + 0] = g0;
        A.d3[(uint64_t)b * 2 + 1] = g1;
        /* softplus(y1) - t1 y1, the logistic loss in a stable form */
        const float sp = fmaxf(y1, 0.f) + log1pf(expf(-fabsf(y1)));
        A.loss[b] = e0 * e0 * valid + A.lambda * (sp - t1 * y1);
    }
}

/* one thread per parameter: mean over the minibatch of delta x input */
__global__ __launch_bounds__(256) void k_train_wgrad(TrainArgs A, float *__restrict__ grad) {
    const int F = A.F;
    const MlpLayout L = mlp_layout(F);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L.n) return;
    const int B = A.B;
    float s = 0.f;
    if (p < L.b1) {
        const int i = p / F, j = p % F;
        for (int b = 0; b < B; b++) s = fmaf(A.d1[(uint64_t)b * kHid + i], A.xh[(uint64_t)b * F + j], s);
    } else if (p < L.w2) {
        const int i = p - L.b1;
        for (int b = 0; b < B; b++) s += A.d1[(uint64_t)b * kHid + i];
    } else if (p < L.b2) {
        const int i = (p - L.w2) / kHid, j = (p - L.w2) % kHid;
        for (int b = 0; b < B; b++) s = fmaf(A.d2[(uint64_t)b * kHid + i], A.a1[(uint64_t)b * kHid + j], s);
    } else if (p < L.w3) {
        const int i = p - L.b2;
        for (int b = 0; b < B; b++) s += A.d2[(uint64_t)b * kHid + i];
    } else if (p < L.b3) {
        const int k = (p - L.w3) / kHid, j = (p - L.w3) % kHid;
        for (int b = 0; b < B; b++) s = fmaf(A.d3[(uint64_t)b * 2 + k], A.a2[(uint64_t)b * kHid + j], s);
    } else {
        const int k = p - L.b3;
        for (int b = 0; b < B; b++) s += A.d3[(uint64_t)b * 2 + k];
    }
    grad[p] = s / (float)B;
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float gg = g[p];
    const float mm = b1 * m[p] + (1.f - b1) * gg;
    const float vv = b2 * v[p] + (1.f - b2) * gg * gg;
    m[p] = mm;
    v[p] = vv;
    w[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

/*
 * Per-query decision from the model outputs y = (log ef, route logit):
 * ef = clamp(ceil(exp(y0) * scale), ef_min, ef_max); route 1 (graph) when
 * y1 > bias. mode 1 forces the graph with the learned ef, mode 2 the graph
 * with a fixed ef, mode 3 the brute path.
 */
__global__ void k_route(const float *__restrict__ y, int nq, int mode, uint32_t ef_fixed,
                        uint32_t ef_min, uint32_t ef_max, float scale, float bias,
                        uint8_t *__restrict__ route, uint32_t *__restrict__ efq) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t ef = ef_fixed;
    uint8_t r = 1;
    if (mode == 3) {
        r = 0;
    } else if (mode == 0 || mode == 1) {
        const float e = expf(fminf(y[2 * q], 20.f)) * scale;
        ef = (uint32_t)fminf(ceilf(e), (float)ef_max);
        r = (mode == 1 || y[2 * q + 1] > bias) ? 1 : 0;
    }
    efq[q] = min(max(ef, ef_min), ef_max);
    route[q] = r;
}

/* FP32 -> BF16 (round to nearest even) inference copy of the parameters,
 * W1 and W2 transposed (W1t[F][64], W2t[64][64]); other blocks in place */
__global__ void k_to_bf16(const float *__restrict__ w, uint16_t *__restrict__ o, int F) {
    const MlpLayout L = mlp_layout(F);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L.n) return;
    int d = p;
    if (p < L.b1) {
        const int i = p / F, j = p % F;
        d = L.w1 + j * kHid + i;
    } else if (p >= L.w2 && p < L.b2) {
        const int i = (p - L.w2) / kHid, j = (p - L.w2) % kHid;
        d = L.w2 + j * kHid + i;
    }
    o[d] = f2bf(w[p]);
}

/*
 * The routed run's decision kernel, one 4-wave block per query: the waves
 * split the centroid distances (each keeps its kNear nearest), wave 0 merges
 * the 4 x kNear candidates with a register sort, then runs the MLP and sets
 * route/ef as in k_route. Features equal those of k_features (same
 * distances, same nearest set). Optionally stores the features and outputs.
 */
__global__ __launch_bounds__(256) void k_learn_route(LearnArgs A) {
    __shared__ __attribute__((aligned(16))) uint16_t ws[kLdsParams];
    __shared__ __attribute__((aligned(16))) float xs[kMaxFeat];
    __shared__ float xh[kMaxFeat];
    __shared__ float hs[kHid];
    __shared__ float wn[4][kNear];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int qi = blockIdx.x;
    const int np = mlp_layout(A.F).n;
    const bool lds = np <= kLdsParams;
    if (lds) stage_params(A.wt, np, ws);
    const float *qv = A.q + (uint64_t)qi * A.dim;
    for (int j = threadIdx.x; j < A.dim; j += blockDim.x) xs[j] = qv[j];
    if (w == 0) {
        const double pass = wave_pass_count(A, A.pofq[qi]);
        if (l == 0) xs[A.dim] = (float)log10(fmax(pass, 1.0) / (double)A.count);
    }
    __syncthreads();
    float best[kNear];
#pragma unroll
    for (int k = 0; k < kNear; k++) best[k] = __uint_as_float(0x7f800000u);
    for (int c0 = 64 * w; c0 < A.ncent; c0 += 256) {
        const int c = min(c0 + l, A.ncent - 1);
        const float v = cent_dist(xs, A.cent, c, A.dim, A.ncent);
        near_insert(best, c0 + l < A.ncent ? v : __uint_as_float(0x7f800000u));
    }
    wave_nearest(best, wn[w]);
    __syncthreads();
    if (w != 0) return; /* no block barrier below */
    {
        /* 4 x kNear candidates -> the kNear smallest, ascending (distances
         * are non-negative, so float bits order as the values) */
        const float cv = wn[l >> 4][l & 15];
        const uint64_t sk = wave_sort64((uint64_t)__float_as_uint(cv));
        if (l < kNear) xs[A.dim + 1 + l] = sqrtf(__uint_as_float((uint32_t)sk));
        wave_lds_sync();
    }
    float y0, y1;
    wave_mlp(xs, A.F, lds ? ws : A.wt, A.mean, A.istd, xh, hs, y0, y1);
    if (A.feat_out)
        for (int j = l; j < A.F; j += 64) A.feat_out[(uint64_t)qi * A.F + j] = xs[j];
    if (l == 0) {
        if (A.y_out) {
            A.y_out[(uint64_t)qi * 2 + 0] = y0;
            A.y_out[(uint64_t)qi * 2 + 1] = y1;
        }
        route_decision(A, qi, y0, y1);
    }
}

} // namespace pv

using namespace pv;

struct FvGpuModel {
    int device = -1;
    int dim = 0, F = 0, ncent = 0, np = 0;
    hipStream_t stream = nullptr;
    float *d_cent = nullptr;
    float *d_w = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr;
    uint16_t *d_wbf = nullptr;
    float *d_mean = nullptr, *d_istd = nullptr;
    std::vector<float> h_mean, h_istd;
    int adam_t = 0;
    /* training scratch */
    DevBuf d_x, d_t, d_idx, d_xh, d_a1, d_a2, d_d1, d_d2, d_d3, d_loss, d_y;
};

namespace {

int model_sync_bf16(FvGpuModel *m, char *errbuf, size_t errlen) {
    hipLaunchKernelGGL(k_to_bf16, dim3((m->np + 255) / 256), dim3(256), 0, m->stream, m->d_w,
                       m->d_wbf, m->F);
    PV_TRY(hipGetLastError());
    return 0;
}

int ensure_train(FvGpuModel *m, int n, int B, char *errbuf, size_t errlen) {
    PV_TRY(m->d_x.ensure((size_t)n * m->F * 4));
    PV_TRY(m->d_t.ensure((size_t)n * 3 * 4));
    PV_TRY(m->d_idx.ensure((size_t)B * 4));
    PV_TRY(m->d_xh.ensure((size_t)B * m->F * 4));
    PV_TRY(m->d_a1.ensure((size_t)B * kHid * 4));
    PV_TRY(m->d_a2.ensure((size_t)B * kHid * 4));
    PV_TRY(m->d_d1.ensure((size_t)B * kHid * 4));
    PV_TRY(m->d_d2.ensure((size_t)B * kHid * 4));
    PV_TRY(m->d_d3.ensure((size_t)B * 2 * 4));
    PV_TRY(m->d_loss.ensure((size_t)B * 4));
    return 0;
}

TrainArgs train_args(FvGpuModel *m, int B, float lambda) {
    TrainArgs A;
    A.x = m->d_x.as<float>();
    A.t = m->d_t.as<float>();
    A.idx = m->d_idx.as<int32_t>();
    A.B = B;
    A.F = m->F;
    A.w = m->d_w;
    A.mean = m->d_mean;
    A.istd = m->d_istd;
    A.lambda = lambda;
    A.xh = m->d_xh.as<float>();
    A.a1 = m->d_a1.as<float>();
    A.a2 = m->d_a2.as<float>();
    A.d1 = m->d_d1.as<float>();
    A.d2 = m->d_d2.as<float>();
    A.d3 = m->d_d3.as<float>();
    A.loss = m->d_loss.as<float>();
    return A;
}

/* forward + backward over samples idx[0..B) into m->d_g; mean loss out */
int grad_step(FvGpuModel *m, int B, float lambda, double *loss, char *errbuf, size_t errlen) {
    TrainArgs A = train_args(m, B, lambda);
    hipLaunchKernelGGL(k_train_fwd, dim3((B + 3) / 4), dim3(256), 0, m->stream, A);
    hipLaunchKernelGGL(k_train_wgrad, dim3((m->np + 255) / 256), dim3(256), 0, m->stream, A,
                       m->d_g);
    PV_TRY(hipGetLastError());
    if (loss) {
        std::vector<float> l(B);
        PV_TRY(hipMemcpyAsync(l.data(), A.loss, (size_t)B * 4, hipMemcpyDeviceToHost, m->stream));
        PV_TRY(hipStreamSynchronize(m->stream));
        double s = 0.0;
        for (float v : l) s += v;
        *loss = s / B;
    }
    return 0;
}

} // namespace

extern "C" int fv_gpu_model_nparams(int dim) { return dim > 0 ? model_nparams(model_nfeat(dim)) : -1; }

extern "C" FvGpuModel *fv_gpu_model_create(int device, int dim, const float *centroids,
                                           int ncent, char *errbuf, size_t errlen) {
    if (dim < 1 || dim > 1024 || centroids == nullptr || ncent < kNear || ncent > 65536) {
        snprintf(errbuf, errlen, "gpu: model: bad arguments (dim=%d ncent=%d)", dim, ncent);
        return nullptr;
    }
    FvGpuModel *m = new FvGpuModel();
    m->device = device < 0 ? 0 : device;
    m->dim = dim;
    m->F = model_nfeat(dim);
    m->ncent = ncent;
    m->np = model_nparams(m->F);
    auto fail = [&](hipError_t e, const char *what) {
        snprintf(errbuf, errlen, "gpu: model: %s: %s", what, hipGetErrorString(e));
        fv_gpu_model_free(m);
        return (FvGpuModel *)nullptr;
    };
    hipError_t e;
    if ((e = hipSetDevice(m->device)) != hipSuccess) return fail(e, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "stream");
    const size_t np = (size_t)m->np;
    if ((e = hipMalloc(&m->d_cent, (size_t)ncent * dim * 4)) != hipSuccess) return fail(e, "alloc");
    if ((e = hipMalloc(&m->d_w, np * 4)) != hipSuccess) return fail(e, "alloc");
    if ((e = hipMalloc(&m->d_m, np * 4)) != hipSuccess) return fail(e, "alloc");
    if ((e = hipMalloc(&m->d_v, np * 4)) != hipSuccess) return fail(e, "alloc");
    if ((e = hipMalloc(&m->d_g, np * 4)) != hipSuccess) return fail(e, "alloc");
    if ((e = hipMalloc(&m->d_wbf, (np + 7) / 8 * 16)) != hipSuccess) return fail(e, "alloc");
    if ((e = hipMemset(m->d_wbf, 0, (np + 7) / 8 * 16)) != hipSuccess) return fail(e, "alloc");
    if ((e = hipMalloc(&m->d_mean, (size_t)m->F * 4)) != hipSuccess) return fail(e, "alloc");
    if ((e = hipMalloc(&m->d_istd, (size_t)m->F * 4)) != hipSuccess) return fail(e, "alloc");
    {
        /* centroids stored transposed, [dim][ncent]: a wave reads coordinate
         * k of 64 consecutive centroids with one coalesced load */
        std::vector<float> ct((size_t)ncent * dim);
        for (int c = 0; c < ncent; c++)
            for (int k = 0; k < dim; k++) ct[(size_t)k * ncent + c] = centroids[(size_t)c * dim + k];
        if ((e = hipMemcpy(m->d_cent, ct.data(), ct.size() * 4, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "copy");
    }
    /* He-initialised weights from a fixed seed; zero biases; identity scaling */
    std::vector<float> w(np, 0.f);
    const MlpLayout L = mlp_layout(m->F);
    std::mt19937_64 rng(0x5eed5eedULL);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int i = 0; i < kHid * m->F; i++) w[L.w1 + i] = (float)(nd(rng) * sqrt(2.0 / m->F));
    for (int i = 0; i < kHid * kHid; i++) w[L.w2 + i] = (float)(nd(rng) * sqrt(2.0 / kHid));
    for (int i = 0; i < 2 * kHid; i++) w[L.w3 + i] = (float)(nd(rng) * sqrt(1.0 / kHid));
    m->h_mean.assign(m->F, 0.f);
    m->h_istd.assign(m->F, 1.f);
    char eb2[256];
    if (fv_gpu_model_set_params(m, w.data(), m->h_mean.data(), m->h_istd.data(), eb2,
                                sizeof(eb2)) != 0) {
        snprintf(errbuf, errlen, "%s", eb2);
        fv_gpu_model_free(m);
        return nullptr;
    }
    return m;
}

extern "C" void fv_gpu_model_free(FvGpuModel *m) {
    if (!m) return;
    if (m->device >= 0) hipSetDevice(m->device);
    if (m->stream) hipStreamSynchronize(m->stream);
    hipFree(m->d_cent);
    hipFree(m->d_w);
    hipFree(m->d_m);
    hipFree(m->d_v);
    hipFree(m->d_g);
    hipFree(m->d_wbf);
    hipFree(m->d_mean);
    hipFree(m->d_istd);
    for (DevBuf *b : {&m->d_x, &m->d_t, &m->d_idx, &m->d_xh, &m->d_a1, &m->d_a2, &m->d_d1,
                      &m->d_d2, &m->d_d3, &m->d_loss, &m->d_y})
        b->release();
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
}

/* params: FP32 [nparams]; mean/std: [F] input standardisation (std 0 -> 1) */
extern "C" int fv_gpu_model_set_params(FvGpuModel *m, const float *params, const float *mean,
                                       const float *std_, char *errbuf, size_t errlen) {
    if (m == nullptr || params == nullptr) PV_FAIL("gpu: model: NULL");
    PV_TRY(hipSetDevice(m->device));
    std::vector<float> is(m->F, 1.f), mu(m->F, 0.f);
    for (int j = 0; j < m->F; j++) {
        if (mean) mu[j] = mean[j];
        if (std_) is[j] = std_[j] > 0.f ? 1.f / std_[j] : 1.f;
    }
    m->h_mean = mu;
    m->h_istd = is;
    PV_TRY(hipMemcpyAsync(m->d_w, params, (size_t)m->np * 4, hipMemcpyHostToDevice, m->stream));
    PV_TRY(hipMemcpyAsync(m->d_mean, mu.data(), (size_t)m->F * 4, hipMemcpyHostToDevice, m->stream));
    PV_TRY(hipMemcpyAsync(m->d_istd, is.data(), (size_t)m->F * 4, hipMemcpyHostToDevice, m->stream));
    PV_TRY(hipMemsetAsync(m->d_m, 0, (size_t)m->np * 4, m->stream));
    PV_TRY(hipMemsetAsync(m->d_v, 0, (size_t)m->np * 4, m->stream));
    m->adam_t = 0;
    if (model_sync_bf16(m, errbuf, errlen) != 0) return -1;
    PV_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

extern "C" int fv_gpu_model_get_params(FvGpuModel *m, float *params, float *mean, float *std_,
                                       char *errbuf, size_t errlen) {
    if (m == nullptr) PV_FAIL("gpu: model: NULL");
    PV_TRY(hipSetDevice(m->device));
    if (params) {
        PV_TRY(hipMemcpyAsync(params, m->d_w, (size_t)m->np * 4, hipMemcpyDeviceToHost, m->stream));
        PV_TRY(hipStreamSynchronize(m->stream));
    }
    for (int j = 0; j < m->F; j++) {
        if (mean) mean[j] = m->h_mean[j];
        if (std_) std_[j] = 1.f / m->h_istd[j];
    }
    return 0;
}

/* one full-batch gradient of the mean loss at the current parameters
 * (targets [n][3]: log ef, route, ef-valid) */
extern "C" int fv_gpu_model_grad(FvGpuModel *m, int n, const float *feats, const float *targets,
                                 float lambda, float *grad, double *loss, char *errbuf,
                                 size_t errlen) {
    if (m == nullptr || n < 1 || feats == nullptr || targets == nullptr || grad == nullptr)
        PV_FAIL("gpu: model grad: bad arguments");
    PV_TRY(hipSetDevice(m->device));
    if (ensure_train(m, n, n, errbuf, errlen) != 0) return -1;
    std::vector<int32_t> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    PV_TRY(hipMemcpyAsync(m->d_x.p, feats, (size_t)n * m->F * 4, hipMemcpyHostToDevice, m->stream));
    PV_TRY(hipMemcpyAsync(m->d_t.p, targets, (size_t)n * 12, hipMemcpyHostToDevice, m->stream));
    PV_TRY(hipMemcpyAsync(m->d_idx.p, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, m->stream));
    if (grad_step(m, n, lambda, loss, errbuf, errlen) != 0) return -1;
    PV_TRY(hipMemcpyAsync(grad, m->d_g, (size_t)m->np * 4, hipMemcpyDeviceToHost, m->stream));
    PV_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

/*
 * Minibatch Adam over (feats, targets) for `epochs` passes. The input
 * standardisation is fitted on `feats` first (mean, std per input). A fixed
 * seed shuffles the samples each epoch; loss_hist[e] receives the mean
 * training loss of epoch e.
 */
extern "C" int fv_gpu_model_train(FvGpuModel *m, int n, const float *feats,
                                  const float *targets, int epochs, int batch, float lr,
                                  float lambda, uint64_t seed, float *loss_hist, char *errbuf,
                                  size_t errlen) {
    if (m == nullptr || n < 1 || feats == nullptr || targets == nullptr || epochs < 1 ||
        batch < 1)
        PV_FAIL("gpu: model train: bad arguments");
    PV_TRY(hipSetDevice(m->device));
    const int B = std::min(batch, n);
    if (ensure_train(m, n, B, errbuf, errlen) != 0) return -1;
    /* standardisation from the training set */
    std::vector<double> mu(m->F, 0.0), var(m->F, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m->F; j++) mu[j] += feats[(uint64_t)i * m->F + j];
    for (int j = 0; j < m->F; j++) mu[j] /= n;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m->F; j++) {
            const double d = feats[(uint64_t)i * m->F + j] - mu[j];
            var[j] += d * d;
        }
    std::vector<float> fm(m->F), fs(m->F);
    for (int j = 0; j < m->F; j++) {
        fm[j] = (float)mu[j];
        const double sd = sqrt(var[j] / n);
        fs[j] = sd > 1e-12 ? (float)sd : 1.f;
    }
    std::vector<float> w(m->np);
    PV_TRY(hipMemcpyAsync(w.data(), m->d_w, (size_t)m->np * 4, hipMemcpyDeviceToHost, m->stream));
    PV_TRY(hipStreamSynchronize(m->stream));
    if (fv_gpu_model_set_params(m, w.data(), fm.data(), fs.data(), errbuf, errlen) != 0)
        return -1;
    PV_TRY(hipMemcpyAsync(m->d_x.p, feats, (size_t)n * m->F * 4, hipMemcpyHostToDevice, m->stream));
    PV_TRY(hipMemcpyAsync(m->d_t.p, targets, (size_t)n * 12, hipMemcpyHostToDevice, m->stream));
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; e++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        double tot = 0.0;
        int steps = 0;
        for (int s = 0; s + B <= n; s += B) {
            PV_TRY(hipMemcpyAsync(m->d_idx.p, perm.data() + s, (size_t)B * 4,
                                  hipMemcpyHostToDevice, m->stream));
            double l = 0.0;
            if (grad_step(m, B, lambda, loss_hist ? &l : nullptr, errbuf, errlen) != 0) return -1;
            tot += l;
            steps++;
            m->adam_t++;
            const float c1 = 1.f - powf(b1, (float)m->adam_t);
            const float c2 = 1.f - powf(b2, (float)m->adam_t);
            hipLaunchKernelGGL(k_adam, dim3((m->np + 255) / 256), dim3(256), 0, m->stream, m->d_w,
                               m->d_g, m->d_m, m->d_v, m->np, lr, b1, b2, eps, c1, c2);
            PV_TRY(hipGetLastError());
        }
        if (loss_hist) loss_hist[e] = steps ? (float)(tot / steps) : 0.f;
    }
    if (model_sync_bf16(m, errbuf, errlen) != 0) return -1;
    PV_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

/* BF16 inference on host features [n][F] -> y [n][2] (log ef, route logit) */
extern "C" int fv_gpu_model_predict(FvGpuModel *m, int n, const float *feats, float *y,
                                    char *errbuf, size_t errlen) {
    if (m == nullptr || n < 1 || feats == nullptr || y == nullptr)
        PV_FAIL("gpu: model predict: bad arguments");
    PV_TRY(hipSetDevice(m->device));
    PV_TRY(m->d_x.ensure((size_t)n * m->F * 4));
    PV_TRY(m->d_y.ensure((size_t)n * 8));
    PV_TRY(hipMemcpyAsync(m->d_x.p, feats, (size_t)n * m->F * 4, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_mlp_infer, dim3((n + 3) / 4), dim3(256), 0, m->stream,
                       m->d_x.as<float>(), n, m->F, m->d_wbf, m->d_mean, m->d_istd,
                       m->d_y.as<float>());
    PV_TRY(hipGetLastError());
    PV_TRY(hipMemcpyAsync(y, m->d_y.p, (size_t)n * 8, hipMemcpyDeviceToHost, m->stream));
    PV_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

/* feature vectors of a staged batch's queries (bitmaps rebuilt on the
 * batch's stream) -> host [nq][F] */
extern "C" int fv_gpu_batch_features(FvGpuIndex *g, FvGpuBatch *b, FvGpuModel *m, float *out,
                                     char *errbuf, size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g || m == nullptr || out == nullptr)
        PV_FAIL("gpu: features: bad arguments");
    if (m->dim != g->dim) PV_FAIL("gpu: features: model dim %d != index dim %d", m->dim, g->dim);
    if (require_bitmaps(b, errbuf, errlen) != 0) return -1;
    PV_TRY(hipSetDevice(g->device));
    hipStream_t s = b->stream;
    PV_TRY(launch_bitmap(g, b, s));
    DevBuf cb, fb;
    PV_TRY(cb.ensure((size_t)std::max(1, b->nprog) * 4));
    uint32_t *pc = cb.as<uint32_t>();
    if (b->nprog > 0)
        hipLaunchKernelGGL(k_prog_count, dim3(b->nprog), dim3(256), 0, s,
                           b->d_bitmap.as<uint32_t>(), g->count_pad / 32, pc);
    PV_TRY(fb.ensure((size_t)b->nq * m->F * 4));
    hipLaunchKernelGGL(k_features, dim3((b->nq + 3) / 4), dim3(256), 0, s, b->d_q, b->nq, g->dim,
                       b->d_pofq, pc, g->count, m->d_cent, m->ncent, fb.as<float>());
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, fb.p, (size_t)b->nq * m->F * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    fb.release();
    cb.release();
    if (e != hipSuccess) PV_FAIL("gpu: features: %s", hipGetErrorString(e));
    return 0;
}

namespace pv {

/*
 * Compaction of a routed run over several query tiles: the brute-routed
 * queries (route 0) of the parent batch, in order, become slots 0 .. nb-1
 * of the view batch; later slots are marked skipped, so every query tile
 * past the packed queries is skipped whole by the brute pipeline. One
 * 1024-thread block (nq <= 4096: four consecutive queries per thread, a
 * block-wide prefix over the per-thread counts); k_gather_rows then copies
 * the packed rows.
 */
struct CompactArgs {
    const uint8_t *route;
    int nq, nqp, dim;
    const float *q;
    const int32_t *pofq;
    const uint32_t *topbq;   /* per-query k (uneven batches) or null */
    float *vq;
    int32_t *vpofq;
    uint32_t *vtopbq;
    uint8_t *vskip;
    int32_t *vmap;
};

__global__ __launch_bounds__(1024) void k_compact(CompactArgs A) {
    __shared__ uint32_t tmp[16];
    __shared__ int32_t smap[kMaxBatchEx];
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int q0 = t * 4;
    uint32_t c = 0;
    bool br[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = q0 + j;
        br[j] = q < A.nq && A.route[q] == 0;
        c += br[j] ? 1u : 0u;
    }
    /* inclusive prefix over the block's 16 waves */
    uint32_t v = wave_scan_incl(c);
    if (l == 63) tmp[w] = v;
    __syncthreads();
    uint32_t add = 0, nb = 0;
    for (int i = 0; i < 16; i++) {
        add += i < w ? tmp[i] : 0u;
        nb += tmp[i];
    }
    uint32_t pos = v - c + add;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (br[j]) smap[pos++] = q0 + j;
    __syncthreads();
    for (int i = t; i < A.nqp; i += blockDim.x) {
        const bool live = i < (int)nb;
        const int src = live ? smap[i] : -1;
        A.vmap[i] = src;
        A.vskip[i] = live ? 0 : 1;
        A.vpofq[i] = live ? A.pofq[src] : -1;
        if (A.vtopbq && i < A.nq) A.vtopbq[i] = live ? A.topbq[src] : 1u;
    }
}

/* the packed query rows (zeros past the packed queries): one wave per row,
 * four rows per block */
__global__ __launch_bounds__(256) void k_gather_rows(const int32_t *__restrict__ vmap, int nqp,
                                                     int dim, const float *__restrict__ q,
                                                     float *__restrict__ vq) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (i >= nqp) return;
    const int src = vmap[i];
    const float *sr = q + (uint64_t)(src < 0 ? 0 : src) * dim;
    float *dr = vq + (uint64_t)i * dim;
    for (int j = l; j < dim; j += 64) dr[j] = src < 0 ? 0.f : sr[j];
}

/* the view's results back to the parent's query positions: block i copies
 * slot i's hits, count, flag and status */
__global__ __launch_bounds__(64) void k_scatter(const int32_t *vmap, int nq, uint32_t stride,
                                                const FvHit *vout, const uint32_t *vcnt,
                                                const uint8_t *vflag, const uint32_t *vstatus,
                                                FvHit *out, uint32_t *cnt, uint8_t *flag,
                                                uint32_t *status) {
    const int i = blockIdx.x;
    const int dst = vmap[i];
    if (dst < 0 || i >= nq) return;
    const uint32_t c = vcnt[i];
    for (uint32_t j = threadIdx.x; j < c && j < stride; j += 64)
        out[(uint64_t)dst * stride + j] = vout[(uint64_t)i * stride + j];
    if (threadIdx.x == 0) {
        cnt[dst] = c;
        flag[dst] = vflag[i];
        status[dst] = vstatus[i];
    }
}

/* an arena with no memory: a buffer bound to it is borrowed (never freed,
 * never grown) */
static Arena borrow_arena;

/* the view batch of b for a compacted brute run (created on first use) */
static int setup_view(FvGpuIndex *g, FvGpuBatch *b, char *errbuf, size_t errlen) {
    if (b->view == nullptr) b->view = new FvGpuBatch();
    FvGpuBatch *v = b->view;
    const int nqp = b->nqt * kQT, dim = g->dim;
    v->owner = g;
    v->nq = b->nq;
    v->nqt = b->nqt;
    v->topb = b->topb;
    v->nprog = b->nprog;
    v->nops = b->nops;
    v->nvals = b->nvals;
    v->all_spans = false;
    v->any_dedup = false;
    v->sum_total = (uint64_t)b->nq * g->count;
    v->max_total = g->count;
    v->max_entry = 0;
    v->has_topbq = b->has_topbq;
    v->inl_ok = false;
    v->inl_col = -1;
    v->d_inl_p = nullptr;
    v->i8_ok = b->i8_ok;
    v->stream = b->stream;
    /* the programs and their bitmaps are the parent's */
    v->d_hdr = b->d_hdr;
    v->d_ops = b->d_ops;
    v->d_vals = b->d_vals;
    v->d_poff = b->d_poff;
    v->d_plen = b->d_plen;
    v->d_bitmap.p = b->d_bitmap.p;
    v->d_bitmap.cap = b->d_bitmap.cap;
    v->d_bitmap.arena = &borrow_arena;
    v->d_qd = b->d_qd;
    v->d_spans = b->d_spans;
    v->d_spanoff = b->d_spanoff;
    /* the view's own inputs: packed queries, programs of queries, per-query
     * k, skip flags */
    const size_t o_q = 0;
    const size_t o_p = (o_q + (size_t)nqp * dim * 4 + 255) & ~(size_t)255;
    const size_t o_k = (o_p + (size_t)nqp * 4 + 255) & ~(size_t)255;
    const size_t o_s = (o_k + (size_t)nqp * 4 + 255) & ~(size_t)255;
    const size_t tot = o_s + (size_t)nqp;
    PV_TRY(v->d_stage.ensure(tot));
    unsigned char *d = v->d_stage.as<unsigned char>();
    v->d_q = (float *)(d + o_q);
    v->d_pofq = (int32_t *)(d + o_p);
    v->d_topbq_p = b->has_topbq ? (const uint32_t *)(d + o_k) : nullptr;
    v->d_vskip = d + o_s;
    PV_TRY(b->d_vmap.ensure((size_t)nqp * 4));
    PV_TRY(v->d_qbf.ensure((size_t)nqp * qbf_dim(g) * 2));
    PV_TRY(v->d_qaux.ensure((size_t)nqp * 16));
    if (g->d_vi8 != nullptr) {
        PV_TRY(v->d_qi8.ensure((size_t)nqp * g->dimp));
        PV_TRY(v->d_qn8.ensure((size_t)nqp * 4));
    }
    v->out_stride = b->topb;
    PV_TRY(v->d_out.ensure((size_t)b->nq * b->topb * sizeof(FvHit)));
    PV_TRY(v->d_cnt.ensure((size_t)b->nq * 4));
    PV_TRY(v->d_flag.ensure((size_t)b->nq));
    PV_TRY(v->d_pcnt.ensure((size_t)nqp * kPoolShards * kPcntStride * 4));
    return 0;
}

} // namespace pv

/*
 * One routed run of a staged batch: bitmaps + query prep, features, BF16
 * MLP, per-query route and ef, then the brute pipeline over the queries
 * routed to it (the others skipped on the device) and the graph search over
 * the rest, both writing k hits per query into the batch outputs.
 */
extern "C" int fv_gpu_batch_run_learned(FvGpuIndex *g, FvGpuBatch *b, FvGpuModel *m,
                                        const FvGpuLearnedOpts *o, char *errbuf,
                                        size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g || o == nullptr)
        PV_FAIL("gpu: learned run: bad arguments");
    if (o->mode < 0 || o->mode > 3) PV_FAIL("gpu: learned run: bad mode %d", o->mode);
    if (o->graph_cus != 0)
        PV_FAIL("gpu: learned run: graph_cus %d: splitting the chip between the graph "
                "searches and the scan is not available (must be 0)", o->graph_cus);
    if (o->k != b->topb) PV_FAIL("gpu: learned run: k %u != staged topb %u", o->k, b->topb);
    /* the graph search returns at most FV_GPU_MAX_EF hits: with a larger k
     * the graph cannot answer (the reference falls back to the CPU cursor,
     * gpu_engine.h:40-42), so graph-only modes fail and the learned route
     * sends every query to brute force */
    if (o->k > FV_GPU_MAX_EF && (o->mode == 1 || o->mode == 2))
        PV_FAIL("gpu: learned run: k %u exceeds the graph search's %d results", o->k,
                FV_GPU_MAX_EF);
    FvGpuLearnedOpts oc = *o;
    if (o->k > FV_GPU_MAX_EF && o->mode == 0) oc.mode = 3;
    /* no logit exceeds a bias of +inf: every query takes brute force, so
     * the run is the brute pipeline alone (no features, model or packing) */
    if (o->mode == 0 && std::isinf(o->route_bias) && o->route_bias > 0.f) oc.mode = 3;
    o = &oc;
    const bool need_model = o->mode == 0 || o->mode == 1;
    if (need_model && m == nullptr) PV_FAIL("gpu: learned run: no model");
    if (need_model && m->dim != g->dim) PV_FAIL("gpu: learned run: model dim != index dim");
    if (o->mode != 3 && g->d_nbr == nullptr) PV_FAIL("gpu: learned run: graph not uploaded");
    if (require_bitmaps(b, errbuf, errlen) != 0) return -1;
    const uint32_t ef_min = std::max<uint32_t>(o->ef_min, o->k);
    const uint32_t ef_max = std::min<uint32_t>(std::max(o->ef_max, ef_min), FV_GPU_MAX_EF);
    const uint32_t ef_fixed = std::min<uint32_t>(std::max(o->ef_fixed, ef_min), FV_GPU_MAX_EF);
    try {
        PV_TRY(hipSetDevice(g->device));
        hipStream_t s = b->stream;
        const int nq = b->nq;
        FvGpuBruteOpts bo = {};
        bo.precision = o->precision;
        PV_TRY(b->d_route.ensure((size_t)nq));
        PV_TRY(b->d_efq.ensure((size_t)nq * 4));
        uint32_t *pc = need_model ? b->d_progcnt.as<uint32_t>() : nullptr;
        if (need_model && pc == nullptr) PV_FAIL("gpu: learned run: batch has no count buffer");
        /* bitmaps + query prep; the bitmap pass also counts passing rows,
         * and up to 256 dims the query prep also computes the router
         * features (all but the selectivity) */
        const bool pre = need_model && g->dim <= 256;
        if (need_model) {
            PV_TRY(b->d_feat.ensure((size_t)nq * m->F * 4));
            PV_TRY(b->d_y.ensure((size_t)nq * 8));
        }
        PrepFeatures pf = {};
        if (pre) {
            pf.cent_t = m->d_cent;
            pf.ncent = m->ncent;
            pf.nfeat = m->F;
            pf.feat = b->d_feat.as<float>();
        }
        PV_TRY(launch_phase1(g, b, scan_precision(g, b, &bo), true, s, pc, pre ? &pf : nullptr));
        LearnArgs A = {};
        A.q = b->d_q;
        A.nq = nq;
        A.dim = g->dim;
        A.pofq = b->d_pofq;
        A.progpart = pc;
        A.nbm = bitmap_blocks(g);
        A.count = g->count;
        A.mode = o->mode;
        A.ef_fixed = ef_fixed;
        A.ef_min = ef_min;
        A.ef_max = ef_max;
        A.scale = o->ef_scale > 0.f ? o->ef_scale : 1.f;
        A.bias = o->route_bias;
        A.route = b->d_route.as<uint8_t>();
        A.efq = b->d_efq.as<uint32_t>();
        if (need_model) {
            A.F = m->F;
            A.cent = m->d_cent;
            A.ncent = m->ncent;
            A.wt = m->d_wbf;
            A.mean = m->d_mean;
            A.istd = m->d_istd;
            A.feat_out = b->d_feat.as<float>();
            A.y_out = b->d_y.as<float>();
            /* with features from phase 1, the decision blocks run inside
             * the brute pipeline's threshold launch */
            if (!pre) hipLaunchKernelGGL(k_learn_route, dim3(nq), dim3(256), 0, s, A);
        } else {
            hipLaunchKernelGGL(k_route, dim3((nq + 255) / 256), dim3(256), 0, s,
                               (const float *)nullptr, nq, o->mode, ef_fixed, ef_min, ef_max,
                               1.f, 0.f, A.route, A.efq);
        }
        PV_TRY(hipGetLastError());
        const uint8_t *route = b->d_route.as<uint8_t>();
        /* several query tiles under the learned route: decide first, pack
         * the brute-routed queries into the leading tiles of the view batch
         * and scan only those (query tiles cost the same whatever share of
         * their queries is routed away) */
        const bool compact = o->mode == 0 && b->nqt > 1;
        const uint32_t ef_launch = o->mode == 2 ? ef_fixed : ef_max;
        const uint64_t mv = o->max_visited ? o->max_visited
                                           : std::max<uint64_t>(200000, (uint64_t)ef_launch * 400);
        GraphExtra x;
        x.route = route;
        x.efq = b->d_efq.as<uint32_t>();
        x.out_k = o->k;
        x.bitmap_ready = true;
        if (compact) {
            if (pre) {
                PV_TRY(launch_route_blocks(A, nq, s));
            }
            /* the graph search of the graph-routed queries on a second
             * stream, beside the packed brute pipeline (disjoint outputs);
             * it is enqueued first, so any buffer it grows is grown before
             * the brute side's kernels are queued */
            if (b->aux_stream == nullptr)
                PV_TRY(hipStreamCreateWithFlags(&b->aux_stream, hipStreamNonBlocking));
            if (b->ev_fork == nullptr)
                PV_TRY(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
            if (b->ev_join == nullptr)
                PV_TRY(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
            PV_TRY(hipEventRecord(b->ev_fork, s));
            PV_TRY(hipStreamWaitEvent(b->aux_stream, b->ev_fork, 0));
            x.stream = b->aux_stream;
            if (run_graph_ex(g, b, ef_launch, mv, 8, 1024, nullptr, x, errbuf, errlen) != 0)
                return -1;
            PV_TRY(hipEventRecord(b->ev_join, b->aux_stream));
            if (setup_view(g, b, errbuf, errlen) != 0) return -1;
            FvGpuBatch *v = b->view;
            CompactArgs C;
            C.route = route;
            C.nq = nq;
            C.nqp = b->nqt * kQT;
            C.dim = g->dim;
            C.q = b->d_q;
            C.pofq = b->d_pofq;
            C.topbq = b->d_topbq_p;
            C.vq = v->d_q;
            C.vpofq = v->d_pofq;
            C.vtopbq = const_cast<uint32_t *>(v->d_topbq_p);
            C.vskip = v->d_vskip;
            C.vmap = b->d_vmap.as<int32_t>();
            hipLaunchKernelGGL(k_compact, dim3(1), dim3(1024), 0, s, C);
            PV_TRY(hipGetLastError());
            hipLaunchKernelGGL(k_gather_rows, dim3((C.nqp + 3) / 4), dim3(256), 0, s, C.vmap,
                               C.nqp, C.dim, C.q, C.vq);
            PV_TRY(hipGetLastError());
            const int prec = scan_precision(g, v, &bo);
            PV_TRY(launch_phase1(g, v, prec, false, s, nullptr, nullptr));
            if (run_brute_ex(g, v, &bo, 0, v->d_vskip, true, errbuf, errlen, nullptr) != 0)
                return -1;
            hipLaunchKernelGGL(k_scatter, dim3(nq), dim3(64), 0, s, b->d_vmap.as<int32_t>(), nq,
                               b->topb, v->d_out.as<FvHit>(), v->d_cnt.as<uint32_t>(),
                               v->d_flag.as<uint8_t>(), qaux_of(v).status, b->d_out.as<FvHit>(),
                               b->d_cnt.as<uint32_t>(), b->d_flag.as<uint8_t>(),
                               qaux_of(b).status);
            PV_TRY(hipGetLastError());
            PV_TRY(hipStreamWaitEvent(s, b->ev_join, 0));
        } else {
            if (run_brute_ex(g, b, &bo, 0, route, true, errbuf, errlen,
                             need_model && pre ? &A : nullptr) != 0)
                return -1;
            if (o->mode != 3 &&
                run_graph_ex(g, b, ef_launch, mv, 8, 1024, nullptr, x, errbuf, errlen) != 0)
                return -1;
        }
        b->out_stride = o->k;
        return 0;
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

/* the last routed run's per-query decisions (synchronous) */
extern "C" int fv_gpu_batch_routes(FvGpuIndex *g, FvGpuBatch *b, uint8_t *route, uint32_t *ef,
                                   float *y, char *errbuf, size_t errlen) {
    if (g == nullptr || b == nullptr || b->owner != g || b->d_route.p == nullptr)
        PV_FAIL("gpu: routes: no routed run on this batch");
    PV_TRY(hipSetDevice(g->device));
    if (route)
        PV_TRY(hipMemcpyAsync(route, b->d_route.p, (size_t)b->nq, hipMemcpyDeviceToHost, b->stream));
    if (ef)
        PV_TRY(hipMemcpyAsync(ef, b->d_efq.p, (size_t)b->nq * 4, hipMemcpyDeviceToHost, b->stream));
    if (y) {
        if (b->d_y.p == nullptr) PV_FAIL("gpu: routes: the last routed run used no model");
        PV_TRY(hipMemcpyAsync(y, b->d_y.p, (size_t)b->nq * 8, hipMemcpyDeviceToHost, b->stream));
    }
    PV_TRY(hipStreamSynchronize(b->stream));
    return 0;
}
