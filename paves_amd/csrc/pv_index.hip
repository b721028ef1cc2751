/*
 * pv_index.hip — GPU-resident index: device choice, upload, free, probes.
 *
 * Reference: FvGpuIndex / pick_device / fv_gpu_index_upload /
 * fv_gpu_index_free (gpu_engine.cu:197-234, 1506-1596, 1692-1773),
 * fv_gpu_available (1489), fv_gpu_tile_variant_qt and fv_gpu_mem_probe
 * (2546-2566).
 *
 * Upload copies the caller's host arrays (they may live in a read-only
 * mmap) into HBM and derives the scan layouts on the GPU: a BF16 copy of
 * the vectors, ||x||^2 per row, the level-0 adjacency in a fixed-stride
 * 0xFFFFFFFF-padded layout (one 128-byte line per node for M0 = 32), and
 * the integer/magnitude facts that decide whether BF16 and FP32 scans are
 * exact. All work is on one stream owned by the index.
 */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>
#include <mutex>

#include "pv_device.h"

namespace pv {

void ev_record(FvGpuIndex *g, int i) {
    if (g->timing) hipEventRecord(g->ev[i], g->ev_stream ? g->ev_stream : g->stream);
}

double ev_elapsed_us(FvGpuIndex *g, int a, int b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g->ev[a], g->ev[b]) != hipSuccess) return 0.0;
    return (double)ms * 1000.0;
}

namespace {

constexpr int kMaxDevices = 32;
int indexes_on[kMaxDevices];   /* live indexes per device */
std::mutex indexes_mu;

constexpr size_t kPickMargin = (size_t)512 << 20;   /* working memory kept free */

/* Device for an index uploaded with device = -1: among the devices with
 * room for `need` plus kPickMargin, the one with the most free HBM (lower
 * ordinal on a tie) — the ABI's contract for -1 (gpu_engine.cu:1518-1545).
 * Free memory is read with each device current; the chosen device is left
 * current, or the caller's when none fits. No device is reset: HIP's
 * primary contexts are shared with the rest of the process. */
int pick_device(size_t need) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -1;
    int prev = 0;
    (void)hipGetDevice(&prev);
    struct Fit {
        size_t free;
        int dev;
    };
    std::vector<Fit> fits;
    for (int d = 0; d < ndev; d++) {
        size_t fr = 0, tot = 0;
        if (hipSetDevice(d) == hipSuccess && hipMemGetInfo(&fr, &tot) == hipSuccess &&
            fr >= need + kPickMargin)
            fits.push_back({fr, d});
    }
    (void)hipGetLastError();
    if (fits.empty()) {
        (void)hipSetDevice(prev);
        return -1;
    }
    const Fit best = *std::max_element(fits.begin(), fits.end(), [](const Fit &a, const Fit &b) {
        return a.free < b.free || (a.free == b.free && a.dev > b.dev);
    });
    (void)hipSetDevice(best.dev);
    return best.dev;
}

/* fp32 -> MFMA-tiled bf16 copy and ||x||^2, plus the largest positive and
 * negative coordinate, max ||x||^2 and an "all integers" flag. Each wave
 * walks rows wave, wave + waves, ... (one row at a time, lanes over its
 * dims) and folds its rows into wave maxima, so the index-wide statistics
 * take one set of atomics per wave. The copy has dimb >= dim dims (a
 * multiple of 128); dims past dim hold zeros, which add exactly nothing to
 * a dot product. */
__global__ __launch_bounds__(256) void k_derive_rows(
    const float *__restrict__ v, uint64_t count, uint64_t count_pad, int dim, int dimb,
    uint16_t *__restrict__ vbf, float *__restrict__ norm,
    unsigned int *__restrict__ stats /* [0]=max x, [1]=max ||x||^2, [2]=nonint, [3]=max -x */) {
    const int l = threadIdx.x & 63;
    const uint64_t waves = (uint64_t)gridDim.x * 4;
    /* MFMA-tiled BF16 position of (row, j): tile row/32, k-step
     * (j mod dimb/2)/8, lane (row mod 32) + 32 (j div dimb/2), element j mod 8 */
    const int half = dimb / 2;
    const int jend = vbf ? dimb : dim;
    float wp = 0.f, wn = 0.f, wss = 0.f;
    int wni = 0;
    for (uint64_t row = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); row < count_pad;
         row += waves) {
        float ss = 0.f, pmax = 0.f, nmax = 0.f;
        int nonint = 0;
        const uint64_t tbase = (row >> 5) * 32ull * dimb;
        for (int j = l; j < jend; j += 64) {
            float x = (row < count && j < dim) ? v[row * dim + j] : 0.f;
            if (vbf) {
                const int g = j / half, w = j % half;
                vbf[tbase + (uint64_t)(w >> 3) * 512 + ((row & 31) + 32 * g) * 8 + (w & 7)] =
                    f2bf(x);
            }
            ss = fmaf(x, x, ss);
            pmax = fmaxf(pmax, x);
            nmax = fmaxf(nmax, -x);
            nonint |= (x != rintf(x)) ? 1 : 0;
        }
#pragma unroll
        for (int o = 32; o; o >>= 1) {
            ss += lane_xor(ss, o);
            pmax = fmaxf(pmax, lane_xor(pmax, o));
            nmax = fmaxf(nmax, lane_xor(nmax, o));
            nonint |= lane_xor(nonint, o);
        }
        if (l == 0) norm[row] = row < count ? ss : 0.f;
        if (row < count) {
            wp = fmaxf(wp, pmax);
            wn = fmaxf(wn, nmax);
            wss = fmaxf(wss, ss);
            wni |= nonint;
        }
    }
    if (l == 0) {
        atomicMax(&stats[0], __float_as_uint(wp));
        atomicMax(&stats[1], __float_as_uint(wss));
        if (wni) atomicOr(&stats[2], 1u);
        atomicMax(&stats[3], __float_as_uint(wn));
    }
}

/* float <-> order-preserving uint (atomicMin / atomicMax on floats) */
__device__ __forceinline__ uint32_t ord_of(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
static inline float float_of_ord(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

/* per-dim minimum and maximum over the rows (order-preserving encodings;
 * lo starts at all ones, hi at zero). Thread t of a block owns dims t,
 * t + 256, ... of a contiguous row range. */
__global__ __launch_bounds__(256) void k_dim_range(const float *__restrict__ v, uint64_t count,
                                                   int dim, uint64_t rows_per_block,
                                                   uint32_t *__restrict__ lo,
                                                   uint32_t *__restrict__ hi) {
    const uint64_t r0 = (uint64_t)blockIdx.x * rows_per_block;
    const uint64_t r1 = min(count, r0 + rows_per_block);
    for (int d = threadIdx.x; d < dim; d += blockDim.x) {
        float mn = __uint_as_float(0x7f800000u), mx = -mn;
        for (uint64_t r = r0; r < r1; r++) {
            const float x = v[r * dim + d];
            mn = fminf(mn, x);
            mx = fmaxf(mx, x);
        }
        if (r0 < r1) {
            atomicMin(&lo[d], ord_of(mn));
            atomicMax(&hi[d], ord_of(mx));
        }
    }
}

/* INT8 copy of an SQ8 index in the MFMA-tiled int8 layout of
 * v_mfma_i32_32x32x32_i8 — tiles of 32 rows, k-steps of 32 dims as 1 KB
 * chunks, slot l (16 bytes) holding row (l & 31), dims 32 s + 16 (l >> 5)
 * .. +16 — over dimp dims (zero codes past dim), and the squared code norm
 * (an integer below 2^24, exact in FP32). Codes: x - centre (exact mode,
 * mid == nullptr) or clamp(rint((x - mid[d]) * inv), -127, 127). One wave
 * per row. fp8: E4M3 codes e4m3_encode((x - mid[d]) * inv) in the same
 * layout, and the squared norm of the decoded codes (FP32, each lane's
 * dims in order, then the butterfly). */
__global__ __launch_bounds__(256) void k_derive_i8(const float *__restrict__ v, uint64_t count,
                                                   uint64_t count_pad, int dim, int dimp,
                                                   int centre, const float *__restrict__ mid,
                                                   float inv, int fp8, int8_t *__restrict__ vi8,
                                                   float *__restrict__ norm8) {
#pragma clang fp contract(off)
    const int l = threadIdx.x & 63;
    const uint64_t row = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= count_pad) return;
    const uint64_t tbase = (row >> 5) * 32ull * dimp;
    int32_t ss = 0;
    float sf = 0.f;
    for (int j = l; j < dimp; j += 64) {
        int code = 0;
        if (row < count && j < dim) {
            const float x = v[row * dim + j];
            if (fp8) {
                code = e4m3_encode((x - mid[j]) * inv);
                const float dv = e4m3_decode((uint32_t)code);
                sf = fmaf(dv, dv, sf);
            } else if (mid == nullptr) {
                code = (int)(x - (float)centre);
            } else {
                const float c = rintf((x - mid[j]) * inv);
                code = (int)fminf(fmaxf(c, -127.f), 127.f);
            }
        }
        const int s = j >> 5, g = (j >> 4) & 1;
        vi8[tbase + (uint64_t)s * 1024 + ((row & 31) + 32 * g) * 16 + (j & 15)] = (int8_t)code;
        ss += code * code;
    }
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        ss += __shfl_xor(ss, o);
        sf += __shfl_xor(sf, o);
    }
    if (l == 0) norm8[row] = fp8 ? sf : (float)ss;
}

/* rank[c][prows[c][i]] = i: each row's position in its column's posting
 * order */
__global__ void k_rank(const uint32_t *__restrict__ prows, uint64_t count, uint64_t nt,
                       uint32_t *__restrict__ rank) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const uint64_t c = i / count;
    rank[c * count + prows[i]] = (uint32_t)(i - c * count);
}

/* level-0 lists [cnt, neigh...] (stride l0_stride) -> padded rows of
 * nbr_stride ids, 0xFFFFFFFF after the count; ids >= count are dropped. */
__global__ void k_pack_l0(const uint32_t *__restrict__ l0, uint32_t l0_stride,
                          uint64_t count, int nbr_stride,
                          uint32_t *__restrict__ nbr) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count * (uint64_t)nbr_stride) return;
    const uint64_t node = i / nbr_stride;
    const uint32_t k = (uint32_t)(i % nbr_stride);
    const uint32_t *lst = l0 + node * l0_stride;
    uint32_t cnt = lst[0];
    if (cnt > l0_stride - 1) cnt = l0_stride - 1;
    uint32_t v = kNoNode;
    if (k < cnt) {
        v = lst[1 + k];
        if (v >= count) v = kNoNode;
    }
    nbr[i] = v;
}

} // namespace
} // namespace pv

using namespace pv;

extern "C" int fv_gpu_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0;
}

extern "C" int fv_gpu_index_device(const FvGpuIndex *g) { return g ? g->device : -1; }

extern "C" int fv_gpu_index_precision(const FvGpuIndex *g) {
    return g ? g->precision : -1;
}

extern "C" void fv_gpu_index_free(FvGpuIndex *g) {
    if (!g) return;
    if (g->device >= 0) hipSetDevice(g->device);
    if (g->stream) hipStreamSynchronize(g->stream);
    for (hipStream_t bs : g->bstreams)
        if (bs) hipStreamSynchronize(bs);
    hipFree(g->d_vecs);
    hipFree(g->d_vbf);
    hipFree(g->d_norm);
    hipFree(g->d_vi8);
    hipFree(g->d_norm8);
    hipFree(g->d_i8mid);
    hipFree(g->d_cols);
    hipFree(g->d_prows);
    hipFree(g->d_rank);
    hipFree(g->d_nbr);
    hipFree(g->d_ones);
    if (g->sync_batch) {
        free_batch_buffers(g->sync_batch);
        delete g->sync_batch;
    }
    for (int i = 0; i < 16; i++)
        if (g->ev[i]) hipEventDestroy(g->ev[i]);
    for (int i = 0; i < 2; i++)
        if (g->fb_ev[i]) hipEventDestroy(g->fb_ev[i]);
    if (g->stream) hipStreamDestroy(g->stream);
    for (hipStream_t bs : g->bstreams)
        if (bs) hipStreamDestroy(bs);
    {
        std::lock_guard<std::mutex> lk(indexes_mu);
        if (g->device >= 0 && g->device < kMaxDevices && indexes_on[g->device] > 0)
            indexes_on[g->device]--;
    }
    delete g;
}

static FvGpuIndex *upload_impl(const float *vectors, const double *cols,
                               const uint32_t *post_rows, const uint32_t *l0,
                               uint32_t l0_stride, uint64_t count, int dim,
                               int ncols, int device, int precision, char *errbuf,
                               size_t errlen) {
#define UP_FAIL(...)                                                          \
    do {                                                                      \
        snprintf(errbuf, errlen, __VA_ARGS__);                                \
        fv_gpu_index_free(g);                                                 \
        return nullptr;                                                       \
    } while (0)
#define UP_TRY(call)                                                          \
    do {                                                                      \
        hipError_t e__ = (call);                                              \
        if (e__ != hipSuccess)                                                \
            UP_FAIL("gpu: index upload failed: %s: %s", #call,                \
                    hipGetErrorString(e__));                                  \
    } while (0)
    FvGpuIndex *g = nullptr;
    if (vectors == nullptr || count == 0 || dim <= 0 || ncols < 0 ||
        count >= 0xFFFFFFFFull || (ncols > 0 && cols == nullptr)) {
        snprintf(errbuf, errlen, "gpu: bad upload arguments (count=%llu dim=%d ncols=%d)",
                 (unsigned long long)count, dim, ncols);
        return nullptr;
    }
    if (precision != FV_GPU_PREC_FP32 && precision != FV_GPU_PREC_SQ8 &&
        precision != FV_GPU_PREC_BF16 && precision != FV_GPU_PREC_FP8) {
        snprintf(errbuf, errlen, "gpu: unknown precision %d", precision);
        return nullptr;
    }
    const uint64_t count_pad = (count + 63) & ~63ull;
    /* the BF16 copy, zero-padded to a multiple of 128 dims, serves every
     * dim up to 1024 */
    const bool want_bf = dim <= 1024;
    const int dimb = want_bf ? (dim + 127) / 128 * 128 : 0;
    const size_t need = count_pad * ((size_t)dim * 4 + (size_t)dimb * 2) + count_pad * 4 +
                        count * (size_t)ncols * 12 + (l0 ? count * (size_t)l0_stride * 8 : 0);
    if (device < 0) device = pick_device(need);
    if (device < 0) {
        snprintf(errbuf, errlen, "gpu: no HIP device with %.1f GB free",
                 (need + ((size_t)512 << 20)) / 1e9);
        return nullptr;
    }
    g = new FvGpuIndex();
    g->device = device;
    g->precision = precision;
    g->count = count;
    g->count_pad = count_pad;
    g->dim = dim;
    g->dimb = dimb;
    g->ncols = ncols;
    g->l0_stride = l0 ? l0_stride : 0;
    UP_TRY(hipSetDevice(device));
    {
        std::lock_guard<std::mutex> lk(indexes_mu);
        if (device < kMaxDevices) indexes_on[device]++;
    }
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            g->ncu = prop.multiProcessorCount;
    }
    UP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    {
        const char *ev = getenv("PAVES_BATCH_STREAMS");
        const int ns = ev ? atoi(ev) : 3;
        g->nstreams = std::min(FvGpuIndex::kMaxBatchStreams, std::max(1, ns));
    }
    for (int i = 0; i < g->nstreams; i++)
        UP_TRY(hipStreamCreateWithFlags(&g->bstreams[i], hipStreamNonBlocking));
    for (int i = 0; i < 16; i++) UP_TRY(hipEventCreate(&g->ev[i]));
    hipStream_t s = g->stream;

    UP_TRY(hipMalloc(&g->d_vecs, count_pad * (size_t)dim * 4));
    UP_TRY(hipMemsetAsync(g->d_vecs, 0, count_pad * (size_t)dim * 4, s));
    UP_TRY(hipMemcpyAsync(g->d_vecs, vectors, count * (size_t)dim * 4,
                          hipMemcpyHostToDevice, s));
    if (want_bf) UP_TRY(hipMalloc(&g->d_vbf, count_pad * (size_t)dimb * 2));
    UP_TRY(hipMalloc(&g->d_norm, count_pad * 4));
    unsigned int *d_stats = nullptr;
    UP_TRY(hipMalloc(&d_stats, 16));
    UP_TRY(hipMemsetAsync(d_stats, 0, 16, s));
    hipLaunchKernelGGL(k_derive_rows,
                       dim3((unsigned)std::min<uint64_t>((count_pad + 3) / 4, 4096)), dim3(256), 0, s,
                       g->d_vecs, count, count_pad, dim, dimb, g->d_vbf, g->d_norm, d_stats);
    {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            hipFree(d_stats);
            UP_FAIL("gpu: derive kernel: %s", hipGetErrorString(e));
        }
    }
    unsigned int hs[4] = {0, 0, 0, 0};
    {
        hipError_t e = hipMemcpyAsync(hs, d_stats, 16, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        hipFree(d_stats);
        if (e != hipSuccess) UP_FAIL("gpu: derive stats: %s", hipGetErrorString(e));
    }
    float pmax, nmax, n2max;
    memcpy(&pmax, &hs[0], 4);
    memcpy(&n2max, &hs[1], 4);
    memcpy(&nmax, &hs[3], 4);
    g->vmax = pmax;
    g->vmin = -nmax;
    g->vmax_abs = fmaxf(pmax, nmax);
    g->xmax_norm = sqrtf(n2max) * (1.0f + 1e-6f);
    g->all_int = hs[2] == 0;

    /* SQ8: an INT8 copy. Exact codes x - centre when every value is an
     * integer, the values (0 included) span at most 255 and dim <= 256
     * (dot products and norms stay below 2^24, exact in FP32); otherwise,
     * up to 1024 dims, quantized codes of the per-dim midrange-centred
     * values under one scale (the largest half-range maps to 127). */
    /* FP8: E4M3 codes of the same midrange-centred rows, the largest
     * half-range scaled to 448 (the format's largest finite value) */
    const bool fp8 = precision == FV_GPU_PREC_FP8;
    if ((precision == FV_GPU_PREC_SQ8 || fp8) && dim <= 1024) {
        const bool exact = !fp8 && g->all_int && g->vmax - g->vmin <= 255.0f && dim <= 256;
        const int dimp = dim <= 64 ? 64 : dim <= 128 ? 128 : (dim + 255) / 256 * 256;
        g->dimp = dimp;
        g->i8_mode = exact ? 1 : (fp8 ? 3 : 2);
        if (exact) {
            g->i8_center = (int)g->vmin + 128;
        } else {
            uint32_t *d_r = nullptr;
            UP_TRY(hipMalloc(&d_r, (size_t)dim * 8));
            UP_TRY(hipMemsetAsync(d_r, 0xFF, (size_t)dim * 4, s));
            UP_TRY(hipMemsetAsync(d_r + dim, 0, (size_t)dim * 4, s));
            const uint64_t rpb = (count + 1023) / 1024;
            hipLaunchKernelGGL(k_dim_range, dim3((unsigned)((count + rpb - 1) / rpb)), dim3(256),
                               0, s, g->d_vecs, count, dim, rpb, d_r, d_r + dim);
            std::vector<uint32_t> hr((size_t)dim * 2);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(hr.data(), d_r, (size_t)dim * 8, hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            hipFree(d_r);
            if (e != hipSuccess) UP_FAIL("gpu: SQ8 range: %s", hipGetErrorString(e));
            std::vector<float> mid((size_t)dimp, 0.f);
            float hmax = 0.f;
            for (int d = 0; d < dim; d++) {
                const float lo = float_of_ord(hr[d]), hi = float_of_ord(hr[dim + d]);
                mid[d] = (lo + hi) * 0.5f;
                hmax = fmaxf(hmax, (hi - lo) * 0.5f);
            }
            g->i8_inv = hmax > 0.f ? (fp8 ? 448.0f : 127.0f) / hmax : 1.0f;
            UP_TRY(hipMalloc(&g->d_i8mid, (size_t)dimp * 4));
            UP_TRY(hipMemcpyAsync(g->d_i8mid, mid.data(), (size_t)dimp * 4,
                                  hipMemcpyHostToDevice, s));
        }
        UP_TRY(hipMalloc(&g->d_vi8, count_pad * (size_t)dimp));
        UP_TRY(hipMalloc(&g->d_norm8, count_pad * 4));
        hipLaunchKernelGGL(k_derive_i8, dim3((unsigned)((count_pad + 3) / 4)), dim3(256), 0, s,
                           g->d_vecs, count, count_pad, dim, dimp, g->i8_center, g->d_i8mid,
                           g->i8_inv, fp8 ? 1 : 0, g->d_vi8, g->d_norm8);
        UP_TRY(hipGetLastError());
    }

    if (ncols > 0) {
        UP_TRY(hipMalloc(&g->d_cols, (size_t)ncols * count * 8));
        UP_TRY(hipMemcpyAsync(g->d_cols, cols, (size_t)ncols * count * 8,
                              hipMemcpyHostToDevice, s));
        if (post_rows != nullptr) { /* optional: only the span-gather route reads them */
            UP_TRY(hipMalloc(&g->d_prows, (size_t)ncols * count * 4));
            UP_TRY(hipMemcpyAsync(g->d_prows, post_rows, (size_t)ncols * count * 4,
                                  hipMemcpyHostToDevice, s));
            /* ranks in posting order, when every column's postings list the
             * non-NULL values ascending and then the NULLs */
            bool sorted = count > 0 && count < 0xFFFFFFFFull;
            std::vector<std::vector<double>> sv((size_t)ncols);
            for (int c = 0; c < ncols && sorted; c++) {
                const double *cv = cols + (size_t)c * count;
                const uint32_t *pr = post_rows + (size_t)c * count;
                std::vector<double> &v = sv[(size_t)c];
                v.reserve(count);
                uint64_t i = 0;
                for (; i < count; i++) {
                    if (pr[i] >= count) { sorted = false; break; }
                    const double x = cv[pr[i]];
                    if (x != x) break;
                    if (!v.empty() && x < v.back()) { sorted = false; break; }
                    v.push_back(x);
                }
                for (; i < count && sorted; i++)
                    if (pr[i] >= count || cv[pr[i]] == cv[pr[i]]) sorted = false;
            }
            if (sorted) {
                UP_TRY(hipMalloc(&g->d_rank, (size_t)ncols * count * 4));
                const uint64_t nt = (uint64_t)ncols * count;
                hipLaunchKernelGGL(k_rank, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s,
                                   g->d_prows, count, nt, g->d_rank);
                UP_TRY(hipGetLastError());
                g->h_sorted = std::move(sv);
            }
        }
    }
    /* bitmap row used by queries without a predicate */
    UP_TRY(hipMalloc(&g->d_ones, count_pad / 8));
    UP_TRY(hipMemsetAsync(g->d_ones, 0xFF, count_pad / 8, s));
    if (l0 != nullptr && l0_stride > 1) {
        const int deg = (int)l0_stride - 1;
        g->nbr_stride = (deg + 7) & ~7;
        uint32_t *d_l0 = nullptr;
        UP_TRY(hipMalloc(&d_l0, count * (size_t)l0_stride * 4));
        hipError_t e = hipMemcpyAsync(d_l0, l0, count * (size_t)l0_stride * 4,
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMalloc(&g->d_nbr, count * (size_t)g->nbr_stride * 4);
        if (e == hipSuccess) {
            const uint64_t n = count * (uint64_t)g->nbr_stride;
            hipLaunchKernelGGL(k_pack_l0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               d_l0, l0_stride, count, g->nbr_stride, g->d_nbr);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        hipFree(d_l0);
        if (e != hipSuccess) UP_FAIL("gpu: graph upload: %s", hipGetErrorString(e));
    }
    UP_TRY(hipStreamSynchronize(s));
    return g;
#undef UP_TRY
#undef UP_FAIL
}

extern "C" FvGpuIndex *fv_gpu_index_upload(const float *vectors, const double *cols,
                                           const uint32_t *post_rows,
                                           const uint32_t *l0, uint32_t l0_stride,
                                           uint64_t count, int dim, int ncols,
                                           int device, int precision, char *errbuf,
                                           size_t errlen) {
    try {
        return upload_impl(vectors, cols, post_rows, l0, l0_stride, count, dim, ncols,
                           device, precision, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error during upload");
        return nullptr;
    }
}

/* Query-tile width of the fast brute path (engine_api.h:190-194): the scan
 * processes 64 queries per pass for every dimension. */
extern "C" int fv_gpu_tile_variant_qt(int dim) { return dim > 0 ? kQT : 0; }

extern "C" int fv_gpu_mem_probe(uint64_t *free_b, uint64_t *total_b) {
    static int probed = -2;
    if (probed == -2) probed = pick_device(0);
    if (probed < 0) return -1;
    size_t f = 0, t = 0;
    if (hipSetDevice(probed) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess)
        return -1;
    if (free_b) *free_b = f;
    if (total_b) *total_b = t;
    return probed;
}

extern "C" void fv_gpu_set_timing(FvGpuIndex *g, int on) {
    if (g) g->timing = on != 0;
}

extern "C" int fv_gpu_feedback(const FvGpuIndex *g, FvGpuFeedback *out) {
    if (g == nullptr || out == nullptr) return -1;
    std::lock_guard<std::mutex> lk(const_cast<FvGpuIndex *>(g)->fb_mu);
    *out = g->fb;
    return 0;
}

extern "C" void fv_gpu_last_stats(const FvGpuIndex *g, FvGpuStats *st) {
    if (g && st) *st = g->last;
}

/* ---------------- copy bandwidth (roofline reference) ---------------- */

namespace pv {
namespace {
typedef uint32_t cp_u32x4 __attribute__((ext_vector_type(4)));
/* grid-stride 16-byte copy: each lane moves consecutive 16-byte words, a
 * wave covers 1 KiB contiguous per step */
__global__ __launch_bounds__(256) void k_copy_calib(const cp_u32x4 *__restrict__ src,
                                                    cp_u32x4 *__restrict__ dst, uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

/* grid-stride 16-byte streaming read (ordinary cached loads, as the scans
 * use), four loads in flight per lane; the words are folded by XOR and
 * stored only if they match a value the fill never produces, so the loads
 * stay live while nothing is written */
__global__ __launch_bounds__(256) void k_read_calib(const cp_u32x4 *__restrict__ src,
                                                    cp_u32x4 *__restrict__ dst, uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    cp_u32x4 acc = {0u, 0u, 0u, 0u};
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const cp_u32x4 a = src[i];
        const cp_u32x4 b = src[i + stride];
        const cp_u32x4 c = src[i + 2 * stride];
        const cp_u32x4 d = src[i + 3 * stride];
        acc ^= a ^ b ^ c ^ d;
    }
    for (; i < n16; i += stride) acc ^= src[i];
    if (acc.x == 0xDEADBEEFu && acc.y == 0x0BADF00Du) dst[0] = acc;
}
} // namespace
} // namespace pv

/* device copy (read_only = false: read + write) or streaming read over
 * `bytes`, timed with events over reps launches; GB/s of bytes moved */
static int mem_bench(int device, uint64_t bytes, int reps, bool read_only, double *gbs,
                     char *errbuf, size_t errlen) {
    if (bytes == 0 || bytes % 16 != 0 || reps < 1 || gbs == nullptr)
        PV_FAIL("gpu: memory bench: bad arguments");
    PV_TRY(hipSetDevice(device));
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, read_only ? 4096 : bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 1, bytes, s);
    float ms = 0.f;
    if (e == hipSuccess) {
        const uint64_t n16 = bytes / 16;
        const unsigned grid = 8192;
        auto launch = [&]() {
            if (read_only)
                hipLaunchKernelGGL(pv::k_read_calib, dim3(grid), dim3(256), 0, s,
                                   (const pv::cp_u32x4 *)a, (pv::cp_u32x4 *)b, n16);
            else
                hipLaunchKernelGGL(pv::k_copy_calib, dim3(grid), dim3(256), 0, s,
                                   (const pv::cp_u32x4 *)a, (pv::cp_u32x4 *)b, n16);
            return hipGetLastError();
        };
        e = launch();
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; r++) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e1) hipEventDestroy(e1);
    if (e0) hipEventDestroy(e0);
    if (s) hipStreamDestroy(s);
    if (b) hipFree(b);
    if (a) hipFree(a);
    if (e != hipSuccess) PV_FAIL("gpu: memory bench: %s", hipGetErrorString(e));
    const double moved = (read_only ? 1.0 : 2.0) * (double)bytes * reps;
    *gbs = ms > 0.f ? moved / ((double)ms * 1e-3) / 1e9 : 0.0;
    return 0;
}

extern "C" int fv_gpu_copy_bench(int device, uint64_t bytes, int reps, double *gbs,
                                 char *errbuf, size_t errlen) {
    return mem_bench(device, bytes, reps, false, gbs, errbuf, errlen);
}

extern "C" int fv_gpu_read_bench(int device, uint64_t bytes, int reps, double *gbs,
                                 char *errbuf, size_t errlen) {
    return mem_bench(device, bytes, reps, true, gbs, errbuf, errlen);
}
