/*
 * pv_noalloc.hip — fv_gpu_brute_batch_noalloc: the synchronous filtered
 * top-B batch (fv_gpu_brute_batch) with no memory allocation.
 *
 * For callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Every device buffer the
 * brute pipeline needs — the packed staging area, BF16 queries, per-query
 * bounds, predicate bitmaps, pre-pass group minima, candidate pool and
 * counters, and the device outputs — is carved from `scratch`, one static
 * device array of kScratchBytes per GPU; host staging comes from a static
 * host array. No hipMalloc, hipFree, hipHostMalloc or other allocating
 * routine is called. A call whose buffers do not fit fails with an error
 * that names the limit (nothing is launched past that point and the
 * stream is drained first).
 *
 * The kernels are the brute pipeline's own; they never read memory they
 * have not written in the same call (counters and status words are
 * cleared by the threshold kernel, bitmaps and minima are written in full),
 * so the scratch needs no clearing between calls. Calls are serialised by
 * a process-wide lock; results are those of fv_gpu_brute_batch on the same
 * arguments (entries past out_cnt[q] in out are unspecified).
 */
#include <string.h>

#include <mutex>

#include "pv_device.h"

static constexpr size_t kScratchBytes = (size_t)64 << 20;     /* device scratch per GPU */
static constexpr size_t kHostScratchBytes = (size_t)4 << 20;  /* host staging */

/* the library's device scratch (one copy per GPU, held until the process
 * ends); host code must use the per-device address from
 * hipGetSymbolAddress, never this name */
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDev = 32;

alignas(256) unsigned char host_scratch[kHostScratchBytes];

std::mutex noalloc_mu;
unsigned char *scratch_ptr[kMaxDev];   /* per device, set once */
FvGpuBatch noalloc_batch[kMaxDev];     /* per device, buffers bound to the arenas */
pv::Arena dev_arena[kMaxDev], host_arena;

/* the scratch address on the current device (first use on a device looks
 * it up with that device current) */
int scratch_on(int dev, char *errbuf, size_t errlen) {
    if (dev < 0 || dev >= kMaxDev) PV_FAIL("gpu: no-alloc: device %d out of range", dev);
    if (scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            PV_FAIL("gpu: no-alloc: scratch address on device %d: %s", dev,
                    hipGetErrorString(e));
        scratch_ptr[dev] = (unsigned char *)p;
    }
    return 0;
}

/* every buffer of the batch empty and bound to the arenas */
void bind_batch(FvGpuBatch *b, pv::Arena *da, pv::Arena *ha) {
    pv::DevBuf *dev[] = {&b->d_qi8, &b->d_qn8, &b->d_stage, &b->d_qbf, &b->d_qaux,
                         &b->d_bitmap, &b->d_gmin, &b->d_pool, &b->d_pcnt, &b->d_out,
                         &b->d_cnt, &b->d_flag, &b->d_visit, &b->d_fspill, &b->d_progcnt,
                         &b->d_feat, &b->d_y, &b->d_route, &b->d_efq};
    for (pv::DevBuf *d : dev) {
        d->p = nullptr;
        d->cap = 0;
        d->arena = da;
    }
    for (pv::PinBuf *h : {&b->h_stage, &b->h_out}) {
        h->p = nullptr;
        h->cap = 0;
        h->arena = ha;
    }
    b->arena_bound = true;
}

/* one pipeline run on the scratch-bound batch; outputs to the caller */
int run_and_copy(FvGpuIndex *g, FvGpuBatch *b, uint32_t stride, FvHit *out, uint32_t *out_cnt,
                 uint8_t *truncated, uint32_t *status, char *errbuf, size_t errlen) {
    FvGpuBruteOpts o = {};
    o.precision = -1;
    if (stride == 1) o.route = 1;
    if (pv::run_brute_ex(g, b, &o, stride, nullptr, false, errbuf, errlen) != 0) return -1;
    const int nq = b->nq;
    PV_TRY(hipMemcpyAsync(out, b->d_out.p, (size_t)nq * b->topb * sizeof(FvHit),
                          hipMemcpyDeviceToHost, b->stream));
    PV_TRY(hipMemcpyAsync(out_cnt, b->d_cnt.p, (size_t)nq * 4, hipMemcpyDeviceToHost, b->stream));
    PV_TRY(hipMemcpyAsync(truncated, b->d_flag.p, (size_t)nq, hipMemcpyDeviceToHost, b->stream));
    PV_TRY(hipMemcpyAsync(status, pv::qaux_of(b).status, (size_t)nq * 4, hipMemcpyDeviceToHost,
                          b->stream));
    PV_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int noalloc_impl(FvGpuIndex *g, int nq, const float *queries, const FvGpuQueryDesc *qd,
                 const uint8_t *dedup, const FvGpuPredNode *nodes, int nnodes,
                 const int32_t *children, int nchildren, const double *invals, int ninvals,
                 const FvGpuSpan *spans, int nspans_total, const uint64_t *spanoff,
                 int nspanoff, uint32_t topb, FvHit *out, uint32_t *out_cnt,
                 uint8_t *truncated, char *errbuf, size_t errlen) {
    if (g == nullptr) PV_FAIL("gpu: NULL index");
    if (nq < 1 || nq > FV_GPU_MAX_BATCH || topb == 0 || topb > FV_GPU_MAX_TOPB ||
        nspans_total > FV_GPU_MAX_SPANS * nq)
        PV_FAIL("gpu: bad brute batch (nq=%d topb=%u)", nq, topb);
    if (out == nullptr || out_cnt == nullptr || truncated == nullptr)
        PV_FAIL("gpu: NULL output arrays");
    std::lock_guard<std::mutex> lk(noalloc_mu);
    const int dev = g->device;
    PV_TRY(hipSetDevice(dev));
    if (scratch_on(dev, errbuf, errlen) != 0) return -1;
    FvGpuBatch *b = &noalloc_batch[dev];
    pv::Arena &da = dev_arena[dev];
    da.base = scratch_ptr[dev];
    da.size = kScratchBytes;
    da.used = 0;
    host_arena.base = host_scratch;
    host_arena.size = kHostScratchBytes;
    host_arena.used = 0;
    bind_batch(b, &da, &host_arena);
    b->stream = g->stream;
    auto exceeded = [&](int rc) {
        if (rc == 0) return 0;
        (void)hipStreamSynchronize(b->stream);
        if (da.used + 4096 >= da.size || host_arena.used + 4096 >= host_arena.size ||
            strstr(errbuf, "out of memory") != nullptr)
            snprintf(errbuf, errlen,
                     "gpu: no-alloc: this batch needs more than the %zu MiB device scratch "
                     "(or %zu MiB host staging); use fv_gpu_brute_batch",
                     kScratchBytes >> 20, kHostScratchBytes >> 20);
        return -1;
    };
    int rc = pv::stage_batch(g, b, nq, queries, qd, dedup, nodes, nnodes, children, nchildren,
                             invals, ninvals, spans, nspans_total, spanoff, nspanoff, topb,
                             errbuf, errlen);
    if (rc != 0) return exceeded(rc);
    uint32_t *status = (uint32_t *)host_arena.take((size_t)nq * 4);
    if (status == nullptr) return exceeded(-1);
    rc = run_and_copy(g, b, 0, out, out_cnt, truncated, status, errbuf, errlen);
    if (rc != 0) return exceeded(rc);
    bool bad = false;
    for (int q = 0; q < nq; q++) bad |= status[q] != 0;
    if (bad) {
        /* pool overflow or oversized re-rank set: the tightest threshold
         * (pre-pass over every row), as fv_gpu_brute_batch does */
        rc = run_and_copy(g, b, 1, out, out_cnt, truncated, status, errbuf, errlen);
        if (rc != 0) return exceeded(rc);
        for (int q = 0; q < nq; q++)
            if (status[q] != 0)
                PV_FAIL("gpu: query %d: candidate set exceeds capacity (status %u)", q, status[q]);
    }
    return 0;
}

} // namespace

extern "C" int fv_gpu_brute_batch_noalloc(FvGpuIndex *g, int nq, const float *queries,
                                          const FvGpuQueryDesc *qd, const uint8_t *dedup,
                                          const FvGpuPredNode *nodes, int nnodes,
                                          const int32_t *children, int nchildren,
                                          const double *invals, int ninvals,
                                          const FvGpuSpan *spans, int nspans_total,
                                          const uint64_t *spanoff, int nspanoff, uint32_t topb,
                                          FvHit *out, uint32_t *out_cnt, uint8_t *truncated,
                                          char *errbuf, size_t errlen) {
    try {
        return noalloc_impl(g, nq, queries, qd, dedup, nodes, nnodes, children, nchildren,
                            invals, ninvals, spans, nspans_total, spanoff, nspanoff, topb, out,
                            out_cnt, truncated, errbuf, errlen);
    } catch (...) {
        snprintf(errbuf, errlen, "gpu: internal error");
        return -1;
    }
}

