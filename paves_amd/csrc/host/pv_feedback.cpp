/*
 * pv_feedback.cpp — the per-index service-time feedback table.
 *
 * The reference worker times every GPU batch on the host and folds it into
 * per-(index, strategy, selectivity bucket) EWMAs that its planner reads
 * (worker.c:296-333, arbiter.h:49-104, fv_feedback_ewma arbiter.h:221-229).
 * The library keeps the same table per index from device event time, so a
 * caller that does not run the reference worker still gets the figures.
 * Host code only (no HIP).
 */
#include <algorithm>

#include "../../../include/paves_gpu.h"

namespace {

/* arbiter.h:221-229: an empty field takes the sample */
inline void ewma(double &field, double sample, double alpha) {
    field = field == 0.0 ? sample : field + alpha * (sample - field);
}

} // namespace

extern "C" int fv_gpu_sel_bucket(int32_t p) {
    if (p >= 200) return 0;
    if (p >= 50) return 1;
    if (p >= 10) return 2;
    if (p >= 2) return 3;
    return 4;
}

extern "C" void fv_gpu_feedback_update(FvGpuFeedback *fb, int strategy, int n,
                                       const int32_t *sel_permille, const uint32_t *want,
                                       double batch_us) {
    if (fb == nullptr || n <= 0 || sel_permille == nullptr || strategy < 0 || strategy > 1)
        return;
    const int si = strategy;
    double wmax = 0.0, wsum[FV_GPU_SEL_BUCKETS] = {0, 0, 0, 0, 0};
    int nb[FV_GPU_SEL_BUCKETS] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        const double w = (si == 1 && want != nullptr)
                             ? (double)want[i] * 1000.0 / std::max<int32_t>(sel_permille[i], 1)
                             : 1.0;
        wmax = std::max(wmax, w);
        const int b = fv_gpu_sel_bucket(sel_permille[i]);
        wsum[b] += w;
        nb[b]++;
    }
    for (int b = 0; b < FV_GPU_SEL_BUCKETS; b++) {
        if (nb[b] == 0) continue;
        ewma(fb->q_us[si][b], (batch_us / std::max(wmax, 1e-9)) * (wsum[b] / nb[b]),
             (double)nb[b] / (nb[b] + 64.0));
    }
    ewma(fb->batch_us[si], batch_us, 0.125);
    ewma(fb->batch_n[si], (double)n, 0.125);
    fb->batches[si]++;
}
