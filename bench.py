#!/usr/bin/env python3
"""Filtered top-10 throughput of the MI355X query path on one GPU.

Workload (the 1M x 128 filtered brute-force configuration): 1,000,000 x 128
integer-valued GMM vectors (256 centres, sigma 20, clipped to [0,218]),
price ~ U[0,1000), category ~ U{0..31}; 10,000 queries from the same
mixture, query i uses filter class i mod 6 (all, price<500, price<100,
category IN (3,7), price<1, category == centre(q) mod 32), served in
micro-batches of 64 through the C ABI's device-resident batch entry points
(fv_gpu_batch_stage once, then fv_gpu_batch_run_brute per micro-batch).

One step = one pass over all 10,000 queries (157 micro-batches). `value` =
queries per second of the BF16-MFMA + exact-FP32-re-rank path with inputs
resident in HBM; the plain FP32 path is timed beside it on the same staged
batches. Results of both are checked against each other (bit-exact) and a
sample against the CPU oracle.

roofline: the dominant kernel is the BF16 full scan (k_scan_bf16<false>);
its average launch duration is measured with HIP events recorded on the
index's stream around that kernel; algorithmic bytes per launch = N x (2*D
bytes of BF16 row + 4 bytes of ||x||^2) + the predicate-bitmap words read
(distinct predicates x N/8).

cpu_baseline: the reference engine (engine.cpp compiled from source under
oracle/_ref) BruteCursor on a bounded sample of the same queries, reading an
index file written in the reference format, all host threads per query.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from paves_amd import _lib as L  # noqa: E402
from paves_amd import engine as E  # noqa: E402
from paves_amd import indexfile, synth  # noqa: E402

METRIC = "filtered top-10 QPS at recall@10≥0.9: learned-ef+BF16 vs plain FP32, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_batches(d, post, n, nq, bsz, k):
    classes = synth.FILTER_CLASSES
    batches = []
    for s in range(0, nq, bsz):
        e = min(nq, s + bsz)
        preds = [synth.filter_pred(classes[i % len(classes)], d["qcent"][i]) for i in range(s, e)]
        b = E.concat([E.serialize(p, n, post) for p in preds], n)
        batches.append((s, e, b))
    return batches


def run_steps(g, staged, opts, steps):
    t0 = time.perf_counter()
    for _ in range(steps):
        for bh in staged:
            g.run_brute(bh, opts)
    g.sync()
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--fp32-steps", type=int, default=0, help="0 = max(2, steps // 4)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the index lives on a single MI355X; see the sharding note in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    t = time.perf_counter()
    d = synth.gmm128(args.n, args.nq)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    log(f"data: {args.n}x128 + {args.nq} queries in {time.perf_counter() - t:.1f}s")
    batches = build_batches(d, post, args.n, args.nq, args.batch, args.k)

    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16, device=0)
    staged = [g.stage(d["queries"][s:e], b, args.k) for s, e, b in batches]
    log(f"staged {len(staged)} micro-batches")
    bf = L.BruteOpts(precision=L.FV_GPU_PREC_BF16, route=0)
    fp = L.BruteOpts(precision=L.FV_GPU_PREC_FP32, route=0)

    # ---- BF16 path: warmup then K timed steps
    run_steps(g, staged, bf, args.warmup)
    g.sync()
    dt = run_steps(g, staged, bf, args.steps)
    qps = args.steps * args.nq / dt
    ms_per_step = dt / args.steps * 1e3
    log(f"bf16: {qps:,.0f} QPS ({ms_per_step:.2f} ms/step)")

    # results of the BF16 path (last step)
    res_bf = [g.fetch(bh, e - s, args.k) for bh, (s, e, _) in zip(staged, batches)]
    overflow = sum(r[3] != 0 for r in res_bf)

    # ---- plain FP32 path on the same staged batches
    fsteps = args.fp32_steps or max(2, args.steps // 4)
    run_steps(g, staged, fp, 1)
    dt32 = run_steps(g, staged, fp, fsteps)
    qps32 = fsteps * args.nq / dt32
    log(f"fp32: {qps32:,.0f} QPS")
    res_fp = [g.fetch(bh, e - s, args.k) for bh, (s, e, _) in zip(staged, batches)]

    # ---- correctness: BF16 == FP32 (both exact), sample == oracle
    exact = True
    for (o1, c1, f1, _), (o2, c2, f2, _) in zip(res_bf, res_fp):
        if not (np.array_equal(c1, c2) and np.array_equal(f1, f2)):
            exact = False
        for q in range(len(c1)):
            c = int(c1[q])
            if not np.array_equal(o1[q, :c], o2[q, :c]):
                exact = False
    sample_ok = None
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        s, e, b = batches[0]
        oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"][s:e], b.qd, b.nodes,
                                      b.children, b.invals, args.k)
        o1, c1, f1, _ = res_bf[0]
        sample_ok = bool(np.array_equal(c1, oc) and np.array_equal(f1, ot) and all(
            np.array_equal(o1[q, : c1[q]], oh[q, : c1[q]]) for q in range(e - s)))
        log(f"oracle check (first micro-batch, {e - s} queries): {sample_ok}")

    # ---- roofline of the dominant kernel (BF16 full scan), HIP events
    g.set_timing(True)
    scan_us, prep_us, sel_us, tot_us, nprog = [], [], [], [], []
    for (s, e, b), bh in zip(batches, staged):
        g.run_brute(bh, bf)
        g.sync()
        st = g.stats()
        if st.route == 1:
            scan_us.append(st.t_scan_us)
            prep_us.append(st.t_prepass_us + st.t_merge_us)
            sel_us.append(st.t_rerank_us)
            tot_us.append(st.t_total_us)
            nprog.append(st.nprog)
    g.set_timing(False)
    avg_scan = float(np.mean(scan_us))
    words_bytes = float(np.mean(nprog)) * args.n / 8.0
    alg_bytes = args.n * (128 * 2 + 4) + words_bytes
    achieved = alg_bytes / (avg_scan * 1e-6) / 1e9
    log(f"scan kernel avg {avg_scan:.1f} us -> {achieved:.0f} GB/s; prepass+tau "
        f"{np.mean(prep_us):.1f} us, select {np.mean(sel_us):.1f} us, total {np.mean(tot_us):.1f} us")

    # ---- CPU baseline: reference BruteCursor on a bounded sample
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(d, cols, rows, vals, args)

    for bh in staged:
        g.free_batch(bh)
    g.close()

    out = {
        "metric": METRIC, "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (seeded GMM, see paves_amd/synth.py)",
        "config": {"workload": "filtered brute-force top-10: 1M x 128 GMM, 10k queries in "
                               "micro-batches of 64, six filter classes",
                   "n": args.n, "dim": 128, "queries": args.nq, "micro_batch": args.batch,
                   "k": args.k, "topb": args.k},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_scan_bf16<false, 4>",
                     "avg_launch_us": round(avg_scan, 2),
                     "alg_bytes_per_launch": int(alg_bytes)},
        "plain_fp32_qps": round(qps32, 1),
        "speedup_vs_plain_fp32": round(qps / qps32, 3),
        "recall_at_10": 1.0 if (exact and sample_ok is not False) else None,
        "ids_bit_exact_bf16_vs_fp32": exact,
        "oracle_sample_ok": sample_ok,
        "overflow_batches": int(overflow),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_baseline(d, cols, rows, vals, args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    try:
        ref = O.RefEngine()
    except RuntimeError as e:
        log(f"cpu baseline skipped: {e}")
        return None
    threads = min(16, os.cpu_count() or 1)
    tmpd = tempfile.mkdtemp(prefix="paves_bench_")
    path = os.path.join(tmpd, "bench.idx")
    try:
        indexfile.write(path, d["base"], cols, post_rows=rows, post_vals=vals)
        idx = ref.open(path)
        classes = synth.FILTER_CLASSES
        # interleave classes; stop after the time budget
        done = 0
        t0 = time.perf_counter()
        order = np.arange(args.nq)
        for qi in order:
            p = synth.filter_pred(classes[qi % len(classes)], d["qcent"][qi])
            ph, keep = ref.compile(idx, p)
            ref.brute(idx, d["queries"][qi], ph, args.k, threads=threads)
            if ph:
                ref.L.fv_pred_free(ph)
            done += 1
            if done % len(classes) == 0 and time.perf_counter() - t0 > args.cpu_seconds:
                break
        el = time.perf_counter() - t0
        ref.close(idx)
        log(f"cpu reference: {done} queries in {el:.1f}s ({threads} threads per query)")
        return {"value": round(done / el, 2), "unit": "queries/s", "cores": threads,
                "kind": "reference",
                "sample": f"first {done} of the 10k queries (classes interleaved), "
                          f"reference BruteCursor (engine.cpp) with {threads} threads per query, "
                          f"same 1M x 128 data read from a reference-format index file"}
    finally:
        try:
            os.unlink(path)
            os.rmdir(tmpd)
        except OSError:
            pass


if __name__ == "__main__":
    main()
