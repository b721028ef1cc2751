#!/usr/bin/env python3
"""Filtered top-10 throughput of the MI355X query path on one GPU.

Workload (the 1M x 128 filtered configuration): 1,000,000 x 128
integer-valued GMM vectors (256 centres, sigma 20, clipped to [0,218]),
price ~ U[0,1000), category ~ U{0..31}; 10,000 queries from the same
mixture, query i uses filter class i mod 6 (all, price<500, price<100,
category IN (3,7), price<1, category == centre(q) mod 32), k = 10, served
through the C ABI's device-resident batch entry points (fv_gpu_batch_stage
once, then a run per batch). Batches are spread over the library's batch
streams, so one batch's small kernels overlap another batch's scan.

One step = one pass over all 10,000 queries. `value` = queries per second
of the learned path: routed runs of 4096 queries (fv_gpu_batch_run_learned,
mode 0) whose route threshold and ef factor are calibrated on held-out
training queries to the fastest setting with recall@10 >= 0.9 in every
filter class (a threshold of +inf when no graph share pays: the run is the
BF16 brute pipeline). Beside it, on the same staged batches: BF16 brute
force at micro-batches of 64 and batches of 4096, the plain FP32 path (FP32
MFMA scan + the same certified re-rank) at both sizes, SQ8 on integer data
(exact INT8) and on a float copy of the workload (quantized codes +
exact re-rank of the top 4 k), the reference's synchronous entry point with
host buffers (PCIe-inclusive, with its per-call breakdown), and the graph
kernel's launch over the workload's first 1024 queries (and unfiltered).
BF16 and FP32 results are checked against each other bit for bit and a
micro-batch against the CPU oracle.

roofline: the dominant kernel of the headline's timed region is the
query-resident BF16 MFMA scan of a routed run of 4096 queries,
k_scan_bf16_qr<8, false, false> (about 80% of each step): MFMA-bound, so
`roofline` is its TFLOP/s against the 2.5 PFLOP/s dense BF16 peak (HIP
events over 10 back-to-back launches on the batch's stream; `python
bench.py --batched-scan-only` is the rocprofv3 target), with `traffic` from
FETCH_SIZE + WRITE_SIZE passes over that command scaled by the copy kernel.
`roofline_microbatch` is the HBM-bound kernel of the micro-batch line: the
BF16 full scan of one 64-query tile,
k_scan_bf16<false, 8, false>, when a batch has the chip to itself. Its
launch duration is measured with HIP events on the batch's stream: for
each scan-route micro-batch the library re-launches that batch's scan
kernel 10 times back to back between two events (fv_gpu_batch_time_scan).
Algorithmic bytes per launch = N x (2*D bytes of BF16 row + 4 bytes of
||x||^2) + the predicate-bitmap words read (distinct predicates x N/8).
`traffic` comes from two rocprofv3 --pmc passes (FETCH_SIZE, then
WRITE_SIZE) over the isolated runs (`--roofline-only`), each scaled by what
the counter reads for a device copy of known size in the same pass; `python
bench.py --roofline-only` alone is the command whose rocprofv3 kernel trace
gives the kernel's average duration. The batches of 4096 run the multi-tile
scan, reported as `batched_scan` (TFLOP/s of the MFMA peak). The graph
kernel's launch is accounted from its own counters (neighbour lists of the
expanded heads, rows and bitmap words of the visited nodes), with a
FETCH_SIZE pass over `--graph-roofline-only`.

cpu_baseline: the reference engine (engine.cpp compiled from source under
oracle/_ref) BruteCursor on a bounded sample of the same queries, reading an
index file written in the reference format, all host threads per query;
cpu_hnsw_baseline: its HnswCursor on the same graph.

`--workload uneven [--dim 128|768]`: one batch of 4,096 queries with k
uniform over {1, 10, 100} and selectivity s = 10^U(-4,0) realised as price <
1000 s, over 1M x 128 (integer GMM) or 2M x 768 (1024-centre GMM, rows
unit-normalised then x10). The launch shape comes from the library's
autotuner (fv_gpu_batch_tune); the run is captured once as a hipGraph and
replayed (one step = one replay); then a stream of distinct batches is
restaged into the captured batch (fv_gpu_batch_restage) and replayed. The
scan of this workload is MFMA-bound, so its roofline is TFLOP/s.

`--workload graph`: filtered HNSW beam search over the same 1M x 128 data
(M = 16, efC = 200), ef in {40, 80, 160, 400}, batches of 64 and 1024, the
exact frontier (spill store) and the bounded one; the best row is gated on
recall@10 >= 0.9 in every filter class.

`--workload build`: HNSW construction of that graph (1M x 128, M = 16,
efC = 200) with level 0 on the GPU (fv_hnsw_build_gpu), seconds per build
(lower is better; one step = one build), against the library's host builder
on the same threads and the compiled reference builder (cpu_baseline);
recall@10 of both graphs at ef 40 and 80. Run it with few steps
(`--steps 2 --warmup 1`).
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import signal
import sqlite3
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from paves_amd import _lib as L  # noqa: E402
from paves_amd import engine as E  # noqa: E402
from paves_amd import hnsw, indexfile, learned, synth  # noqa: E402

METRIC = "filtered top-10 QPS at recall@10≥0.9: learned-ef+BF16 vs plain FP32, 1 GPU"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFS = 157.3    # dense FP32 matrix (MFMA) peak of MI355X
SCAN_KERNEL = "k_scan_bf16<false, 8, false>"   # main scan: 8 waves per WG, bitmap predicates
COPY_BYTES = 1 << 28   # 256 MiB device copy used to calibrate the PMC byte counters
QR_KERNEL = "k_scan_bf16_qr<8, false, false>"  # query-resident scan, batches of 8+ tiles
GRAPH_KERNEL = "k_graph"   # filtered level-0 beam search, one 4-wave WG per query
GRAPH_NBR = 32             # neighbour-list slots read per expanded head (M0 = 32)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_batches(d, post, n, nq, bsz, k, entries=None):
    classes = synth.FILTER_CLASSES
    batches = []
    for s in range(0, nq, bsz):
        e = min(nq, s + bsz)
        preds = [synth.filter_pred(classes[i % len(classes)], d["qcent"][i]) for i in range(s, e)]
        b = E.concat([E.serialize(p, n, post) for p in preds], n,
                     entries=None if entries is None else entries[s:e])
        batches.append((s, e, b))
    return batches


def run_steps(g, staged, opts, steps):
    t0 = time.perf_counter()
    for _ in range(steps):
        for bh in staged:
            g.run_brute(bh, opts)
    g.sync()
    return time.perf_counter() - t0


def measure_scan(g, batches, staged, bf, n, reps=10):
    """Main-scan kernel duration. Per scan-route micro-batch: the library
    re-launches the batch's scan kernel `reps` times back to back between
    two HIP events on the batch's stream (fv_gpu_batch_time_scan), the
    figure rocprofv3's kernel trace reports as the kernel's duration. The
    per-stage times of isolated runs (events around each kernel of one
    batch, so each includes its own dispatch) are reported beside it."""
    g.set_timing(True)
    scan_us, prep_us, sel_us, tot_us, nprog, idx = [], [], [], [], [], []
    for i, ((s, e, b), bh) in enumerate(zip(batches, staged)):
        g.run_brute(bh, bf)
        g.sync()
        st = g.stats()
        if st.route == 1:
            scan_us.append(st.t_scan_us)
            prep_us.append(st.t_prepass_us + st.t_merge_us)
            sel_us.append(st.t_rerank_us)
            tot_us.append(st.t_total_us)
            nprog.append(st.nprog)
            idx.append(i)
    g.set_timing(False)
    b2b = [g.time_scan(staged[i], bf, reps) for i in idx]
    avg_scan = float(np.mean(b2b))
    alg_bytes = n * (128 * 2 + 4) + float(np.mean(nprog)) * n / 8.0
    return {"avg_scan_us": avg_scan, "isolated_scan_us": float(np.mean(scan_us)),
            "scan_batches": len(idx),
            "alg_bytes": alg_bytes, "launches": len(idx) * reps,
            "prepass_tau_us": float(np.mean(prep_us)), "select_us": float(np.mean(sel_us)),
            "total_us": float(np.mean(tot_us))}


def copy_calibration(reps=20):
    """The library's k_copy_calib over COPY_BYTES (read COPY_BYTES, write
    COPY_BYTES per launch); returns the read+write rate in GB/s."""
    import ctypes
    lib = L.load()
    gbs = ctypes.c_double(0.0)
    eb = L.errbuf()
    if lib.fv_gpu_copy_bench(0, COPY_BYTES, reps, ctypes.byref(gbs), eb, len(eb)) != 0:
        raise RuntimeError(eb.value.decode())
    return gbs.value


def pmc_pass(counter, args, child=None, kernel=None):
    """One rocprofv3 --pmc pass over a child `bench.py` run (default:
    `--roofline-only --copy-calib`, the scan kernel); returns the median KiB
    per dispatch of the measured kernel and of the copy kernel."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    out = tempfile.mkdtemp(prefix="paves_pmc_", dir="/tmp")
    if child is None:
        child = ["--roofline-only", "--copy-calib", "--n", str(args.n),
                 "--nq", str(min(args.nq, 1024)), "--batch", str(args.batch), "--k", str(args.k)]
    kernel = kernel or SCAN_KERNEL
    cmd = [prof, "--pmc", counter, "-d", out, "-o", "pmc", "--",
           sys.executable, os.path.abspath(__file__)] + child
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                             stderr=subprocess.DEVNULL, start_new_session=True)
        try:
            p.wait(timeout=180)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            log(f"pmc {counter}: timed out")
            return None
        dbs = [os.path.join(r, f) for r, _, fs in os.walk(out) for f in fs if f.endswith(".db")]
        if not dbs:
            log(f"pmc {counter}: no database (rc {p.returncode})")
            return None
        c = sqlite3.connect(dbs[0])
        rows = c.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                         "where counter_name = ? group by dispatch_id", (counter,)).fetchall()
        c.close()
    finally:
        shutil.rmtree(out, ignore_errors=True)
    scan = [v for _, k, v in rows if kernel.replace(" ", "") in k.replace(" ", "")]
    copy = [v for _, k, v in rows if "k_copy_calib" in k]
    if not scan or not copy:
        log(f"pmc {counter}: scan {len(scan)} / copy {len(copy)} dispatches found")
        return None
    return {"scan_kib": float(np.median(scan)), "copy_kib": float(np.median(copy))}


def measure_traffic(args, child=None, kernel=None):
    """HBM bytes per scan launch: FETCH_SIZE + WRITE_SIZE, each scaled by
    the counter's reading for the known-size copy (COPY_BYTES each way)."""
    f = pmc_pass("FETCH_SIZE", args, child=child, kernel=kernel)
    w = pmc_pass("WRITE_SIZE", args, child=child, kernel=kernel) if f else None
    if not f or not w:
        return None, None
    kib = COPY_BYTES / 1024.0
    fs, ws = kib / f["copy_kib"], kib / w["copy_kib"]
    fetch = f["scan_kib"] * fs * 1024.0
    write = w["scan_kib"] * ws * 1024.0
    info = {"fetch_kib_raw": round(f["scan_kib"], 1), "write_kib_raw": round(w["scan_kib"], 1),
            "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
            "fetch_bytes": int(fetch), "write_bytes": int(write)}
    return fetch + write, info


def graph_launch(g, queries, b, ef, mv, dim, fcap=1024):
    """One beam-search launch over a batch: its kernel duration (HIP events
    around the search kernel) and its algorithmic bytes from the kernel's
    own counters — one neighbour list (GRAPH_NBR ids) per expanded head,
    one FP32 vector row per visited node and, for a filtered query, one
    predicate-bitmap word per visited node."""
    g.graph(queries, b, ef, mv, wmax=8, fcap=fcap)   # warm-up (buffers, code)
    g.set_timing(True)
    _, _, _, dbg = g.graph(queries, b, ef, mv, wmax=8, fcap=fcap, debug=True)
    us = g.stats().t_graph_us
    g.set_timing(False)
    heads = dbg[:, 0].astype(np.float64)
    vis = dbg[:, 1].astype(np.float64)
    filt = (b.qd["proot"] >= 0).astype(np.float64)
    alg = float((heads * GRAPH_NBR * 4).sum() + (vis * (dim * 4 + 4 * filt)).sum())
    gbs = alg / (us * 1e-6) / 1e9
    return {"kernel": GRAPH_KERNEL, "queries": len(queries), "ef": ef,
            "frontier": "exact" if fcap > 0 else "bounded",
            "heads_expanded": int(heads.sum()), "nodes_visited": int(vis.sum()),
            "alg_bytes_per_launch": int(alg), "avg_launch_us": round(us, 2),
            "achieved_GBs": round(gbs, 1), "peak_GBs": HBM_PEAK_GBS,
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "ns_per_visited_node": round(us * 1e3 / max(1.0, float(vis.sum())), 4)}


def graph_roofline_only(args):
    """The beam-search launch of graph_launch over the first 1024 queries of
    the workload (six classes), the graph and entry points read from
    --graph-file (written by the parent run) or built here; prints one JSON
    line."""
    d = synth.gmm128(args.n, args.nq)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    if args.graph_file:
        z = np.load(args.graph_file)
    else:
        f = hnsw.build(d["base"], M=16, ef_construction=200, threads=args.threads)
        z = {"l0": f["l0"], "ent": hnsw.entries(f, d["base"], d["queries"][:1024],
                                                threads=args.threads)}
    nq = len(z["ent"])
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, args.n, post) for p in preds], args.n, entries=z["ent"])
    g = E.GpuIndex(d["base"], cols, rows, l0=z["l0"], precision=L.FV_GPU_PREC_BF16, device=0)
    m = graph_launch(g, d["queries"][:nq], b, args.graph_ef, max(200_000, 400 * args.graph_ef),
                     128)
    copy_calibration(reps=3)
    g.close()
    print(json.dumps(m), flush=True)


def graph_traffic(args, graph, ent):
    """HBM bytes of the beam-search launch: one FETCH_SIZE pass over the
    child run of graph_roofline_only, scaled by the copy kernel's reading
    (the kernel writes only its results: ef hits per query)."""
    path = os.path.join(tempfile.mkdtemp(prefix="paves_graph_", dir="/tmp"), "g.npz")
    np.savez(path, l0=graph["l0"], ent=ent)
    try:
        f = pmc_pass("FETCH_SIZE", args, child=["--graph-roofline-only", "--graph-file", path,
                                                "--graph-ef", str(args.graph_ef),
                                                "--n", str(args.n), "--nq", str(args.nq)],
                     kernel=GRAPH_KERNEL)
    finally:
        shutil.rmtree(os.path.dirname(path), ignore_errors=True)
    if not f:
        return None
    fs = COPY_BYTES / 1024.0 / f["copy_kib"]
    return {"fetch_kib_raw": round(f["scan_kib"], 1), "fetch_scale": round(fs, 4),
            "fetch_bytes": int(f["scan_kib"] * fs * 1024.0)}


def roofline_only(args):
    """Isolated scan-route runs only (for rocprofv3); prints one JSON line."""
    d = synth.gmm128(args.n, args.nq)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    batches = build_batches(d, post, args.n, args.nq, args.batch, args.k)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16, device=0)
    staged = [g.stage(d["queries"][s:e], b, args.k) for s, e, b in batches]
    bf = L.BruteOpts(precision=L.FV_GPU_PREC_BF16, route=0)
    for bh in staged:            # warm-up
        g.run_brute(bh, bf)
    g.sync()
    m = measure_scan(g, batches, staged, bf, args.n)
    if args.copy_calib:
        copy_calibration(reps=3)
    for bh in staged:
        g.free_batch(bh)
    g.close()
    print(json.dumps(m), flush=True)


def batched_scan_only(args):
    """The multi-tile scans of one batch of --learned-batch queries alone
    (for rocprofv3): the query-resident kernel (library default) and the
    two-tiles-per-WG kernel, each timed by measure_batched_scan; prints one
    JSON line."""
    d = synth.gmm128(args.n, args.learned_batch)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    batches = build_batches(d, post, args.n, args.learned_batch, args.learned_batch, args.k)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16, device=0)
    staged = [g.stage(d["queries"][s:e], b, args.k) for s, e, b in batches]
    bf = L.BruteOpts(precision=L.FV_GPU_PREC_BF16, route=1)
    m = measure_batched_scan(g, staged, batches, bf, args, reps=10)
    if args.copy_calib:
        copy_calibration(reps=3)
    for bh in staged:
        g.free_batch(bh)
    g.close()
    print(json.dumps(m), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--workload", choices=["filtered", "uneven", "graph", "build"],
                    default="filtered",
                    help="filtered: the 1M x 128 workload (default); uneven: one 4096-query "
                         "batch with mixed k and selectivity; graph: filtered HNSW beam "
                         "search, ef sweep with recall@10; build: HNSW construction of the "
                         "graph workload's graph (level 0 on the GPU)")
    ap.add_argument("--dim", type=int, choices=[128, 768], default=128,
                    help="uneven workload: 1M x 128 or 2M x 768")
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--stream-batches", type=int, default=4,
                    help="uneven workload: distinct batches restaged through one capture")
    ap.add_argument("--batch", type=int, default=64, help="micro-batch of the brute baselines")
    ap.add_argument("--learned-batch", type=int, default=4096,
                    help="queries per routed run of the learned path (packed brute tiles)")
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--train-queries", type=int, default=200_000,
                    help="training queries for the learned router (0 = skip the learned path)")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads (graph build, entry points, CPU baselines)")
    ap.add_argument("--fp32-steps", type=int, default=0, help="0 = max(2, steps // 4)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes (traffic = null)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the isolated scan runs (the rocprofv3 target)")
    ap.add_argument("--copy-calib", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--batched-scan-only", action="store_true",
                    help="time the multi-tile scans of one --learned-batch batch (for rocprofv3)")
    ap.add_argument("--graph-roofline-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--graph-file", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--graph-ef", type=int, default=80,
                    help="ef of the beam-search roofline launch (1024 queries, six classes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the index lives on a single MI355X; sharding it across GPUs is not built)",
              file=sys.stderr)
        sys.exit(2)
    args.n_set = args.n is not None
    if args.n is None:
        args.n = 1_000_000
    if args.workload == "uneven":
        uneven_path(args)
        return
    if args.workload == "graph":
        graph_path(args)
        return
    if args.workload == "build":
        build_path(args)
        return
    if args.roofline_only:
        roofline_only(args)
        return
    if args.batched_scan_only:
        batched_scan_only(args)
        return
    if args.graph_roofline_only:
        graph_roofline_only(args)
        return
    learned_on = args.train_queries > 0

    traffic, traffic_info = (None, None)
    traffic_qr, traffic_qr_info = (None, None)
    if not args.no_traffic:
        t = time.perf_counter()
        traffic, traffic_info = measure_traffic(args)
        log(f"pmc traffic: {traffic_info} ({time.perf_counter() - t:.0f}s)")
        if learned_on:
            t = time.perf_counter()
            traffic_qr, traffic_qr_info = measure_traffic(
                args, child=["--batched-scan-only", "--copy-calib", "--n", str(args.n),
                             "--learned-batch", str(args.learned_batch), "--k", str(args.k)],
                kernel=QR_KERNEL)
            log(f"pmc traffic (query-resident scan): {traffic_qr_info} "
                f"({time.perf_counter() - t:.0f}s)")

    t = time.perf_counter()
    d = synth.gmm128(args.n, args.nq)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    log(f"data: {args.n}x128 + {args.nq} queries in {time.perf_counter() - t:.1f}s")

    graph, ent, setup = None, None, {}
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16, device=0)
    if learned_on:
        # the graph, built with level 0 on this GPU and kept there
        t = time.perf_counter()
        graph = hnsw.build_gpu(g, d["base"], M=16, ef_construction=200, threads=args.threads,
                               keep=True)
        setup["graph_build_s"] = round(time.perf_counter() - t, 2)
        setup["graph_builder"] = "fv_hnsw_build_gpu (level 0 on the GPU)"
        t = time.perf_counter()
        ent = hnsw.entries(graph, d["base"], d["queries"], threads=args.threads)
        setup["entry_points_us_per_query"] = round((time.perf_counter() - t) / args.nq * 1e6, 3)
        log(f"graph: M=16 efC=200 built in {setup['graph_build_s']}s (level 0 on the GPU, "
            f"{args.threads} host threads), entry points "
            f"{setup['entry_points_us_per_query']} us/query")
    batches = build_batches(d, post, args.n, args.nq, args.batch, args.k, ent)
    batchesL = build_batches(d, post, args.n, args.nq, args.learned_batch, args.k, ent)

    staged = [g.stage(d["queries"][s:e], b, args.k) for s, e, b in batches]
    stagedL = [g.stage(d["queries"][s:e], b, args.k) for s, e, b in batchesL]
    log(f"staged {len(staged)} micro-batches of {args.batch} and {len(stagedL)} batches of "
        f"{args.learned_batch}")
    bf = L.BruteOpts(precision=L.FV_GPU_PREC_BF16, route=0)
    fp = L.BruteOpts(precision=L.FV_GPU_PREC_FP32, route=0)

    # ---- BF16 brute path: micro-batches of 64, then batches of --learned-batch
    run_steps(g, staged, bf, args.warmup)
    g.sync()
    dt = run_steps(g, staged, bf, args.steps)
    qps_bf = args.steps * args.nq / dt
    log(f"bf16 brute, micro-batches of {args.batch}: {qps_bf:,.0f} QPS "
        f"({dt / args.steps * 1e3:.2f} ms/step)")
    res_bf = [g.fetch(bh, e - s, args.k) for bh, (s, e, _) in zip(staged, batches)]
    overflow = sum(r[3] != 0 for r in res_bf)
    truth = [o[q, : c[q]]["node"].copy() for (o, c, _, _) in res_bf for q in range(len(c))]
    run_steps(g, stagedL, bf, args.warmup)
    dtL = run_steps(g, stagedL, bf, args.steps)
    qps_bfL = args.steps * args.nq / dtL
    log(f"bf16 brute, batches of {args.learned_batch}: {qps_bfL:,.0f} QPS")
    res_bfL = [g.fetch(bh, e - s, args.k) for bh, (s, e, _) in zip(stagedL, batchesL)]
    # the hits the reference worker asks for: max(4 k, 64) (exec.c:286)
    wk = max(4 * args.k, 64)
    stagedW = [g.stage(d["queries"][s:e], b, wk) for s, e, b in batchesL]
    run_steps(g, stagedW, bf, max(1, args.warmup))
    dtW = run_steps(g, stagedW, bf, max(2, args.steps // 4))
    qps_bfW = max(2, args.steps // 4) * args.nq / dtW
    log(f"bf16 brute, batches of {args.learned_batch}, top-{wk} (the worker's request): "
        f"{qps_bfW:,.0f} QPS")
    for bh in stagedW:
        g.free_batch(bh)

    # ---- plain FP32 path (FP32 MFMA x.q + certified exact re-rank) on the
    # same staged batches, at both batch sizes
    fsteps = args.fp32_steps or max(2, args.steps // 4)
    run_steps(g, staged, fp, 1)
    dt32 = run_steps(g, staged, fp, fsteps)
    qps32 = fsteps * args.nq / dt32
    res_fp = [g.fetch(bh, e - s, args.k) for bh, (s, e, _) in zip(staged, batches)]
    run_steps(g, stagedL, fp, 1)
    dt32L = run_steps(g, stagedL, fp, fsteps)
    qps32L = fsteps * args.nq / dt32L
    log(f"plain fp32 (FP32 MFMA) brute: {qps32:,.0f} QPS (batches of {args.batch}), "
        f"{qps32L:,.0f} QPS (batches of {args.learned_batch})")
    fp32_scan = measure_fp32_scan(g, staged, fp, args.n)

    # ---- SQ8 (exact INT8 MFMA scan of x - centre) on a second index
    sq8 = sq8_path(d, cols, rows, batches, res_bf, args, batchesL)
    # ---- SQ8 over non-integer data: quantized codes + exact re-rank
    sq8q, fp8q = sq8_quant_path(args)

    # ---- the reference's synchronous entry point (host buffers in and out:
    # predicate compile, H2D staging, pipeline, D2H), one pass, for the
    # PCIe-inclusive rate, with its per-stage breakdown
    qps_sync, sync_detail = sync_abi(g, d, batches, args)
    # the same entry point at the reference's largest batch (FV_GPU_MAX_BATCH)
    batches1k = batchesL if args.learned_batch == 1024 else build_batches(
        d, post, args.n, args.nq, 1024, args.k, ent)
    qps_syncL, sync_detailL = sync_abi(g, d, batches1k, args)
    sync_detail["batches_of_1024"] = dict(sync_detailL, qps=round(qps_syncL, 1))

    # ---- correctness: BF16 == FP32 (both exact), batch sizes agree, sample == oracle
    exact = same_results(res_bf, res_fp) and same_results(
        [r for r in res_bf], split_results(res_bfL, batchesL, batches))
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

    # ---- learned route + ef over BF16 brute / graph (the headline path)
    lrn = None
    if learned_on:
        lrn = learned_path(g, graph, d, post, stagedL, batchesL, truth, args, setup)

    # ---- roofline of the brute-force kernel (BF16 full scan, one query
    # tile), HIP events; and the batched scan's MFMA rate
    m = measure_scan(g, batches, staged, bf, args.n)
    avg_scan, alg_bytes = m["avg_scan_us"], m["alg_bytes"]
    achieved = alg_bytes / (avg_scan * 1e-6) / 1e9
    log(f"scan kernel avg {avg_scan:.1f} us -> {achieved:.0f} GB/s (isolated launches "
        f"{m['isolated_scan_us']:.1f} us); prepass+tau "
        f"{m['prepass_tau_us']:.1f} us, select {m['select_us']:.1f} us, "
        f"total {m['total_us']:.1f} us (isolated)")
    batched_scan = measure_batched_scan(g, stagedL, batchesL, bf, args)
    copy_gbs = None
    try:
        copy_gbs = copy_calibration()
        log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")
    except RuntimeError as ex:
        log(f"device copy not measured: {ex}")

    # ---- CPU baselines: reference BruteCursor and HnswCursor, bounded samples
    cpu, cpu_hnsw = None, None
    if not args.no_cpu:
        cpu = cpu_baseline(d, cols, rows, vals, args)
        if graph is not None:
            cpu_hnsw = graph_cpu_baseline(d, cols, rows, vals, graph, ent, truth, args,
                                          ef=lrn["cpu_hnsw_ef"] if lrn else 400)

    for bh in staged + stagedL:
        g.free_batch(bh)
    g.close()

    if lrn is not None:
        value, steps_ms, rec = lrn["qps"], lrn["ms_per_step"], lrn["recall_at_10"]
    else:
        value, steps_ms, rec = qps_bf, dt / args.steps * 1e3, (1.0 if exact else None)
    plain = max(qps32, qps32L)
    roof_main = None
    if lrn is not None and batched_scan is not None:
        # the dominant kernel of the timed region: the learned path's routed
        # runs spend most of each step in the query-resident scan
        # full-batch scan time, scaled by each run's share of a full batch
        runs = sum((e - s) / args.learned_batch for s, e, _ in batchesL)
        share = batched_scan["avg_launch_us"] * runs / (steps_ms * 1e3)
        qr_alg = args.n * (128 * 2 + 4)
        roof_main = {"bound": "mfma", "achieved": batched_scan["achieved_TFLOPs"],
                     "peak": 2500.0, "unit": "TFLOP/s", "frac": batched_scan["frac"],
                     "traffic": int(traffic_qr) if traffic_qr else None,
                     "kernel": QR_KERNEL, "avg_launch_us": batched_scan["avg_launch_us"],
                     "flops_per_launch": batched_scan["flops_per_launch"],
                     "alg_bytes_per_launch": int(qr_alg),
                     "traffic_detail": traffic_qr_info,
                     "share_of_step": round(share, 3),
                     "note": f"the query-resident BF16 MFMA scan of one routed run of "
                             f"{args.learned_batch} queries (2 x N x 128 FLOP per query; "
                             "HIP events over 10 back-to-back launches on the batch's stream); "
                             "algorithmic bytes = N x (256 B BF16 row + 4 B norm); the "
                             "predicate bitmaps (N/8 B per distinct program) are left out"}
    if roof_main is None:
        roof_main = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": int(traffic) if traffic else None, "kernel": SCAN_KERNEL,
                     "avg_launch_us": round(avg_scan, 2)}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(steps_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (seeded GMM, see paves_amd/synth.py)",
        "config": {"workload": (f"filtered top-10, learned route/ef over BF16 brute and HNSW, "
                                f"routed runs of {args.learned_batch} queries: 1M x 128 GMM, "
                                f"10k queries, six filter classes") if lrn else
                               (f"filtered brute-force top-10: 1M x 128 GMM, 10k queries in "
                                f"micro-batches of {args.batch}, six filter classes"),
                   "n": args.n, "dim": 128, "queries": args.nq, "micro_batch": args.batch,
                   "learned_batch": args.learned_batch if lrn else None,
                   "k": args.k, "topb": args.k,
                   "graph": "HNSW M=16 (M0=32) efC=200" if lrn else None,
                   "train_queries": args.train_queries if lrn else 0},
        "roofline": roof_main,
        "roofline_microbatch": {
                     "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": int(traffic) if traffic else None, "kernel": SCAN_KERNEL,
                     "avg_launch_us": round(avg_scan, 2),
                     "isolated_launch_us": round(m["isolated_scan_us"], 2),
                     "alg_bytes_per_launch": int(alg_bytes),
                     "launches": m["launches"],
                     "copy_peak_GBs": round(copy_gbs, 1) if copy_gbs else None,
                     "frac_of_copy_peak": round(achieved / copy_gbs, 4) if copy_gbs else None,
                     "traffic_detail": traffic_info,
                     "note": "the brute-force kernel (one 64-query tile per workgroup, the "
                             "micro-batch shape); the learned path's batches of "
                             f"{args.learned_batch} run the multi-tile scan reported in "
                             "batched_scan",
                     "overlapped": {
                         "kernel": "k_scan_bf16<false, 4, false>",
                         "note": "in the timed region batches overlap on the library's streams "
                                 "and each scan takes the shared-chip shape (4-wave WGs); this "
                                 "is the aggregate scan rate there: algorithmic scan bytes of "
                                 "all scan-route micro-batches per step / step time",
                         "aggregate_GBs": round(m["scan_batches"] * alg_bytes * args.steps / dt
                                                / 1e9, 1)}},
        "batched_scan": batched_scan,
        "recall_at_10": round(rec, 4) if rec is not None else None,
        "bf16_brute_qps": round(qps_bf, 1),
        "bf16_brute_qps_batched": round(qps_bfL, 1),
        "bf16_brute_qps_batched_worker_topb": {"topb": wk, "qps": round(qps_bfW, 1)},
        "plain_fp32_qps": round(plain, 1),
        "plain_fp32": {"kind": "FP32 MFMA x.q (v_mfma_f32_32x32x2_f32) + certified exact "
                               "FP32 re-rank", "qps_micro_batches": round(qps32, 1),
                       "qps_batched": round(qps32L, 1), "scan": fp32_scan},
        "speedup_vs_plain_fp32": round(value / plain, 3),
        "sync_abi_qps_host_buffers": round(qps_sync, 1),
        "sync_abi_breakdown": sync_detail,
        "sq8": sq8,
        "sq8_quantized": sq8q,
        "fp8": fp8q,
        "ids_bit_exact_bf16_vs_fp32": exact,
        "oracle_sample_ok": sample_ok,
        "overflow_batches": int(overflow),
        "learned": lrn,
        "setup": setup,
        "cpu_baseline": cpu,
        "cpu_hnsw_baseline": cpu_hnsw,
    }
    print(json.dumps(out))


def same_results(ra, rb):
    """Two lists of (out, cnt, flag, rc) agree in counts, flags, ids and distances."""
    for (o1, c1, f1, _), (o2, c2, f2, _) in zip(ra, rb):
        if not (np.array_equal(c1, c2) and np.array_equal(f1, f2)):
            return False
        for q in range(len(c1)):
            c = int(c1[q])
            if not np.array_equal(o1[q, :c], o2[q, :c]):
                return False
    return True


def split_results(res, batches_from, batches_to):
    """Results fetched per batch of one size, re-cut to another batch size."""
    o = np.concatenate([r[0] for r in res])
    c = np.concatenate([r[1] for r in res])
    f = np.concatenate([r[2] for r in res])
    return [(o[s:e], c[s:e], f[s:e], 0) for s, e, _ in batches_to]


def measure_fp32_scan(g, staged, fp, n, reps=5, nb=8):
    """The FP32 MFMA scan kernel alone (back-to-back launches on a few
    micro-batches): duration and FLOP rate against the FP32 matrix peak."""
    us = float(np.mean([g.time_scan(bh, fp, reps) for bh in staged[:nb]]))
    flops = 2.0 * n * 128 * 64
    tf = flops / (us * 1e-6) / 1e12
    log(f"fp32 MFMA scan: {us:.1f} us per 64-query launch -> {tf:.1f} TFLOP/s")
    return {"kernel": "k_scan_f32m<false, 8>", "avg_launch_us": round(us, 2),
            "flops_per_launch": flops, "achieved_TFLOPs": round(tf, 2),
            "peak_TFLOPs": FP32_PEAK_TFS, "frac": round(tf / FP32_PEAK_TFS, 4)}


def measure_batched_scan(g, stagedL, batchesL, bf, args, reps=5):
    """The main scan of a full batch of args.learned_batch queries (the
    multi-query-tile kernel) alone: MFMA rate and HBM rate."""
    full = [bh for bh, (s, e, _) in zip(stagedL, batchesL) if e - s == args.learned_batch]
    if not full:
        return None
    nq = args.learned_batch
    flops = 2.0 * args.n * 128 * nq

    def timed(opts):
        g.run_brute(full[0], opts)
        g.sync()
        us = g.time_scan(full[0], opts, reps)
        return us, flops / (us * 1e-6) / 1e12

    # the library's default shape for a batch of >= 8 query tiles with
    # predicate bitmaps is the query-resident kernel (qtw 8)
    us, tf = timed(L.BruteOpts(precision=bf.precision, route=bf.route, qtw=8))
    us2, tf2 = timed(L.BruteOpts(precision=bf.precision, route=bf.route, qtw=2))
    log(f"batched scan ({nq} queries): query-resident {us:.1f} us -> {tf:.0f} TFLOP/s; "
        f"two tiles per WG {us2:.1f} us -> {tf2:.0f} TFLOP/s")
    return {"kernel": "k_scan_bf16_qr<8> (query-resident: one 64-query tile per wave, "
                      "the library default for batches of >= 8 query tiles)",
            "queries": nq, "avg_launch_us": round(us, 2), "flops_per_launch": flops,
            "achieved_TFLOPs": round(tf, 1), "peak_TFLOPs": 2500.0,
            "frac": round(tf / 2500.0, 4),
            "two_tiles_per_wg": {"kernel": "k_scan_bf16_mq<false, 2, 4, false>",
                                 "avg_launch_us": round(us2, 2), "achieved_TFLOPs": round(tf2, 1),
                                 "frac": round(tf2 / 2500.0, 4)}}


def sync_abi(g, d, batches, args):
    """fv_gpu_brute_batch (host buffers in and out) over every micro-batch,
    once, after one untimed call that sizes the library's buffers for this
    batch shape (a worker calls it repeatedly); with event timing, the
    device time of its pipeline per call."""
    s, e, b = batches[0]
    g.brute(d["queries"][s:e], b, args.k)
    t0 = time.perf_counter()
    for s, e, b in batches:
        g.brute(d["queries"][s:e], b, args.k)
    el = time.perf_counter() - t0
    qps = args.nq / el
    g.set_timing(True)
    dev = []
    for s, e, b in batches[:32]:
        g.brute(d["queries"][s:e], b, args.k)
        dev.append(g.stats().t_total_us)
    g.set_timing(False)
    per_call = el / len(batches) * 1e6
    dev_us = float(np.mean(dev))
    log(f"fv_gpu_brute_batch (host buffers, synchronous, {batches[0][1] - batches[0][0]} "
        f"queries per call): {qps:,.0f} QPS; {per_call:.1f} us per call, of which "
        f"{dev_us:.1f} us device pipeline")
    return qps, {"us_per_call": round(per_call, 1), "device_pipeline_us": round(dev_us, 1),
                 "host_and_transfer_us": round(per_call - dev_us, 1),
                 "note": "per call: predicate compile and packing on the host, one H2D copy, "
                         "the device pipeline, D2H of results, synchronisation"}


GRAPH_EFS = (40, 80, 160, 400)


def graph_path(args):
    """Filtered HNSW beam search: ef sweep, recall@10 against exact truth."""
    nq = min(args.nq, 4096)
    classes = synth.FILTER_CLASSES
    t = time.perf_counter()
    d = synth.gmm128(args.n, nq)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    graph = hnsw.build(d["base"], M=16, ef_construction=200, threads=args.threads)
    build_s = time.perf_counter() - t
    ent = hnsw.entries(graph, d["base"], d["queries"], threads=args.threads)
    log(f"data + graph M=16 efC=200: {build_s:.1f}s ({args.threads} threads)")
    g = E.GpuIndex(d["base"], cols, rows, l0=graph["l0"], precision=L.FV_GPU_PREC_FP32, device=0)
    # exact truth: the library's brute force (bit-identical to the oracle)
    truth = []
    for s in range(0, nq, 1024):
        e = min(nq, s + 1024)
        preds = [synth.filter_pred(classes[i % 6], d["qcent"][i]) for i in range(s, e)]
        b = E.concat([E.serialize(p, args.n, post) for p in preds], args.n)
        out, cnt, _ = g.brute(d["queries"][s:e], b, args.k)
        truth += [out[j, : cnt[j]]["node"].copy() for j in range(e - s)]
    sweep, best = [], None
    # frontier: "exact" = LDS frontier of 1024 with the spill store (no entry
    # the CPU cursor would pop is lost); "bounded" = 1024 entries, overflow
    # dropped as in the reference GPU kernel
    for bsz, frontier in ((64, "exact"), (1024, "exact"), (1024, "bounded")):
        fcap = 1024 if frontier == "exact" else -1024
        batches = build_batches(d, post, args.n, nq, bsz, args.k, ent)
        staged = [g.stage(d["queries"][s:e], b, args.k) for s, e, b in batches]
        for ef in GRAPH_EFS:
            mv = max(200_000, ef * 400)
            for bh in staged:
                g.run_graph(bh, ef, mv, 8, fcap)
            g.sync()
            t0 = time.perf_counter()
            for bh in staged:
                g.run_graph(bh, ef, mv, 8, fcap)
            g.sync()
            dt = time.perf_counter() - t0
            per = [[] for _ in classes]
            for (s, e, _), bh in zip(batches, staged):
                o, c, f, rc = g.fetch(bh, e - s, ef)
                for j in range(e - s):
                    tr = truth[s + j]
                    got = o[j, : min(int(c[j]), args.k)]["node"]
                    per[(s + j) % 6].append(O_recall(got, tr))
            rec = float(np.mean([x for p in per for x in p]))
            row = {"batch": bsz, "frontier": frontier, "ef": ef, "qps": round(nq / dt, 1),
                   "recall_at_10": round(rec, 4),
                   "recall_per_class": {classes[i]: round(float(np.mean(per[i])), 4)
                                        for i in range(6)}}
            sweep.append(row)
            log(f"batch {bsz} {frontier} ef {ef}: {nq / dt:,.0f} QPS, recall@10 {rec:.4f} "
                + " ".join(f"{classes[i]}={np.mean(per[i]):.3f}" for i in range(6)))
            worst = min(float(np.mean(p)) for p in per)
            row["worst_class_recall"] = round(worst, 4)
            if worst >= 0.9 and (best is None or row["qps"] > best["qps"]):
                best = row
        for bh in staged:
            g.free_batch(bh)
    # roofline of the launch shape of the reported row (its ef and frontier)
    b0 = build_batches(d, post, args.n, min(nq, 1024), 1024, args.k, ent)[0][2]
    ref_ef = best["ef"] if best else args.graph_ef
    ref_fcap = -1024 if best and best["frontier"] == "bounded" else 1024
    roof = graph_launch(g, d["queries"][:min(nq, 1024)], b0, ref_ef,
                        max(200_000, 400 * ref_ef), 128, fcap=ref_fcap)
    log(f"graph kernel roofline: {roof}")
    cpu = None if args.no_cpu else graph_cpu_baseline(d, cols, rows, vals, graph, ent, truth, args)
    g.close()
    res = {
        "metric": "filtered HNSW beam-search QPS at recall@10>=0.9 in every filter class (best "
                  "ef and frontier of the sweep), 1 GPU",
        "value": best["qps"] if best else None, "unit": "queries/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": round(nq / best["qps"] * 1e3, 3) if best else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded GMM, see paves_amd/synth.py)",
        "config": {"workload": f"filtered HNSW: {args.n} x 128 GMM, graph M=16 (M0=32) efC=200, "
                               f"{nq} queries, six filter classes via predicate bitmaps, "
                               "batches of 64 and 1024, ef in {40,80,160,400}, exact and "
                               "bounded frontiers",
                   "n": args.n, "dim": 128, "queries": nq, "graph_build_s": round(build_s, 1),
                   "best": best},
        "sweep": sweep,
        "roofline": {"bound": "hbm", "achieved": roof["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": roof["frac"], "traffic": None, "detail": roof},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def graph_recall(g, d, post, graph, truth, ef, args, nq):
    """recall@10 per class of 1024-query batches at ef (exact frontier)."""
    classes = synth.FILTER_CLASSES
    ent = hnsw.entries(graph, d["base"], d["queries"][:nq], threads=args.threads)
    batches = build_batches(d, post, args.n, nq, 1024, args.k, ent)
    per = [[] for _ in classes]
    mv = max(200_000, ef * 400)
    for s, e, b in batches:
        out, cnt, _ = g.graph(d["queries"][s:e], b, ef, mv)
        for j in range(e - s):
            got = out[j, : min(int(cnt[j]), args.k)]["node"]
            per[(s + j) % 6].append(O_recall(got, truth[s + j]))
    return {classes[i]: round(float(np.mean(per[i])), 4) for i in range(6)}, \
        round(float(np.mean([x for p in per for x in p])), 4)


def reference_build_baseline(x, args, M=16, efc=200):
    """The reference builder (engine.cpp fv_builder_finish, oracle/_ref) on
    the host threads: HNSW construction plus the index-file write, vectors
    only (no columns). Rows are added before the clock starts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import ctypes as C
    try:
        ref = O.RefEngine()
    except RuntimeError as e:
        log(f"cpu baseline skipped: {e}")
        return None
    threads = min(16, os.cpu_count() or 1)
    n, dim = x.shape
    tmpd = tempfile.mkdtemp(prefix="paves_bench_")
    path = os.path.join(tmpd, "ref.idx")
    try:
        att = np.zeros(1, np.int32)
        b = ref.L.fv_builder_create(dim, 0, att.ctypes.data, n)
        sc = np.zeros(1, np.float64)
        nul = np.zeros(1, np.uint8)
        xv = np.ascontiguousarray(x, np.float32)
        for i in range(n):
            ref.L.fv_builder_add_row(b, xv[i].ctypes.data, i, sc.ctypes.data, nul.ctypes.data)
        eb = C.create_string_buffer(256)
        t0 = time.perf_counter()
        rc = ref.L.fv_builder_finish(b, path.encode(), 1, M, efc, threads, eb, 256)
        el = time.perf_counter() - t0
        ref.L.fv_builder_destroy(b)
        if rc != 0:
            log(f"cpu baseline failed: {eb.value.decode()}")
            return None
        log(f"reference builder: {el:.1f}s on {threads} threads")
        return {"value": round(el, 2), "unit": "s", "cores": threads, "kind": "reference",
                **host_info(),
                "sample": f"the whole workload: reference fv_builder_finish (engine.cpp HNSW "
                          f"build, M={M}, efC={efc}, {threads} threads) over the same {n} x {dim} "
                          "rows, including its index-file write"}
    finally:
        try:
            os.unlink(path)
        except OSError:
            pass
        try:
            os.rmdir(tmpd)
        except OSError:
            pass


def build_path(args):
    """HNSW construction of the graph workload's graph (1M x 128, M = 16,
    efC = 200): host phase (upper levels, entry points) plus level 0 on the
    GPU, timed end to end from host vectors to the graph on the device and
    in host memory. Compared with the library's host builder (the
    reference builder's semantics) on the same threads and with the compiled
    reference builder (cpu_baseline); both graphs are searched for recall."""
    nq = min(args.nq, 4096)
    M, efc = 16, 200
    d = synth.gmm128(args.n, nq)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    x = d["base"]
    g = E.GpuIndex(x, cols, rows, precision=L.FV_GPU_PREC_FP32, device=0)
    for _ in range(args.warmup):
        hnsw.build_gpu(g, x, M, efc, threads=args.threads, keep=False)
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        dev = hnsw.build_gpu(g, x, M, efc, threads=args.threads, keep=True)
    el = (time.perf_counter() - t0) / steps
    st = dev["stats"]
    log(f"device build: {el:.2f}s per build ({st})")
    t1 = time.perf_counter()
    host = hnsw.build(x, M, efc, threads=args.threads)
    host_s = time.perf_counter() - t1
    log(f"host builder: {host_s:.2f}s on {args.threads} threads")
    # exact truth from the library's brute force (bit-identical to the oracle)
    classes = synth.FILTER_CLASSES
    truth = []
    for s in range(0, nq, 1024):
        e = min(nq, s + 1024)
        preds = [synth.filter_pred(classes[i % 6], d["qcent"][i]) for i in range(s, e)]
        b = E.concat([E.serialize(p, args.n, post) for p in preds], args.n)
        out, cnt, _ = g.brute(d["queries"][s:e], b, args.k)
        truth += [out[j, : cnt[j]]["node"].copy() for j in range(e - s)]
    recall = {}
    for ef in (40, 80):
        pd, md = graph_recall(g, d, post, dev, truth, ef, args, nq)
        recall[f"device_ef{ef}"] = {"mean": md, "per_class": pd}
    hg = E.GpuIndex(x, cols, rows, l0=host["l0"], precision=L.FV_GPU_PREC_FP32, device=0)
    try:
        for ef in (40, 80):
            ph, mh = graph_recall(hg, d, post, host, truth, ef, args, nq)
            recall[f"host_ef{ef}"] = {"mean": mh, "per_class": ph}
    finally:
        hg.close()
    log("recall@10: " + ", ".join(f"{k} {v['mean']}" for k, v in recall.items()))
    # searches of the build: one FP32 row per visited node, one neighbour
    # line per expanded head, over their event time
    alg = float(st["visited"]) * 128 * 4 + float(st["heads_expanded"]) * GRAPH_NBR * 4
    gbs = alg / max(st["search_s"], 1e-9) / 1e9
    roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "detail": {"kernel": GRAPH_KERNEL + " (efConstruction searches of the build)",
                       "nodes_visited": st["visited"], "heads_expanded": st["heads_expanded"],
                       "alg_bytes": int(alg), "search_s": round(st["search_s"], 4),
                       "prune_s": round(st["prune_s"], 4), "link_s": round(st["link_s"], 4),
                       "ns_per_visited_node": round(st["search_s"] * 1e9 /
                                                    max(1, st["visited"]), 4)}}
    g.close()
    cpu = None if args.no_cpu else reference_build_baseline(x, args, M, efc)
    res = {
        "metric": "HNSW construction time, 1M x 128, M=16, efC=200 (host vectors to graph), 1 GPU",
        "value": round(el, 3), "unit": "s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(el * 1e3, 1), "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded GMM, see paves_amd/synth.py)",
        "config": {"workload": f"HNSW build: {args.n} x 128 GMM, M=16 (M0=32), efC=200; host "
                               f"phase on {args.threads} threads, level 0 on the GPU in "
                               f"{st['batches']} batches; recall on {nq} queries, six filter "
                               "classes, batches of 1024",
                   "n": args.n, "dim": 128, "host_phase_s": round(st["host_s"], 3),
                   "device_phase_s": round(st["device_s"], 3), "batches": st["batches"],
                   "host_builder_s": round(host_s, 3), "host_threads": args.threads,
                   "recall_at_10": recall},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def O_recall(got, truth):
    tr = set(int(x) for x in truth)
    return len(tr & set(int(x) for x in got)) / len(tr) if tr else 1.0


def graph_cpu_baseline(d, cols, rows, vals, graph, ent, truth, args, ef=400):
    """Reference HnswCursor (engine.cpp, oracle/_ref) at the given ef on a
    bounded sample, one query per host thread, reading a reference-format
    index file carrying the same graph. Each thread opens the index itself:
    the reference cursor borrows a per-index visited array (not shared
    between threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import threading
    try:
        ref = O.RefEngine()
    except RuntimeError as e:
        log(f"cpu baseline skipped: {e}")
        return None
    threads = min(16, os.cpu_count() or 1)
    tmpd = tempfile.mkdtemp(prefix="paves_bench_")
    path = os.path.join(tmpd, "graph.idx")
    classes = synth.FILTER_CLASSES
    try:
        indexfile.write(path, d["base"], cols, M=16, efc=200, levels=graph["levels"],
                        l0=graph["l0"], upper_offs=graph["upper_offs"], upper=graph["upper"],
                        entry_node=graph["entry_node"], max_level=graph["max_level"],
                        post_rows=rows, post_vals=vals)
        idxs = [ref.open(path) for _ in range(threads)]
        recs = [[] for _ in range(threads)]
        t0 = time.perf_counter()

        def worker(w):
            idx = idxs[w]
            for qi in range(w, len(truth), threads):
                if time.perf_counter() - t0 > args.cpu_seconds:
                    return
                p = synth.filter_pred(classes[qi % 6], d["qcent"][qi])
                ph, keep = ref.compile(idx, p)
                try:
                    hits = ref.hnsw(idx, d["queries"][qi], ph, ef, args.k)
                finally:
                    if ph:
                        ref.L.fv_pred_free(ph)
                recs[w].append(O_recall([h[0] for h in hits][: args.k], truth[qi]))

        ths = [threading.Thread(target=worker, args=(w,)) for w in range(threads)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        el = time.perf_counter() - t0
        for idx in idxs:
            ref.close(idx)
        recs = [x for r in recs for x in r]
        done = len(recs)
        log(f"cpu reference HNSW ef={ef}: {done} queries in {el:.1f}s, "
            f"recall {np.mean(recs):.4f}")
        return {"value": round(done / el, 2), "unit": "queries/s", "cores": threads,
                "kind": "reference", "recall_at_10": round(float(np.mean(recs)), 4),
                "ef": ef, **host_info(),
                "sample": f"first {done} queries (classes interleaved), reference HnswCursor "
                          f"(engine.cpp) at ef={ef}, one query per thread on {threads} threads, "
                          "same graph read from a reference-format index file"}
    finally:
        try:
            os.unlink(path)
            os.rmdir(tmpd)
        except OSError:
            pass


def uneven_batch(d, n, nq, seed, post):
    """k ~ U{1,10,100}, selectivity 10^U(-4,0) as price < 1000 s."""
    rng = np.random.default_rng(seed)
    kq = rng.choice(np.array([1, 10, 100], np.uint32), nq)
    sel = 10.0 ** rng.uniform(-4.0, 0.0, nq)
    preds = [E.cmp(synth.COL_PRICE, E.CMP_LT, float(1000.0 * s)) for s in sel]
    return kq, sel, E.concat([E.serialize(p, n, post) for p in preds], n)


def uneven_path(args):
    dim = args.dim
    n = args.n if args.n_set else (1_000_000 if dim == 128 else 2_000_000)
    nq = 4096
    ns = args.stream_batches
    t = time.perf_counter()
    d = synth.gmm128(n, nq * (1 + ns)) if dim == 128 else synth.gmm768(n, nq * (1 + ns))
    allq = d["queries"]
    d["queries"] = allq[:nq]
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    kq, sel, b = uneven_batch(d, n, nq, 20240617 + dim, post)
    stream = []
    for i in range(ns):
        kqi, _, bi = uneven_batch(d, n, nq, 20240618 + dim + i, post)
        kqi[0] = 100                  # every batch's largest k: the staged batch's
        stream.append((allq[(i + 1) * nq:(i + 2) * nq], kqi, bi))
    log(f"data: {n}x{dim} + {nq} uneven queries (+ {ns} more batches) in "
        f"{time.perf_counter() - t:.1f}s")
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16, device=0)
    kq[0] = 100
    bh = g.stage_ex(d["queries"], b, kq)
    base = L.BruteOpts(precision=L.FV_GPU_PREC_BF16, route=1, inline_pred=1)
    t = time.perf_counter()
    best, best_us, ntried = g.tune(bh, base, reps=3)
    tune_s = time.perf_counter() - t
    shape = {"qtw": best.qtw, "waves": best.waves, "wg_per_cu": best.wg_per_cu,
             "order": {1: "chunked", 2: "sweep", 3: "query-major"}.get(best.order, "auto"),
             "sample_stride": best.sample_stride}
    log(f"autotuner: {ntried} shapes in {tune_s:.1f}s, best {shape} at {best_us:.0f} us")
    opts = L.BruteOpts(precision=best.precision, route=best.route, inline_pred=best.inline_pred,
                       wg_per_cu=best.wg_per_cu, qtw=best.qtw, order=best.order,
                       waves=best.waves, sample_stride=best.sample_stride, graph=1)
    for _ in range(max(2, args.warmup)):      # eager run, then capture
        g.run_brute(bh, opts)
    g.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.run_brute(bh, opts)
    g.sync()
    dt = (time.perf_counter() - t0) / args.steps
    qps = nq / dt
    log(f"hipGraph replay: {dt * 1e3:.3f} ms per 4096-query launch -> {qps:,.0f} QPS")
    out, cnt, flag, rc = g.fetch(bh, nq, 100)

    # a stream of distinct batches through the one capture: each is restaged
    # into the batch (host compile and packing, one H2D copy), replayed and
    # fetched; the same opts keep the capture
    restream = None
    if ns > 0:
        sb = g.stage_ex(d["queries"], b, kq)
        for _ in range(2):
            g.run_brute(sb, opts)
        g.sync()
        kept, t0 = [], time.perf_counter()
        for qs_i, kq_i, b_i in stream:
            kept.append(g.restage(sb, qs_i, b_i, 100, kq_i))
            g.run_brute(sb, opts)
            g.fetch(sb, nq, 100)
        dts = time.perf_counter() - t0
        g.free_batch(sb)
        restream = {"batches": ns, "queries_per_batch": nq, "qps": round(ns * nq / dts, 1),
                    "ms_per_batch": round(dts / ns * 1e3, 3), "capture_kept": bool(all(kept)),
                    "note": "per batch: fv_gpu_batch_restage (predicate compile, packing, one "
                            "H2D copy), one replay of the captured run, fetch of the results"}
        log(f"restaged stream: {ns} distinct batches, {dts / ns * 1e3:.2f} ms per batch "
            f"({ns * nq / dts:,.0f} QPS), capture kept: {all(kept)}")

    # per-kernel times of one eager run with the same shape (HIP events)
    eager = L.BruteOpts(precision=best.precision, route=best.route, inline_pred=best.inline_pred,
                        wg_per_cu=best.wg_per_cu, qtw=best.qtw, order=best.order,
                        waves=best.waves, sample_stride=best.sample_stride)
    g.set_timing(True)
    scans = []
    for _ in range(3):
        g.run_brute(bh, eager)
        g.sync()
        st = g.stats()
        scans.append((st.t_scan_us, st.t_prepass_us + st.t_merge_us, st.t_rerank_us,
                      st.t_total_us))
    g.set_timing(False)
    scan_us, pre_us, sel_us, tot_us = (float(np.mean([s[i] for s in scans])) for i in range(4))
    flops = 2.0 * n * dim * nq
    tflops = flops / (scan_us * 1e-6) / 1e12
    gy = (nq // 64 + best.qtw - 1) // best.qtw
    hbm_gbs = n * (2 * dim + 4 + 8) * gy / (scan_us * 1e-6) / 1e9
    log(f"scan {scan_us:.0f} us ({tflops:.0f} TFLOP/s; {gy} data passes, {hbm_gbs:.0f} GB/s), "
        f"pre-pass+tau {pre_us:.0f} us, select {sel_us:.0f} us, eager total {tot_us:.0f} us")

    sample_ok, cpu = None, None
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        rng = np.random.default_rng(7)
        m = 48 if dim == 128 else 12
        qs_all = rng.choice(nq, m, replace=False)
        ok = rc == 0
        t = time.perf_counter()
        for k in (1, 10, 100):
            qs = qs_all[kq[qs_all] == k]
            if len(qs) == 0:
                continue
            oh, oc, ot, _ = O.brute_batch(d["base"], cols, d["queries"][qs], b.qd[qs], b.nodes,
                                          b.children, b.invals, k)
            for j, q in enumerate(qs):
                c = int(cnt[q])
                ok &= c == int(oc[j]) and int(flag[q]) == int(ot[j])
                ok &= bool(np.array_equal(out[q, :c], oh[j, :c]))
        cpu_s = time.perf_counter() - t
        sample_ok = bool(ok)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(m / cpu_s, 2), "unit": "queries/s", "cores": threads,
               "kind": "port", "sample": f"{m} of the 4096 queries, each at its own k, "
                                         f"oracle full scan (OpenMP over queries)"}
        log(f"oracle check ({m} queries): {sample_ok}; oracle {m / cpu_s:.2f} QPS")
    g.free_batch(bh)
    g.close()
    res = {
        "metric": f"uneven-batch filtered top-k QPS ({nq} queries, k in {{1,10,100}}, "
                  "selectivity 10^U(-4,0)), 1 GPU",
        "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": max(2, args.warmup), "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (seeded GMM, see paves_amd/synth.py)",
        "config": {"workload": f"uneven batch: {n} x {dim} GMM, {nq} queries, k in {{1,10,100}}, "
                               "selectivity log-uniform in [1e-4,1], one hipGraph replay",
                   "n": n, "dim": dim, "queries": nq, "autotuned": shape,
                   "shapes_tried": ntried, "tune_s": round(tune_s, 2)},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 4), "traffic": None,
                     "kernel": ("k_scan_bf16_qr" if best.qtw == 8 else
                                "k_scan_bf16_mq" if best.qtw > 1 else "k_scan_bf16"),
                     "avg_launch_us": round(scan_us, 1), "flops_per_launch": flops,
                     "data_passes": gy, "hbm_GBs": round(hbm_gbs, 1)},
        "prepass_tau_us": round(pre_us, 1), "select_us": round(sel_us, 1),
        "restaged_stream": restream,
        "oracle_sample_ok": sample_ok, "fetch_rc": rc,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def sq8_path(d, cols, rows, batches, res_bf, args, batchesL=None):
    """The same micro-batches on an index uploaded with FV_GPU_PREC_SQ8:
    integer data spanning <= 255 values scan as int8 (x - centre) with
    v_mfma_i32_32x32x32_i8 — exact, half the bytes of BF16. Throughput,
    results against the BF16 run, and the INT8 scan kernel's rate
    (back-to-back launches, as for the BF16 roofline)."""
    g8 = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_SQ8, device=0)
    staged8 = [g8.stage(d["queries"][s:e], b, args.k) for s, e, b in batches]
    sq = L.BruteOpts(precision=L.FV_GPU_PREC_SQ8, route=0)
    run_steps(g8, staged8, sq, max(1, args.warmup))
    dt8 = run_steps(g8, staged8, sq, args.steps)
    qps8 = args.steps * args.nq / dt8
    res8 = [g8.fetch(bh, e - s, args.k) for bh, (s, e, _) in zip(staged8, batches)]
    same = all(np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and
               all(np.array_equal(a[0][q, : a[1][q]], b[0][q, : b[1][q]])
                   for q in range(len(a[1]))) for a, b in zip(res8, res_bf))
    scan_us, nprog, used = [], [], 0
    for bh in staged8[:40]:
        g8.run_brute(bh, sq)
        g8.sync()
        st = g8.stats()
        used += st.precision == L.FV_GPU_PREC_SQ8
        if st.route == 1 and st.precision == L.FV_GPU_PREC_SQ8:
            nprog.append(st.nprog)
            scan_us.append(g8.time_scan(bh, sq, 10))
    for bh in staged8:
        g8.free_batch(bh)
    # batches of 1024 on the SQ8 index: the library's default (the exact
    # query-resident BF16 scan) and the INT8 code scan asked for
    batched = None
    if batchesL:
        stagedL8 = [g8.stage(d["queries"][s:e], b, args.k) for s, e, b in batchesL]
        batched = {}
        for tag, o in (("default", L.BruteOpts(route=1)),
                       ("int8_asked", L.BruteOpts(route=1, precision=L.FV_GPU_PREC_SQ8))):
            run_steps(g8, stagedL8, o, max(1, args.warmup))
            dtb = run_steps(g8, stagedL8, o, max(2, args.steps // 4))
            g8.run_brute(stagedL8[0], o)
            g8.sync()
            batched[tag] = {"qps": round(max(2, args.steps // 4) * args.nq / dtb, 1),
                            "precision": int(g8.stats().precision)}
        log(f"sq8 index, batches of {args.learned_batch}: default (query-resident BF16, "
            f"exact) {batched['default']['qps']:,.0f} QPS; INT8 scan asked "
            f"{batched['int8_asked']['qps']:,.0f} QPS")
        for bh in stagedL8:
            g8.free_batch(bh)
    g8.close()
    if not scan_us:
        log("sq8: the INT8 scan was not taken")
        return {"qps": round(qps8, 1), "int8_scan_taken": False, "exact_vs_bf16": same}
    us = float(np.mean(scan_us))
    alg = args.n * (128 + 4) + float(np.mean(nprog)) * args.n / 8.0
    gbs = alg / (us * 1e-6) / 1e9
    log(f"sq8 (INT8 scan): {qps8:,.0f} QPS, results equal BF16: {same}; scan kernel "
        f"{us:.1f} us -> {gbs:.0f} GB/s")
    return {"qps": round(qps8, 1), "int8_scan_taken": True, "batches_int8": used,
            "batches_of_1024": batched,
            "exact_vs_bf16": same, "kernel": "k_scan_i8<false, 8, 2, false>",
            "avg_launch_us": round(us, 2), "alg_bytes_per_launch": int(alg),
            "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def quant_leg(d, cols, rows, batches, prec, args, nq, truth=None):
    """One reduced-precision index over the float data: throughput of its
    scan (SQ8 int8 codes or FP8 E4M3 codes, top 4 k re-ranked exactly),
    recall@k against the exact BF16 results (computed here when not given),
    and the code scan kernel's rate (10 back-to-back launches per batch)."""
    g = E.GpuIndex(d["base"], cols, rows, precision=prec, device=0)
    staged = [g.stage(d["queries"][s:e], b, args.k) for s, e, b in batches]
    res = {}
    legs = [("q", prec)] + ([] if truth is not None else [("bf16", L.FV_GPU_PREC_BF16)])
    for name, pr in legs:
        o = L.BruteOpts(precision=pr, route=0)
        run_steps(g, staged, o, max(1, args.warmup))
        dt = run_steps(g, staged, o, args.steps)
        res[name] = (args.steps * nq / dt,
                     [g.fetch(bh, e - s, args.k) for bh, (s, e, _) in zip(staged, batches)])
    if truth is None:
        truth = res["bf16"]
    rec = []
    for (o8, c8, _, _), (ob, cb, _, _) in zip(res["q"][1], truth[1]):
        for q in range(len(c8)):
            tr = set(int(x) for x in ob[q, : cb[q]]["node"])
            got = set(int(x) for x in o8[q, : c8[q]]["node"])
            rec.append(len(tr & got) / len(tr) if tr else 1.0)
    sq = L.BruteOpts(precision=prec, route=0)
    scan_us, nprog = [], []
    for bh in staged[:24]:
        g.run_brute(bh, sq)
        g.sync()
        st = g.stats()
        if st.route == 1 and st.precision == prec:
            nprog.append(st.nprog)
            scan_us.append(g.time_scan(bh, sq, 10))
    for bh in staged:
        g.free_batch(bh)
    g.close()
    us = float(np.mean(scan_us)) if scan_us else None
    alg = args.n * (128 + 4) + (float(np.mean(nprog)) if nprog else 0.0) * args.n / 8.0
    gbs = alg / (us * 1e-6) / 1e9 if us else None
    kern = "k_scan_i8<false, 8, 2, false, true>" if prec == L.FV_GPU_PREC_FP8 else \
        "k_scan_i8<false, 8, 2, false>"
    return {"qps": round(res["q"][0], 1), "recall_at_10": round(float(np.mean(rec)), 4),
            "overfetch": 4, "kernel": kern,
            "avg_launch_us": round(us, 2) if us else None, "alg_bytes_per_launch": int(alg),
            "achieved_GBs": round(gbs, 1) if gbs else None,
            "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}, truth, \
        (res["bf16"][0] if "bf16" in res else None)


def sq8_quant_path(args, nq=4096):
    """SQ8 and FP8 indexes over non-integer data: a 1M x 128 unit-norm
    mixture (the 768-d generator at 128 dims, FP32 values), nq queries of
    the six filter classes in micro-batches of 64. SQ8 keeps int8 codes of
    the midrange-centred rows (INT8 MFMA scan), FP8 E4M3 codes of the same
    (FP8 MFMA scan); both rank by code distance and re-rank the top 4 k
    exactly. Throughput and recall@k against the exact BF16 path on the same
    data, and each code scan kernel's rate."""
    t = time.perf_counter()
    d = synth.gmm768(args.n, nq, dim=128, ncent=1024, seed=synth.BASE_SEED + 128)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    batches = build_batches(d, post, args.n, nq, args.batch, args.k)
    log(f"quantized indexes: {args.n}x128 float data in {time.perf_counter() - t:.1f}s")
    sq8, truth, bf_qps = quant_leg(d, cols, rows, batches, L.FV_GPU_PREC_SQ8, args, nq)
    fp8, _, _ = quant_leg(d, cols, rows, batches, L.FV_GPU_PREC_FP8, args, nq, truth)
    log(f"sq8 quantized: {sq8['qps']:,.0f} QPS (BF16 exact on the same data {bf_qps:,.0f}), "
        f"recall@{args.k} {sq8['recall_at_10']}; scan {sq8['avg_launch_us']} us -> "
        f"{sq8['achieved_GBs']} GB/s")
    log(f"fp8: {fp8['qps']:,.0f} QPS, recall@{args.k} {fp8['recall_at_10']}; scan "
        f"{fp8['avg_launch_us']} us -> {fp8['achieved_GBs']} GB/s")
    sq8.update({"bf16_exact_qps": round(bf_qps, 1), "queries": nq,
                "data": f"{args.n} x 128 unit-norm Gaussian mixture, FP32 values"})
    fp8.update({"queries": nq, "data": "the same"})
    return sq8, fp8


def learned_path(g, graph, d, post, staged, batches, truth, args, setup):
    """Train the router on queries drawn apart from the evaluated ones,
    then time the routed path (mode 0) over the staged micro-batches; also
    the fixed-ef fallback (graph, ef = 400, FP32 distances) on a sample."""
    classes = synth.FILTER_CLASSES
    t = time.perf_counter()
    tq, tc = synth.more_queries(d, args.train_queries, seed=synth.BASE_SEED + 1)
    tsers = [E.serialize(synth.filter_pred(classes[i % 6], tc[i]), args.n, post)
             for i in range(args.train_queries)]
    setup["train_queries_gen_s"] = round(time.perf_counter() - t, 2)
    model, info = learned.train_router(g, graph, d["base"], tq, tsers, k=args.k, log=log,
                                       micro=args.learned_batch)
    setup["router"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in info.items()}
    # safety factor on the predicted ef, from training queries held apart
    nval = min(4096, args.train_queries)
    vq, vs = tq[-nval:], tsers[-nval:]
    vcls = np.arange(args.train_queries - nval, args.train_queries) % 6
    vent = hnsw.entries(graph, d["base"], vq, threads=args.threads)
    vtruth = learned.exact_truth(g, vq, vs, args.k)
    # route threshold and ef factor: fastest held-out setting whose recall
    # reaches the target in every filter class
    bias, scale, btable = learned.calibrate_bias(g, model, vq, vs, vent, vtruth, k=args.k,
                                                 micro=args.learned_batch, reps=5, groups=vcls,
                                                 scales=(1.0, 1.5, 2.0, 3.0))
    opts = L.LearnedOpts(k=args.k, mode=0, ef_scale=scale, route_bias=bias, max_visited=50_000)
    log(f"router: ef scale {scale}, route bias {bias:.3g}; held-out calibration (graph share, "
        f"ef scale, QPS, worst-class recall): " +
        ", ".join(f"({t[0]:.2f}, {t[1]}, {t[3]:,.0f}, {t[4]:.3f})" for t in btable))

    def step():
        for bh in staged:
            g.run_learned(bh, model, opts)

    for _ in range(args.warmup):
        step()
    g.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    g.sync()
    dt = time.perf_counter() - t0
    qps = args.steps * args.nq / dt
    rec, ng, efs, per = [], 0, [], [[] for _ in classes]
    for (s, e, _), bh in zip(batches, staged):
        o, c, f, rc = g.fetch(bh, e - s, args.k)
        r, ef = g.routes(bh, e - s)
        for j in range(e - s):
            tr = set(int(x) for x in truth[s + j])
            got = set(int(x) for x in o[j, : c[j]]["node"])
            rr = len(tr & got) / len(tr) if tr else 1.0
            rec.append(rr)
            per[(s + j) % 6].append(rr)
            if r[j]:
                ng += 1
                efs.append(int(ef[j]))
    log(f"learned route+ef: {qps:,.0f} QPS, recall@10 {np.mean(rec):.4f}, "
        f"graph-routed {ng / args.nq:.3f}")
    # fixed-ef fallback on the graph (plain FP32 distances), bounded sample
    nfx = min(args.nq, 1024)
    fq = d["queries"][:nfx]
    fs = [E.serialize(synth.filter_pred(classes[i % 6], d["qcent"][i]), args.n, post)
          for i in range(nfx)]
    fent = np.concatenate([b.qd["entry"] for _, _, b in batches])[:nfx].astype(np.uint32)
    fx = learned.run_mode(g, model, fq, fs, fent, truth[:nfx],
                          L.LearnedOpts(k=args.k, mode=2, ef_fixed=400), micro=args.batch)
    log(f"fixed ef=400 graph (FP32): {fx[0]:,.0f} QPS, recall {fx[1]:.4f} ({nfx} queries)")
    # the beam-search kernel on its roofline: one launch over the first
    # batch of the workload (six classes) at ef = graph_ef
    s0, e0, b0 = batches[0]
    if e0 - s0 > 1024:
        # the synchronous graph entry point takes at most FV_GPU_MAX_BATCH
        e0 = s0 + 1024
        b0 = E.concat([E.serialize(synth.filter_pred(classes[i % 6], d["qcent"][i]), args.n, post)
                       for i in range(s0, e0)], args.n,
                      entries=np.asarray(batches[0][2].qd["entry"])[:1024])
    mvg = max(200_000, 400 * args.graph_ef)
    gk = graph_launch(g, d["queries"][s0:e0], b0, args.graph_ef, mvg, 128)
    gk["traffic"] = None if args.no_traffic else graph_traffic(args, graph,
                                                              np.asarray(b0.qd["entry"]))
    gk["note"] = ("the six filter classes: the launch lasts as long as its slowest search "
                  "(selective classes visit up to max_visited nodes)")
    bu = E.concat([E.serialize(None, args.n)] * (e0 - s0), args.n, entries=b0.qd["entry"])
    gk["unfiltered"] = graph_launch(g, d["queries"][s0:e0], bu, args.graph_ef, mvg, 128)
    for tag, m in (("six classes", gk), ("unfiltered", gk["unfiltered"])):
        log(f"graph kernel, {tag} (ef {args.graph_ef}, {e0 - s0} queries): "
            f"{m['avg_launch_us']:.0f} us, {m['nodes_visited']:,} nodes visited -> "
            f"{m['achieved_GBs']:.0f} GB/s ({m['ns_per_visited_node']} ns/node)")
    log(f"graph kernel traffic (six classes): {gk['traffic']}")
    model.close()
    return {"qps": qps, "ms_per_step": dt / args.steps * 1e3, "recall_at_10": float(np.mean(rec)),
            "recall_per_class": {classes[i]: round(float(np.mean(per[i])), 4) for i in range(6)},
            "graph_routed_frac": ng / args.nq, "mean_ef_graph_routed": float(np.mean(efs)) if efs else 0.0,
            "ef_scale": scale, "route_bias": bias if np.isfinite(bias) else None,
            "route_disabled": not np.isfinite(bias),
            "bias_calibration": [{"graph_share_target": t[0], "ef_scale": t[1],
                                  "qps": round(t[3], 1), "worst_class_recall_at_10": round(t[4], 4),
                                  "graph_routed_frac": round(t[5], 4)} for t in btable],
            "fixed_ef400_graph_qps": round(fx[0], 1), "fixed_ef400_graph_recall": round(fx[1], 4),
            "fixed_ef400_sample": nfx, "cpu_hnsw_ef": 80,
            "graph_kernel": gk,
            "batch": args.learned_batch,
            "note": "each routed run decides every query's route and ef on the GPU, packs the "
                    "brute-routed queries into the leading query tiles (the scan runs only "
                    "those), then runs the graph search for the rest"}


def host_info():
    """nproc and the CPU model of the host the baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model,
            "sched_affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else None}


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
                "kind": "reference", **host_info(),
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
