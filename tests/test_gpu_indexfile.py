"""A reference-written PAVES index file served by the GPU path: the file is
read by paves_amd/indexfile.py (no reference code), uploaded as it lies
(vectors, attribute columns with NULLs, posting rows, level-0 lists at
stride 2M+1 converted to padded CSR on the GPU), and searched. Brute force
equals the reference BruteCursor on the same file (ids and distances, six
filter classes); the one-head graph search from the reference's own entry
points (fv_index_graph_entry) equals the reference HnswCursor; and the
reference host glue's graph batch entry (fv_gpu_exec_graph_batch,
engine.cpp:1315-1339, engine.cpp built with -DFV_GPU and linked against
this library) returns results that satisfy the contract and reach the
CPU cursor's recall."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import have_reference_build
from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import indexfile, synth

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not have_reference_build(), reason="oracle/_ref not built")]

N, NQ = 20_000, 24


@pytest.fixture(scope="module")
def world(tmp_path_factory):
    d = synth.gmm128(N, NQ, seed=91)
    cols = synth.columns(d)
    cols[0, ::41] = np.nan
    path = str(tmp_path_factory.mktemp("gidx") / "ref.idx")
    ref = O.RefEngine()
    ref.build(path, d["base"], cols, M=16, efc=100, threads=8)
    return d, cols, path, ref


def test_file_brute_and_graph_equal_reference_cursors(world):
    d, cols, path, ref = world
    f = indexfile.read(path)
    g = E.GpuIndex(f["vectors"], f["cols"], f["post_rows"], l0=f["l0"],
                   precision=L.FV_GPU_PREC_BF16, device=0)
    h = ref.open(path)
    post = E.Postings(f["post_rows"], f["post_vals"])
    try:
        ent = np.array([ref.graph_entry(h, q) for q in d["queries"]], np.uint32)
        for cls in synth.FILTER_CLASSES:
            print(f"case indexfile cls={cls}", flush=True)
            preds = [synth.filter_pred(cls, d["qcent"][i]) for i in range(NQ)]
            b = E.concat([E.serialize(p, N, post) for p in preds], N, entries=ent)
            out, cnt, trunc = g.brute(d["queries"], b, 10)
            gout, gcnt, gexh = g.graph(d["queries"], b, 32, 200000, wmax=1, fcap=1024)
            for qi in range(NQ):
                ph, keep = ref.compile(h, preds[qi])
                want = ref.brute(h, d["queries"][qi], ph, 10)
                got = [(int(x["node"]), float(x["dist"])) for x in out[qi, : cnt[qi]]]
                assert got == want, (cls, qi)
                wg = ref.hnsw(h, d["queries"][qi], ph, 32, int(gcnt[qi]))
                gg = [(int(x["node"]), float(x["dist"])) for x in gout[qi, : gcnt[qi]]]
                assert gg == wg, (cls, qi)
                if ph:
                    ref.L.fv_pred_free(ph)
    finally:
        ref.close(h)
        g.close()


def test_reference_glue_graph_batch(world):
    """engine.cpp's fv_gpu_exec_graph_batch (GpuConcat + fv_gpu_graph_batch)
    on this library, against its own CPU cursors on the same file."""
    d, cols, path, ref = world
    gpu = O.RefEngine(gpu=True)
    gpu.L.fv_engine_gpu_precision(L.FV_GPU_PREC_BF16)
    hg, hc = gpu.open(path), ref.open(path)
    ef, mv = 48, 200000
    try:
        assert gpu.L.fv_index_gpu_ensure(hg, -1) == 0
        for cls in synth.FILTER_CLASSES:
            print(f"case glue graph cls={cls}", flush=True)
            keep, bps = [], (O.FvGpuBatchPred * NQ)()
            for qi in range(NQ):
                p = synth.filter_pred(cls, d["qcent"][qi])
                ph, k = gpu.compile(hg, p)
                keep.append((ph, k))
                arrs = (np.zeros(16, dtype=O.PRED_DT), np.zeros(16, np.int32),
                        np.zeros(16, np.float64), np.zeros(64, dtype=L.SPAN_DT))
                keep.append(arrs)
                tot = C.c_uint64()
                assert gpu.L.fv_pred_serialize(hg, ph, arrs[0].ctypes.data, 16,
                                               arrs[1].ctypes.data, 16, arrs[2].ctypes.data, 16,
                                               arrs[3].ctypes.data, 64, C.byref(bps[qi]),
                                               C.byref(tot)) == 0
            ent = np.array([gpu.graph_entry(hg, q) for q in d["queries"]], np.uint32)
            qs = np.ascontiguousarray(d["queries"], np.float32)
            out = np.zeros((NQ, ef), dtype=L.HIT_DT)
            cnt = np.zeros(NQ, np.uint32)
            exh = np.zeros(NQ, np.uint8)
            rc = gpu.L.fv_gpu_exec_graph_batch(hg, NQ, qs.ctypes.data, ent.ctypes.data, bps, ef,
                                               mv, out.ctypes.data, cnt.ctypes.data,
                                               exh.ctypes.data)
            assert rc == 0
            rec_gpu, rec_cpu = [], []
            for qi in range(NQ):
                p = synth.filter_pred(cls, d["qcent"][qi])
                mask = None if p is None else E.eval_pred(p, cols)
                c = int(cnt[qi])
                ids = out[qi, :c]["node"].astype(np.int64)
                assert c <= ef and len(set(ids.tolist())) == c
                if mask is not None:
                    assert mask[ids].all()
                keys = (out[qi, :c]["dist"].view(np.uint32).astype(np.uint64) << np.uint64(32)) \
                    | ids.astype(np.uint64)
                assert np.all(keys[1:] > keys[:-1])
                for k in range(c):
                    assert out[qi, k]["dist"] == O.l2sq(d["queries"][qi], d["base"][ids[k]])
                pc, kk = ref.compile(hc, p)
                truth = [n for n, _ in ref.brute(hc, d["queries"][qi], pc, 10)]
                cpu = [n for n, _ in ref.hnsw(hc, d["queries"][qi], pc, ef, 10)]
                if pc:
                    ref.L.fv_pred_free(pc)
                rec_gpu.append(O.recall(ids[:10], truth))
                rec_cpu.append(O.recall(cpu, truth))
            assert np.mean(rec_gpu) >= min(0.9, np.mean(rec_cpu) - 0.05), (cls, rec_gpu, rec_cpu)
            for ph, k in keep[::2]:
                if ph:
                    gpu.L.fv_pred_free(ph)
    finally:
        gpu.close(hg)
        ref.close(hc)
