"""The learned router's GPU kernels (pv_learn.hip) against the FP64 oracle
(oracle/mlp_ref.py): training gradient and loss, BF16 inference, feature
vectors; and that training fits a known target."""
import numpy as np
import pytest

from paves_amd import _lib as L
from paves_amd import engine as E
from paves_amd import learned, synth

import mlp_ref as R

pytestmark = pytest.mark.gpu

DIM = 128


def _random_problem(seed, B):
    rng = np.random.default_rng(seed)
    F = R.nfeat(DIM)
    p = (rng.standard_normal(R.nparams(F)) * 0.1).astype(np.float32)
    x = (rng.standard_normal((B, F)) * 3 + 1).astype(np.float32)
    t = np.stack([rng.standard_normal(B), rng.integers(0, 2, B), rng.integers(0, 2, B)],
                 1).astype(np.float32)
    mean = (rng.standard_normal(F) * 0.5).astype(np.float32)
    std = rng.uniform(0.5, 3.0, F).astype(np.float32)
    return p, x, t, mean, std, F


@pytest.fixture(scope="module")
def model():
    rng = np.random.default_rng(3)
    m = learned.Model(DIM, rng.uniform(0, 200, (256, DIM)).astype(np.float32))
    yield m
    m.close()


@pytest.mark.parametrize("B,lam", [(37, 1.0), (256, 0.3)])
def test_gradient_matches_fp64_oracle(model, B, lam):
    p, x, t, mean, std, F = _random_problem(B, B)
    model.set_params(p, mean, std)
    g, loss = model.grad(x, t, lam)
    # the device standardises with 1/std in FP32: the oracle uses the same
    istd = (1.0 / std.astype(np.float32)).astype(np.float64)
    ref_g = R.grad(p, x, t, mean.astype(np.float64), 1.0 / istd, F, lam)
    ref_l = R.loss(p, x, t, mean.astype(np.float64), 1.0 / istd, F, lam)
    assert abs(loss - ref_l) <= 1e-4 * abs(ref_l) + 1e-6
    scale = np.abs(ref_g).max()
    assert np.all(np.abs(g - ref_g) <= 2e-4 * scale + 1e-3 * np.abs(ref_g)), \
        np.abs(g - ref_g).max() / scale


def test_gradient_finite_difference_on_device(model):
    """Central difference of the device loss along the device gradient
    equals the gradient's norm (a check that needs no oracle)."""
    B = 64
    p, x, t, mean, std, F = _random_problem(5, B)
    model.set_params(p, mean, std)
    g, _ = model.grad(x, t, 1.0)
    gn = float(np.linalg.norm(g.astype(np.float64)))
    v = (g / gn).astype(np.float32)
    h = 1e-3
    model.set_params(p + h * v, mean, std)
    _, lp = model.grad(x, t, 1.0)
    model.set_params(p - h * v, mean, std)
    _, lm = model.grad(x, t, 1.0)
    fd = (lp - lm) / (2 * h)
    assert abs(fd - gn) <= 1e-2 * gn, (fd, gn)


def test_bf16_inference_matches_oracle(model):
    p, x, t, mean, std, F = _random_problem(11, 101)
    model.set_params(p, mean, std)
    y = model.predict(x)
    istd = (1.0 / std).astype(np.float64)
    ref = R.predict_bf16(p, x, mean.astype(np.float64), 1.0 / istd, F)
    assert np.allclose(y, ref, rtol=1e-3, atol=1e-3 * np.abs(ref).max())


def test_features_match_oracle():
    n, nq = 30_000, 24
    d = synth.gmm128(n, nq, seed=13)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    cent = learned.kmeans(d["base"], k=64, iters=3, sample=5000)
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16)
    m = learned.Model(128, cent)
    bh = g.stage(d["queries"], b, 10)
    try:
        f = m.batch_features(g, bh, nq)
        cnt = [n if p is None else int(E.eval_pred(p, cols).sum()) for p in preds]
        ref = R.features(d["queries"], cnt, n, cent)
        assert np.array_equal(f[:, :128], d["queries"])
        assert np.allclose(f[:, 128], ref[:, 128], rtol=0, atol=1e-6)
        assert np.allclose(f[:, 129:], ref[:, 129:], rtol=1e-5, atol=1e-3)
    finally:
        g.free_batch(bh)
        m.close()
        g.close()


def test_training_fits_known_target(model):
    """Targets from a fixed random teacher network: training must cut the
    loss by 4x and halve the constant predictor's error on held-out
    samples."""
    rng = np.random.default_rng(21)
    F = R.nfeat(DIM)
    n = 40960
    x = rng.standard_normal((n, F)).astype(np.float32)
    teacher = (rng.standard_normal(R.nparams(F)) * 0.15)
    y, _ = R.forward(teacher, x, np.zeros(F), np.ones(F), F)
    t = np.stack([y[:, 0], (y[:, 1] > 0).astype(np.float64), np.ones(n)], 1).astype(np.float32)
    p0 = (rng.standard_normal(R.nparams(F)) * 0.05).astype(np.float32)
    model.set_params(p0)
    k = 32768
    hist = model.train(x[:k], t[:k], epochs=30, batch=1024, lr=3e-3, lam=0.5, seed=4)
    assert hist[-1] < hist[0] / 4, hist
    pred = model.predict(x[k:])
    mse = float(np.mean((pred[:, 0] - t[k:, 0]) ** 2))
    const = float(np.var(t[k:, 0]))
    assert mse < 0.5 * const, (mse, const)
    acc = float(np.mean((pred[:, 1] > 0) == (t[k:, 1] > 0.5)))
    assert acc > 0.8, acc


def test_routed_run_decisions_match_model():
    """The routed run's fused decision kernel (bitmap-pass counts, features,
    BF16 MLP) gives the outputs the separate feature and predict kernels
    give, and route/ef follow them; brute-routed queries get the exact
    brute-force answer, graph-routed ones valid graph hits."""
    n, nq = 30_000, 40
    d = synth.gmm128(n, nq, seed=17)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    from paves_amd import hnsw
    gr = hnsw.build(d["base"], M=8, ef_construction=48, threads=8)
    ent = hnsw.entries(gr, d["base"], d["queries"])
    b = E.concat([E.serialize(p, n, post) for p in preds], n, entries=ent)
    g = E.GpuIndex(d["base"], cols, rows, l0=gr["l0"], precision=L.FV_GPU_PREC_BF16)
    rng = np.random.default_rng(2)
    m = learned.Model(128, learned.kmeans(d["base"], k=64, iters=2, sample=4000))
    F = R.nfeat(128)
    p = (rng.standard_normal(R.nparams(F)) * 0.1).astype(np.float32)
    bh = g.stage(d["queries"], b, 10)
    try:
        f = m.batch_features(g, bh, nq)
        m.set_params(p, f.mean(0), f.std(0) + 1e-3)
        want = m.predict(f)
        # bias 0 on a random model: a mix of routes
        for rep in range(2):    # the second run checks the count reset
            g.run_learned(bh, m, L.LearnedOpts(k=10, mode=0, ef_min=10, ef_max=64))
            g.sync()
            route, ef, y = g.routes(bh, nq, with_y=True)
            assert np.array_equal(y, want), rep
            exp_ef = np.clip(np.ceil(np.exp(np.minimum(want[:, 0], 20.0))), 10, 64)
            assert np.array_equal(ef, exp_ef.astype(np.uint32))
            assert np.array_equal(route, (want[:, 1] > 0).astype(np.uint8))
        out, cnt, flag, rc = g.fetch(bh, nq, 10)
        assert rc == 0
        bo, bc, bt = g.brute(d["queries"], E.concat([E.serialize(p, n, post) for p in preds], n), 10)
        for q in range(nq):
            ids = out[q, : cnt[q]]["node"]
            if route[q] == 0:
                assert cnt[q] == bc[q] and np.array_equal(ids, bo[q, : bc[q]]["node"])
            else:
                assert len(set(ids.tolist())) == len(ids)
                if preds[q] is not None:
                    assert E.eval_pred(preds[q], cols)[ids.astype(np.int64)].all()
    finally:
        g.free_batch(bh)
        m.close()
        g.close()


@pytest.mark.skip(reason="splitting the chip between a routed run's graph searches and its "
                  "scan hung on the GPU (persistent graph grid) and crashed the host process "
                  "(CU-masked stream); cause not found, so the split is not available and "
                  "graph_cus != 0 fails the call")
def test_routed_run_persistent_graph_grid_equals_shared():
    """A routed run over many query tiles with the chip split between the
    graph searches (graph_cus persistent workgroups, one per CU, taking the
    graph-routed queries in turn) and the packed brute scan (its grid sized
    to the other CUs) returns what the shared run returns: same routes,
    hits, counts and flags; brute-routed queries equal the exact brute
    force."""
    n, nq = 40_000, 1024
    d = synth.gmm128(n, nq, seed=19)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    from paves_amd import hnsw
    g = E.GpuIndex(d["base"], cols, rows, precision=L.FV_GPU_PREC_BF16)
    gr = hnsw.build_gpu(g, d["base"], M=8, ef_construction=48, threads=8, keep=True)
    ent = hnsw.entries(gr, d["base"], d["queries"])
    b = E.concat([E.serialize(p, n, post) for p in preds], n, entries=ent)
    rng = np.random.default_rng(3)
    m = learned.Model(128, learned.kmeans(d["base"], k=64, iters=2, sample=4000))
    F = R.nfeat(128)
    p = (rng.standard_normal(R.nparams(F)) * 0.1).astype(np.float32)
    bh = g.stage(d["queries"], b, 10)
    try:
        f = m.batch_features(g, bh, nq)
        m.set_params(p, f.mean(0), f.std(0) + 1e-3)
        ref = None
        for cus in (0, 32, 64, 32):
            print(f"case routed run, persistent graph grid of {cus} CUs", flush=True)
            g.run_learned(bh, m, L.LearnedOpts(k=10, mode=0, ef_min=10, ef_max=64,
                                               max_visited=20_000, graph_cus=cus))
            g.sync()
            res = g.fetch(bh, nq, 10)
            assert res[3] == 0
            route, ef = g.routes(bh, nq)
            if ref is None:
                ref = (res, route, ef)
                assert 0 < int(route.sum()) < nq
                continue
            assert np.array_equal(route, ref[1]) and np.array_equal(ef, ref[2])
            for x, y in zip(res[:3], ref[0][:3]):
                assert np.array_equal(x, y), cus
        out, cnt, _, _ = ref[0]
        bo, bc, _ = g.brute(d["queries"], E.concat([E.serialize(q, n, post) for q in preds], n), 10)
        for q in np.flatnonzero(ref[1] == 0):
            assert cnt[q] == bc[q] and np.array_equal(out[q, : cnt[q]]["node"], bo[q, : bc[q]]["node"])
    finally:
        g.free_batch(bh)
        m.close()
        g.close()


def test_graph_cus_rejected():
    """graph_cus != 0 (a split of the chip) fails the call before any
    device work, with a message; 0 runs."""
    n, nq = 5_000, 128
    d = synth.gmm128(n, nq, seed=23)
    g = E.GpuIndex(d["base"], None, None, precision=L.FV_GPU_PREC_BF16)
    b = E.concat([E.serialize(None, n) for _ in range(nq)], n)
    bh = g.stage(d["queries"], b, 10)
    try:
        with pytest.raises(E.GpuError, match="graph_cus"):
            g.run_learned(bh, None, L.LearnedOpts(k=10, mode=3, graph_cus=32))
        g.run_learned(bh, None, L.LearnedOpts(k=10, mode=3))
        g.sync()
        assert g.fetch(bh, nq, 10)[3] == 0
    finally:
        g.free_batch(bh)
        g.close()
