"""Pins the FP64 MLP oracle (oracle/mlp_ref.py): its analytic gradient
against central finite differences, its BF16 rounding against known
answers. CPU only."""
import numpy as np

import mlp_ref as R


def _problem(seed=0, B=7, dim=8):
    rng = np.random.default_rng(seed)
    F = R.nfeat(dim)
    p = rng.standard_normal(R.nparams(F)) * 0.3
    x = rng.standard_normal((B, F)) * 2 + 1
    t = np.stack([rng.standard_normal(B), rng.integers(0, 2, B), rng.integers(0, 2, B)], 1)
    mean = rng.standard_normal(F) * 0.5
    std = rng.uniform(0.5, 2.0, F)
    return p, x, t.astype(np.float64), mean, std, F


def test_gradient_matches_finite_differences():
    p, x, t, mean, std, F = _problem()
    g = R.grad(p, x, t, mean, std, F, 0.7)
    rng = np.random.default_rng(1)
    for k in rng.choice(len(p), 60, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss(pp, x, t, mean, std, F, 0.7) - R.loss(pm, x, t, mean, std, F, 0.7)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_parameter_count():
    assert R.nparams(R.nfeat(128)) == 13634   # 145 -> 64 -> 64 -> 2


def test_bf16_known_answers():
    v = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.0, 3.0e38], np.float32)
    # 1 + 2^-8 rounds to even (1.0); 1 + 1.5*2^-8 rounds up to 1 + 2^-7
    assert list(R.bf16(v)[:4]) == [1.0, 1.0, 1.0078125, 1.0078125]
    assert R.bf16(v)[4] == -2.0
