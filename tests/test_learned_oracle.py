"""Pins the FP64 MLP oracle (oracle/mlp_ref.py): its analytic gradient
against central finite differences; its forward pass, standardisation,
loss and feature builder against hand-worked known answers (a sparse
network whose outputs are computed by hand, centroids at known distances);
its BF16 rounding against known answers. CPU only."""
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


def _sparse_net(F):
    """W1[0, 0] = 2, b1[0] = -1; W2[0, 0] = 3; W3 = [[0.5], [-1]] on unit 0;
    b3 = [0.25, 1]; everything else zero."""
    p = np.zeros(R.nparams(F))
    W1, b1, W2, b2, W3, b3 = R.unpack(p, F)   # views into p
    W1[0, 0] = 2.0
    b1[0] = -1.0
    W2[0, 0] = 3.0
    W3[0, 0] = 0.5
    W3[1, 0] = -1.0
    b3[:] = [0.25, 1.0]
    return p


def test_forward_and_loss_known_answers():
    F = R.nfeat(1)
    p = _sparse_net(F)
    x = np.zeros((3, F))
    x[0, 0] = 2.0      # x_hat 2: z1 = 3, h2 = 9, y = [4.75, -8]
    x[1, 0] = 0.0      # z1 = -1: relu zero, y = b3
    x[2, 0] = 5.0      # with mean 1, std 2 below: x_hat 2 again
    mean = np.zeros(F)
    std = np.ones(F)
    y, _ = R.forward(p, x[:2], mean, std, F)
    assert y.tolist() == [[4.75, -8.0], [0.25, 1.0]]
    mean[0], std[0] = 1.0, 2.0
    y2, _ = R.forward(p, x[2:], mean, std, F)
    assert y2.tolist() == [[4.75, -8.0]]
    # loss: valid (4.75 - 4)^2 + lam (softplus(-8) - 0 * -8); the second
    # sample invalid (t2 = 0) with label 1: lam (softplus(1) - 1)
    t = np.array([[4.0, 0.0, 1.0], [9.0, 1.0, 0.0]])
    lam = 0.5
    want = (0.5625 + lam * np.log1p(np.exp(-8.0)) + lam * (np.log1p(np.e) - 1.0)) / 2
    assert abs(R.loss(p, x[:2], t, np.zeros(F), np.ones(F), F, lam) - want) < 1e-15


def test_features_known_answers():
    # 17 centroids on the first axis at 0..16: the 16 nearest to the origin
    # are at distances 0..15; selectivity 50 / 1000 and an empty set
    cent = np.zeros((17, 2))
    cent[:, 0] = np.arange(17)
    f = R.features(np.zeros((2, 2)), [50, 0], 1000, cent)
    assert f.shape == (2, 2 + 1 + 16)
    assert f[0, :2].tolist() == [0.0, 0.0]
    assert abs(f[0, 2] - np.log10(0.05)) < 1e-15
    assert f[1, 2] == -3.0
    assert f[0, 3:].tolist() == [float(k) for k in range(16)]
    # a query between centroids: distances |q - k|, ascending
    f = R.features(np.array([[7.5, 0.0]]), [1000], 1000, cent)
    assert f[0, 2] == 0.0
    assert f[0, 3:].tolist() == [0.5, 0.5, 1.5, 1.5, 2.5, 2.5, 3.5, 3.5, 4.5, 4.5, 5.5, 5.5,
                                  6.5, 6.5, 7.5, 7.5]
