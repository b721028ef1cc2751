"""FP64 restatement of the learned router's MLP (TEST INFRASTRUCTURE ONLY:
the checker for paves_amd/csrc/pv_learn.hip; never used by the product).

Layout and loss as in include/paves_gpu.h (fv_gpu_model_*): parameters
W1[64][F] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]; inputs standardised
x_hat = (x - mean) / std; h1 = relu(W1 x_hat + b1); h2 = relu(W2 h1 + b2);
y = W3 h2 + b3; per-sample loss valid (y0 - t0)^2 + lambda (softplus(y1)
- t1 y1); mean over samples. Pinned by central finite differences and
hand-worked known answers for forward, loss and features
(tests/test_learned_oracle.py).
"""
from __future__ import annotations

import numpy as np

HID = 64
NEAR = 16


def nfeat(dim: int) -> int:
    return dim + 1 + NEAR


def nparams(F: int) -> int:
    return HID * F + HID + HID * HID + HID + 2 * HID + 2


def unpack(p: np.ndarray, F: int):
    o = 0
    W1 = p[o:o + HID * F].reshape(HID, F); o += HID * F
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    W3 = p[o:o + 2 * HID].reshape(2, HID); o += 2 * HID
    b3 = p[o:o + 2]
    return W1, b1, W2, b2, W3, b3


def forward(p, x, mean, std, F):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), F)
    xh = (np.asarray(x, np.float64) - mean) / std
    z1 = xh @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3.T + b3
    return y, (xh, z1, h1, z2, h2)


def loss(p, x, t, mean, std, F, lam):
    y, _ = forward(p, x, mean, std, F)
    e0 = y[:, 0] - t[:, 0]
    z = y[:, 1]
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    return float(np.mean(e0 * e0 * t[:, 2] + lam * (sp - t[:, 1] * z)))


def grad(p, x, t, mean, std, F, lam):
    """Analytic gradient of `loss` (FP64 back-propagation)."""
    p = np.asarray(p, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, F)
    y, (xh, z1, h1, z2, h2) = forward(p, x, mean, std, F)
    B = x.shape[0]
    g0 = 2 * (y[:, 0] - t[:, 0]) * t[:, 2]
    g1 = lam * (1 / (1 + np.exp(-y[:, 1])) - t[:, 1])
    d3 = np.stack([g0, g1], 1) / B
    d2 = (d3 @ W3) * (z2 > 0)
    d1 = (d2 @ W2) * (z1 > 0)
    return np.concatenate([(d1.T @ xh).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0),
                           (d3.T @ h2).ravel(), d3.sum(0)])


def bf16(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def predict_bf16(p, x, mean, std, F):
    """Inference with BF16-rounded parameters (FP64 arithmetic)."""
    y, _ = forward(bf16(p), np.asarray(x, np.float32), mean, std, F)
    return y


def features(queries, sel_count, count, centroids):
    """[q | log10(max(pass,1)/count) | sqrt of the 16 smallest squared
    distances to the centroids, ascending] per query (FP64)."""
    q = np.asarray(queries, np.float64)
    c = np.asarray(centroids, np.float64)
    d2 = ((q[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    near = np.sqrt(np.sort(d2, 1)[:, :NEAR])
    sel = np.log10(np.maximum(np.asarray(sel_count, np.float64), 1.0) / count)
    return np.concatenate([q, sel[:, None], near], 1)
