"""CPU oracle of the scheduler's network (test infrastructure only, like the rest of oracle/):
the 12 -> 32 -> 32 -> 3 MLP with ReLU hidden layers in float64, its loss, its gradient by hand
written backpropagation, and Adam. It restates the arithmetic the GPU kernels implement
(grafeo-cuda_amd/csrc/kernels/mlp.hip); there is no reference counterpart (the learned scheduler
is an addition). Pinned by tests/test_oracle_mlp.py against torch.autograd in float64 and against
central finite differences.

Parameter layout (1571 values): W1[32][12], b1[32], W2[32][32], b2[32], W3[3][32], b3[3].
Inputs are normalised as (x - mean) * scale. Loss over a batch: sum(mask * (y - t)^2) / sum(mask).
"""

from __future__ import annotations

import numpy as np

IN, H, OUT = 12, 32, 3
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]
PARAMS = sum(int(np.prod(s)) for _, s in SIZES)  # 1571


def unpack(p: np.ndarray) -> dict:
    out, i = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[i:i + k], dtype=np.float64).reshape(shape)
        i += k
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def forward(p, x, mean=None, scale=None):
    """(y, cache) for a batch x[count][12]."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    mean = np.zeros(IN) if mean is None else np.asarray(mean, np.float64)
    scale = np.ones(IN) if scale is None else np.asarray(scale, np.float64)
    xn = (x - mean) * scale
    a1 = xn @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (w, xn, a1, h1, a2, h2)


def loss(p, x, t, mask, mean=None, scale=None) -> float:
    y, _ = forward(p, x, mean, scale)
    mask = np.asarray(mask, np.float64)
    return float(np.sum(mask * (y - np.asarray(t, np.float64)) ** 2) / np.sum(mask))


def loss_grad(p, x, t, mask, mean=None, scale=None):
    """(loss, dL/dp) by backpropagation."""
    y, (w, xn, a1, h1, a2, h2) = forward(p, x, mean, scale)
    mask = np.asarray(mask, np.float64)
    obs = np.sum(mask)
    e = (y - np.asarray(t, np.float64)) * mask
    value = float(np.sum(e * e) / obs)
    dy = 2.0 * e / obs                        # [count][3]
    g = {"W3": dy.T @ h2, "b3": dy.sum(axis=0)}
    dh2 = (dy @ w["W3"]) * (a2 > 0)           # [count][32]
    g["W2"], g["b2"] = dh2.T @ h1, dh2.sum(axis=0)
    dh1 = (dh2 @ w["W2"]) * (a1 > 0)
    g["W1"], g["b1"] = dh1.T @ xn, dh1.sum(axis=0)
    return value, pack(g)


def adam(p, x, t, mask, steps, lr, beta1=0.9, beta2=0.999, eps=1e-8, mean=None, scale=None):
    """`steps` full-batch Adam steps (bias-corrected); returns (params, final loss)."""
    p = np.asarray(p, np.float64).copy()
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    for step in range(1, steps + 1):
        _, g = loss_grad(p, x, t, mask, mean, scale)
        m1 = beta1 * m1 + (1 - beta1) * g
        m2 = beta2 * m2 + (1 - beta2) * g * g
        p -= lr * (m1 / (1 - beta1 ** step)) / (np.sqrt(m2 / (1 - beta2 ** step)) + eps)
    return p, loss(p, x, t, mask, mean, scale)
