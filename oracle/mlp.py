"""CPU restatement of the surrogate network (test infrastructure only).

No counterpart exists in the reference: WInDS has no learned model.  The
network is y = W_L s(... s(W_1 x + b_1) ...) + b_L with s(z) = z * sigmoid(z)
(SiLU) and the loss MSE = sum (y - t)^2 / (B * n_out).  forward/backward run
in FP64; with bf16=True the weights, the layer inputs and the back-propagated
deltas are rounded to bfloat16 (round to nearest even) where the device
kernels round them, so the device gradients can be compared at FP32
accumulation accuracy.  The FP64 gradient is pinned by central finite
differences (tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = np.where(finite, r, u & 0xFFFF0000).astype(np.uint32)
    return out.view(np.float32).astype(np.float64)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(params, X, bf=False):
    """params: [(W [n_out, n_in], b [n_out])...]; returns (Y, cache)."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    a = q(X)
    cache = []
    for i, (W, b) in enumerate(params):
        z = a @ q(W).T + np.asarray(b, dtype=np.float64)
        cache.append((a, z))
        if i + 1 < len(params):
            a = q(z * _sig(z))
        else:
            a = z
    return a, cache


def loss_and_grads(params, X, T, bf=False):
    """MSE loss and [(dW, db)...] of the same shapes as params."""
    q = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    Y, cache = forward(params, X, bf)
    B, n = Y.shape
    r = Y - np.asarray(T, dtype=np.float64)
    loss = float(np.sum(r * r) / (B * n))
    dz = 2.0 * r / (B * n)
    grads = [None] * len(params)
    for i in range(len(params) - 1, -1, -1):
        a, _ = cache[i]
        dq = q(dz)
        grads[i] = (dq.T @ a, dz.sum(axis=0))
        if i > 0:
            W = q(params[i][0])
            da = dq @ W
            zp = cache[i - 1][1]
            s = _sig(zp)
            dz = da * (s * (1.0 + zp * (1.0 - s)))
    return loss, grads


def flatten(pairs) -> np.ndarray:
    return np.concatenate([np.concatenate([np.ravel(W), np.ravel(b)]) for W, b in pairs])


def loss_only(params, X, T):
    Y, _ = forward(params, X)
    r = Y - T
    return float(np.sum(r * r) / r.size)
