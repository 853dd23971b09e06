"""FP64 numpy restatement of the surrogate's circular 3x3 convolutions.

TEST INFRASTRUCTURE ONLY: the checker for the BF16 MFMA kernels.  Layout
conventions follow include/sps_nn.h: NHWC activations, weights [tap][out][in]
with tap = 3*r + s and offsets (r-1, s-1), periodic wrap in H and W.

The backward formulas here are pinned in tests/test_conv_oracle.py against
central finite differences of the forward in FP64.
"""
from __future__ import annotations

import numpy as np


def _shift(x, dr, ds):
    """value at (h+dr, w+ds) for every (h, w), periodic"""
    return np.roll(np.roll(x, -dr, axis=1), -ds, axis=2)


def conv3x3(x, w, b=None):
    N, H, W, C = x.shape
    y = np.zeros((N, H, W, w.shape[1]))
    for t in range(9):
        y += _shift(x, t // 3 - 1, t % 3 - 1) @ w[t].T
    if b is not None:
        y += b
    return y


def conv3x3_bwd_data(dy, w):
    dx = np.zeros(dy.shape[:3] + (w.shape[2],))
    for t in range(9):
        dx += _shift(dy, -(t // 3 - 1), -(t % 3 - 1)) @ w[t]
    return dx


def conv3x3_bwd_weight(x, dy):
    K, C = dy.shape[3], x.shape[3]
    dw = np.zeros((9, K, C))
    for t in range(9):
        xs = _shift(x, t // 3 - 1, t % 3 - 1).reshape(-1, C)
        dw[t] = dy.reshape(-1, K).T @ xs
    return dw


def silu(v):
    return v / (1.0 + np.exp(-v))


def silu_grad(v):
    s = 1.0 / (1.0 + np.exp(-v))
    return s * (1.0 + v * (1.0 - s))


def upsample2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def pool2(x, mode="avg"):
    N, H, W, C = x.shape
    s = x.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))
    return s / 4.0 if mode == "avg" else s


def bf16_round(a):
    """Round FP32/FP64 values to the nearest BF16 (ties to even), as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def to_bf16_bits(a):
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(bits):
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)
