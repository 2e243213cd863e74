"""CPU oracle of the learned-correction MLP — TEST INFRASTRUCTURE ONLY.

numpy restatement of y = W3 relu(W2 relu(W1 x + b1) + b2) + b3 (+ S x with the optional linear skip
layer S) with mean-squared-error loss L = sum((y - t)^2) / (B * d_out), its analytic gradient
(backpropagation), and a finite-difference check that pins the backpropagation formulas in float64.
``bf16=True`` rounds every GEMM operand to bfloat16 (round to nearest even) the way the GPU kernels
stage operands in LDS, so that the GPU gradients can be compared at float32 accumulation accuracy.

Parameter layout (flat float32): [W1 (h x d_in) | b1 | W2 (h x h) | b2 | W3 (d_out x h) | b3] and,
with skip=True, S (d_out x d_in) appended.
"""
from __future__ import annotations

import numpy as np


def shapes(d_in: int, d_out: int, h: int = 128, skip: bool = False):
    s = [(h, d_in), (h,), (h, h), (h,), (d_out, h), (d_out,)]
    return s + [(d_out, d_in)] if skip else s


def unpack(theta: np.ndarray, d_in: int, d_out: int, h: int = 128, skip: bool = False):
    out, o = [], 0
    for s in shapes(d_in, d_out, h, skip):
        k = int(np.prod(s))
        out.append(theta[o:o + k].reshape(s))
        o += k
    return out


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even) and back to float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _mm(a, b, bf16):
    if bf16:
        return bf16_round(a).astype(np.float64) @ bf16_round(b).astype(np.float64)
    return a @ b


def forward(theta, X, d_in, d_out, h=128, bf16=False, skip=False):
    W = unpack(theta, d_in, d_out, h, skip)
    W1, b1, W2, b2, W3, b3 = W[:6]
    z1 = _mm(X, W1.T, bf16) + b1
    h1 = np.maximum(z1, 0)
    z2 = _mm(h1, W2.T, bf16) + b2
    h2 = np.maximum(z2, 0)
    y = _mm(h2, W3.T, bf16) + b3
    if skip:
        y = y + _mm(X, W[6].T, bf16)
    return y, (h1, h2)


def loss_and_grad(theta, X, T, d_in, d_out, h=128, bf16=False, skip=False):
    W = unpack(theta, d_in, d_out, h, skip)
    W1, b1, W2, b2, W3, b3 = W[:6]
    B = X.shape[0]
    y, (h1, h2) = forward(theta, X, d_in, d_out, h, bf16, skip)
    diff = y - T
    loss = float((diff ** 2).sum() / (B * d_out))
    dy = 2.0 * diff / (B * d_out)
    gW3 = _mm(dy.T, h2, bf16)
    gb3 = dy.sum(axis=0)
    dh2 = _mm(dy, W3, bf16) * (h2 > 0)
    gW2 = _mm(dh2.T, h1, bf16)
    gb2 = dh2.sum(axis=0)
    dh1 = _mm(dh2, W2, bf16) * (h1 > 0)
    gW1 = _mm(dh1.T, X, bf16)
    gb1 = dh1.sum(axis=0)
    parts = [gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3]
    if skip:
        parts.append(_mm(dy.T, X, bf16).ravel())
    return loss, np.concatenate(parts)


def finite_difference(theta, X, T, d_in, d_out, h=128, idx=None, eps=1e-6, skip=False):
    """Central differences of the float64 loss at the given parameter indices."""
    theta = np.asarray(theta, dtype=np.float64)
    out = []
    for i in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += eps
        tm[i] -= eps
        lp, _ = loss_and_grad(tp, X, T, d_in, d_out, h, skip=skip)
        lm, _ = loss_and_grad(tm, X, T, d_in, d_out, h, skip=skip)
        out.append((lp - lm) / (2 * eps))
    return np.array(out)
