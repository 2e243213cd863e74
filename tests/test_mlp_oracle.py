"""The MLP oracle pinned on the CPU: backpropagation against central finite differences in
float64, and the bfloat16 operand rounding against its definition."""
from __future__ import annotations

import numpy as np

from oracle import mlp_oracle as mo


def _setup(d_in=5, d_out=3, h=7, B=6, seed=0):
    rng = np.random.default_rng(seed)
    n = sum(int(np.prod(s)) for s in mo.shapes(d_in, d_out, h))
    theta = rng.standard_normal(n) * 0.5
    X = rng.standard_normal((B, d_in))
    T = rng.standard_normal((B, d_out))
    return theta, X, T, d_in, d_out, h


def test_gradient_matches_finite_differences():
    theta, X, T, d_in, d_out, h = _setup()
    _, g = mo.loss_and_grad(theta, X, T, d_in, d_out, h)
    idx = np.arange(theta.size)  # every parameter: W1, b1, W2, b2, W3, b3
    fd = mo.finite_difference(theta, X, T, d_in, d_out, h, idx=idx, eps=1e-6)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.abs(g).max())


def test_loss_is_mean_squared_error():
    theta, X, T, d_in, d_out, h = _setup(seed=1)
    y, _ = mo.forward(theta, X, d_in, d_out, h)
    loss, _ = mo.loss_and_grad(theta, X, T, d_in, d_out, h)
    assert np.isclose(loss, np.mean((y - T) ** 2))
    # With W3 = 0 the output is b3 for every sample.
    W = mo.unpack(theta.copy(), d_in, d_out, h)
    W[4][:] = 0.0
    theta2 = np.concatenate([w.ravel() for w in W])
    y2, _ = mo.forward(theta2, X, d_in, d_out, h)
    assert np.allclose(y2, W[5][None, :])


def test_bf16_rounding():
    # Exactly representable values stay; ties round to even; 1 + 2^-9 rounds down to 1.
    a = np.array([1.0, 1.5, -2.25, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, 3.0e38], dtype=np.float32)
    r = mo.bf16_round(a)
    assert r[0] == 1.0 and r[1] == 1.5 and r[2] == -2.25
    assert r[3] == 1.0            # tie (1 + 2^-8) -> even mantissa 1.0
    assert r[4] == 1 + 2 ** -6    # tie (1 + 3*2^-8) -> even mantissa 1 + 4*2^-8
    assert r[5] == 1.0
    assert np.isfinite(r[6])
    # Operands already in bfloat16 make the bf16 forward identical to the float64 forward.
    theta, X, T, d_in, d_out, h = _setup(seed=2)
    theta = mo.bf16_round(theta).astype(np.float64)
    Xb = mo.bf16_round(X).astype(np.float64)
    # First layer only: h1 is then exactly representable? not in general, so compare layer 1.
    W1, b1 = mo.unpack(theta, d_in, d_out, h)[:2]
    z_ref = Xb @ W1.T + b1
    z_bf = mo._mm(Xb, W1.T, True) + b1
    assert np.allclose(z_ref, z_bf, rtol=0, atol=1e-12)


def test_skip_layer_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    d, h = 4, 6
    n = sum(int(np.prod(q)) for q in mo.shapes(d, d, h, skip=True))
    theta = rng.standard_normal(n) * 0.5
    X = rng.standard_normal((5, d))
    T = rng.standard_normal((5, d))
    _, g = mo.loss_and_grad(theta, X, T, d, d, h, skip=True)
    fd = mo.finite_difference(theta, X, T, d, d, h, idx=np.arange(n), eps=1e-6, skip=True)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.abs(g).max())
    # With W3 = 0 and b3 = 0 the network is the skip layer alone.
    W = mo.unpack(theta.copy(), d, d, h, skip=True)
    W[4][:] = 0.0
    W[5][:] = 0.0
    th2 = np.concatenate([w.ravel() for w in W])
    y, _ = mo.forward(th2, X, d, d, h, skip=True)
    assert np.allclose(y, X @ W[6].T)
