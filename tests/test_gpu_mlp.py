"""GPU MLP of the learned coarse correction: MFMA GEMM, forward, backward and training against
the numpy oracle (operands rounded to bfloat16 in both, FP32 accumulation on the GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from polysolve_amd.linear import Mlp, gemm_bf16

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K,trans_a", [(32, 32, 8, False), (64, 64, 32, False), (70, 45, 37, False),
                                           (128, 200, 96, False), (33, 17, 130, True), (128, 128, 256, True)])
def test_gemm_bf16_matches_numpy(gpu, M, N, K, trans_a):
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.standard_normal((K, M) if trans_a else (M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C = gemm_bf16(A, B, trans_a=trans_a)
    Ab = mo.bf16_round(A).astype(np.float64)
    ref = (Ab.T if trans_a else Ab) @ mo.bf16_round(B).astype(np.float64)
    # Exact BF16 products accumulated in FP32: error <= K * eps32 * sum|a b|.
    bound = K * 2.0 ** -23 * (np.abs(Ab.T if trans_a else Ab) @ np.abs(mo.bf16_round(B).astype(np.float64)))
    assert np.all(np.abs(C - ref) <= bound + 1e-30)


def _net(d_in=24, d_out=9, h=128, seed=5):
    net = Mlp(d_in, d_out, h, seed=seed)
    theta = net.get_parameters().astype(np.float64)
    return net, theta, d_in, d_out, h


def test_forward_matches_oracle(gpu):
    net, theta, d_in, d_out, h = _net()
    rng = np.random.default_rng(0)
    X = rng.standard_normal((300, d_in)).astype(np.float32)
    Y = net.forward(X)
    ref, _ = mo.forward(theta, X.astype(np.float64), d_in, d_out, h, bf16=True)
    # Hidden activations are rounded to bfloat16 from FP32 on the GPU and from float64 in the
    # oracle; a value on a rounding boundary may land one bf16 ulp apart.
    assert np.linalg.norm(Y - ref) <= 2e-3 * np.linalg.norm(ref)


def test_gradients_match_oracle(gpu):
    net, theta, d_in, d_out, h = _net(seed=6)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((256, d_in)).astype(np.float32)
    T = rng.standard_normal((256, d_out)).astype(np.float32)
    loss, g = net.loss_and_grad(X, T)
    lref, gref = mo.loss_and_grad(theta, X.astype(np.float64), T.astype(np.float64), d_in, d_out, h, bf16=True)
    assert abs(loss - lref) <= 1e-3 * lref
    offs = np.cumsum([0] + [int(np.prod(s)) for s in mo.shapes(d_in, d_out, h)])
    for i in range(6):
        a, b = g[offs[i]:offs[i + 1]], gref[offs[i]:offs[i + 1]]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), f"parameter block {i}"
    # And the float64 (unrounded) gradient is within bfloat16 accuracy.
    _, g64 = mo.loss_and_grad(theta, X.astype(np.float64), T.astype(np.float64), d_in, d_out, h, bf16=False)
    assert np.linalg.norm(g - g64) <= 3e-2 * np.linalg.norm(g64)


def test_parameters_roundtrip_and_errors(gpu):
    net, theta, d_in, d_out, h = _net()
    t = np.arange(net.num_parameters, dtype=np.float32) * 1e-4
    net.set_parameters(t)
    assert np.array_equal(net.get_parameters(), t)
    with pytest.raises(Exception):
        net.set_parameters(t[:-1])


def test_training_fits_a_smooth_map(gpu):
    rng = np.random.default_rng(2)
    d_in, d_out = 8, 4
    Wt = rng.standard_normal((d_out, d_in)) / np.sqrt(d_in)
    X = rng.standard_normal((4096, d_in)).astype(np.float32)
    T = (np.tanh(X @ Wt.T) + 0.1 * (X[:, :d_out] ** 2)).astype(np.float32)
    net = Mlp(d_in, d_out, 64, seed=3)
    l0 = float(np.mean((net.forward(X) - T) ** 2))
    last = net.train(X, T, epochs=60, batch=256, lr=2e-3, seed=4)
    l1 = float(np.mean((net.forward(X) - T) ** 2))
    assert l1 < 0.05 * l0, (l0, last, l1)
