"""Parity at the benchmark's size: the 64^3 heterogeneous elasticity system (823,875 DOF, seed 0).

On this system each apply workgroup walks many of the 9.5k bank matrices through the persistent
grid-stride loop with two matrices in flight, which the small systems of the other parity tests
never reach with the default launch.  Here the HIP path is compared with the CPU oracle on the
same domains: the coarse map bit for bit, a strided sample of the FP64 bank inverses, the FP64 MAS
apply on two vectors, and the mixed (FP16) apply against the FP64 one.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from polysolve_amd import MASSolver, problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hex64():
    p = problems.config2(seed=0)
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 3, "aggregation": "morton", "coarse_solver": "bank"}})
    s.set_coordinates(p.coords)
    s.factorize(p.A)
    offsets, perm = s.partition()
    O = oracle.MasOracle(p.A, 3, offsets, perm)
    assert O.status == 0
    return p, s, O, perm


def test_hex64_coarse_map_matches_oracle(gpu, hex64):
    p, s, O, perm = hex64
    li = s.level_info()
    assert li["n"] == 823875 and li["nnzb"] == 7081535
    assert li["level_num"] == O.info["level_num"]
    assert [li[f"cco{i}"] for i in range(4)] == [O.info[f"cco{i}"] for i in range(4)]
    om, _ = O.coarse_map()
    L = O.info["level_num"]
    assert np.array_equal(s.coarse_map()[:, :L], om[perm, :L])


def test_hex64_bank_inverses_sample_match_oracle(gpu, hex64):
    _, s, O, _ = hex64
    _, inv_ref = O.matrices()
    inv = s.bank_inverses(3)
    assert inv.shape == inv_ref.shape and inv.shape[0] > 9000
    # Every 37th matrix (coprime to the 8-matrix launch groups), plus the last of every level.
    take = np.unique(np.concatenate([np.arange(0, inv.shape[0], 37), [inv.shape[0] - 1]]))
    scale = np.abs(inv_ref[take]).max(axis=1, keepdims=True)
    assert (np.abs(inv[take] - inv_ref[take]) / scale).max() < 1e-9


def test_hex64_apply_fp64_matches_oracle(gpu, hex64):
    p, s, O, _ = hex64
    rng = np.random.default_rng(64)
    for r in (rng.standard_normal(p.n), p.b.copy()):
        z = s.apply_preconditioner(r)
        zr = O.apply(r)
        assert np.linalg.norm(z - zr) <= 1e-10 * np.linalg.norm(zr)


def test_hex64_apply_mixed_close_to_fp64(gpu, hex64):
    p, s, _, _ = hex64
    rng = np.random.default_rng(65)
    r = rng.standard_normal(p.n)
    z64 = s.apply_preconditioner(r)
    z16 = s.apply_preconditioner(r, mixed=True)
    assert np.linalg.norm(z16 - z64) <= 5e-3 * np.linalg.norm(z64)
    assert r @ z16 > 0
