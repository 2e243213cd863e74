"""Learned coarse correction of the mixed path: the lattice coarse space against a numpy
construction, the network's skip layer against the oracle, the gain of the trained correction
on a drifting sequence of heterogeneous elasticity systems, and the plain-MAS fallback."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from polysolve_amd import MASSolver, problems
from polysolve_amd.linear import Mlp
from tests import helpers

pytestmark = pytest.mark.gpu


def _solver(params):
    s = MASSolver()
    prm = {"block_dim": 3, "precision": "mixed", "relative_tolerance": 1e-8, "absolute_tolerance": 0.0,
           "max_iter": 3000}
    prm.update(params)
    s.set_parameters({"MAS": prm})
    return s


@pytest.mark.parametrize("k,block_path", [(4, True), (10, False)])
def test_lattice_restriction_and_prolongation_match_numpy(gpu, k, block_path):
    """k = 4: every node block touches few lattice points (per-block reduction); k = 10 on an 8^3
    mesh: blocks touch more than 64 points and the per-point entry lists are used."""
    p = problems.hex_elasticity(8, E=problems.random_moduli(8, 11))
    s = _solver({"lattice_points": k})
    s.set_coordinates(p.coords)
    s.factorize(p.A)
    P = helpers.lattice_prolongation(p.A, p.coords, k)
    rng = np.random.default_rng(0)
    v = rng.standard_normal(p.n)
    u = s.lattice_restrict(v)
    ref = P.T @ v
    assert u.shape == ref.shape
    assert np.linalg.norm(u - ref) <= 1e-6 * np.linalg.norm(ref)
    w = rng.standard_normal(P.shape[1])
    e = s.lattice_prolong(w)
    assert np.linalg.norm(e - P @ w) <= 1e-6 * np.linalg.norm(P @ w)
    # Dirichlet rows receive no correction.
    assert np.all(e[P.getnnz(axis=1) == 0] == 0.0)
    assert s.get_info()["learned"]["block_restriction"] == block_path


def test_mlp_skip_layer_gradients_match_oracle(gpu):
    d, h = 24, 32
    net = Mlp(d, d, h, seed=3, skip=True)
    theta = net.get_parameters().astype(np.float64)
    rng = np.random.default_rng(4)
    theta[-d * d:] = rng.standard_normal(d * d) * 0.1  # a non-zero skip layer
    net.set_parameters(theta.astype(np.float32))
    theta = net.get_parameters().astype(np.float64)
    X = rng.standard_normal((128, d)).astype(np.float32)
    T = rng.standard_normal((128, d)).astype(np.float32)
    Y = net.forward(X)
    ref, _ = mo.forward(theta, X.astype(np.float64), d, d, h, bf16=True, skip=True)
    assert np.linalg.norm(Y - ref) <= 2e-3 * np.linalg.norm(ref)
    loss, g = net.loss_and_grad(X, T)
    lref, gref = mo.loss_and_grad(theta, X.astype(np.float64), T.astype(np.float64), d, d, h, bf16=True, skip=True)
    assert abs(loss - lref) <= 1e-3 * lref
    offs = np.cumsum([0] + [int(np.prod(q)) for q in mo.shapes(d, d, h, skip=True)])
    for i in range(7):
        a, b = g[offs[i]:offs[i + 1]], gref[offs[i]:offs[i + 1]]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), f"parameter block {i}"


NE = 16
TRAIN, EVAL = range(1, 7), (12, 20)


@pytest.fixture(scope="module")
def drift_sequence():
    moduli = list(problems.config5_moduli(NE, steps=max(EVAL) + 1, seed=3))
    return {t: problems.hex_elasticity(NE, E=moduli[t], name=f"drift{t}") for t in list(TRAIN) + list(EVAL)}


@pytest.fixture(scope="module")
def trained_solver(drift_sequence):
    s = _solver({"lattice_points": 5, "learned_collect": True, "learned_random_pairs": 120,
                 "learned_snapshots": 64})
    for t in TRAIN:
        p = drift_sequence[t]
        s.set_coordinates(p.coords)
        s.factorize(p.A)
        x = np.zeros(p.n)
        s.solve(p.b, x)
        assert s.get_info()["solver_status"] == "Reach relative tolerance"
    info = s.get_info()["learned"]
    assert info["pairs"] > 600, info
    loss = s.learned_train(epochs=200, batch=64, lr=3e-3, seed=0)
    assert np.isfinite(loss)
    s.set_parameters({"MAS": {"learned_collect": False}})
    return s


def _solve(s, p, learned):
    s.set_parameters({"MAS": {"learned_correction": learned}})
    s.set_coordinates(p.coords)
    s.factorize(p.A)
    x = np.zeros(p.n)
    s.solve(p.b, x)
    info = s.get_info()
    rel = np.linalg.norm(p.b - p.A @ x) / np.linalg.norm(p.b)
    return info, rel, x


def test_learned_correction_reduces_iterations(gpu, drift_sequence, trained_solver):
    """The trained network's gain over plain MAS on later systems of the drifting sequence, at
    the same FP64 accuracy."""
    s = trained_solver
    plain, learned = [], []
    for t in EVAL:
        p = drift_sequence[t]
        i0, r0, x0 = _solve(s, p, False)
        i1, r1, x1 = _solve(s, p, True)
        assert i0["solver_status"] == i1["solver_status"] == "Reach relative tolerance"
        assert r0 <= 1e-8 and r1 <= 1e-8
        assert i1["learned"]["used"] and not i1["learned"]["fallback"], i1["learned"]
        assert not i0["learned"]["used"]
        assert np.linalg.norm(x1 - x0) <= 1e-6 * np.linalg.norm(x0)
        plain.append(i0["solver_iter"])
        learned.append(i1["solver_iter"])
    print("plain MAS iterations", plain, "learned", learned)
    assert sum(learned) <= 0.75 * sum(plain), (plain, learned)


def test_exact_and_linear_corrections(gpu, drift_sequence):
    """Ablations of the network: the exact lattice coarse solve (P^T A P)^-1 P^T r installed in its
    skip layer, and a skip-layer-only network trained on the same pairs, both inside the same
    flexible PCG.  Each reaches the FP64 tolerance with fewer iterations than plain MAS."""
    p = drift_sequence[EVAL[0]]
    s = _solver({"lattice_points": 5, "learned_collect": True, "learned_random_pairs": 120,
                 "learned_snapshots": 64, "learned_network": "linear"})
    for t in TRAIN:
        q = drift_sequence[t]
        s.set_coordinates(q.coords)
        s.factorize(q.A)
        s.solve(q.b, np.zeros(q.n))
    s.learned_train(epochs=200, batch=64, lr=3e-3, seed=0)
    s.set_parameters({"MAS": {"learned_collect": False}})
    plain, _, x0 = _solve(s, p, False)
    lin, r_lin, x_lin = _solve(s, p, True)
    assert lin["learned"]["network"] == "linear" and lin["learned"]["used"]
    ms = s.learned_exact()
    assert ms > 0
    ex, r_ex, x_ex = _solve(s, p, True)  # refactorises the same matrix: the installed map stays
    assert ex["learned"]["network"] == "galerkin" and ex["learned"]["used"]
    print("plain", plain["solver_iter"], "linear", lin["solver_iter"], "exact", ex["solver_iter"])
    for info, rel, x in ((lin, r_lin, x_lin), (ex, r_ex, x_ex)):
        assert info["solver_status"] == "Reach relative tolerance" and rel <= 1e-8
        assert not info["learned"]["fallback"]
        assert info["solver_iter"] < plain["solver_iter"]
        assert np.linalg.norm(x - x0) <= 1e-6 * np.linalg.norm(x0)


@pytest.mark.parametrize("single_reduction", [False, True])
def test_learned_fallback_to_plain_mas_on_stall(gpu, drift_sequence, trained_solver, single_reduction):
    """A correction scaled far beyond what was trained stops the residual from decreasing; plain
    MAS takes over and the solve still reaches the FP64 tolerance.  Both PCG iterations (the
    single-reduction one assembles u = M r + P c in its SpMV and restarts with beta = 0)."""
    s = trained_solver
    s.set_parameters({"MAS": {"learned_output_scale": 1e6, "learned_stall_segments": 2,
                              "single_reduction": single_reduction}})
    try:
        info, rel, _ = _solve(s, drift_sequence[EVAL[0]], True)
    finally:
        s.set_parameters({"MAS": {"learned_output_scale": 1.0, "learned_stall_segments": 3,
                                  "single_reduction": False}})
    assert info["pcg_iteration"] == ("single_reduction" if single_reduction else "classic")
    assert info["learned"]["fallback"], info["learned"]
    assert info["learned"]["fallback_iteration"] > 0
    assert info["solver_status"] == "Reach relative tolerance"
    assert rel <= 1e-8


def test_learned_correction_without_network_is_plain(gpu, drift_sequence):
    s = _solver({"learned_correction": True})
    p = drift_sequence[EVAL[0]]
    s.set_coordinates(p.coords)
    s.factorize(p.A)
    x = np.zeros(p.n)
    s.solve(p.b, x)
    info = s.get_info()
    assert info["solver_status"] == "Reach relative tolerance"
    assert not info["learned"]["used"] and "no trained network" in info["learned"]["note"]


def test_info_stays_json_when_training_diverges(gpu):
    """A learning rate far too large drives the training loss to a non-finite value; get_info must
    still be strict JSON (the value becomes null), since a polysolve caller parses it after every
    solve."""
    import json

    from polysolve_amd import _native

    p = problems.hex_elasticity(6, E=problems.random_moduli(6, 21))
    s = _solver({"lattice_points": 3, "learned_collect": True, "learned_random_pairs": 16})
    s.set_coordinates(p.coords)
    s.factorize(p.A)
    x = np.zeros(p.n)
    s.solve(p.b, x)
    loss = s.learned_train(epochs=3, batch=16, lr=1e30, seed=0)
    assert not np.isfinite(loss)

    def refuse(tok):
        raise ValueError(f"non-standard JSON constant {tok}")

    rc, text = _native.read_string(_native.lib().psmas_get_info, s._h)
    assert rc == _native.PSMAS_OK
    info = json.loads(text, parse_constant=refuse)
    assert info["learned"]["train_loss"] is None
