"""Batches of uneven systems (the 64-system workload's shape at test size) and the autotuner."""
from __future__ import annotations

import numpy as np
import pytest

from polysolve_amd import MASSolver, problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uneven_batch():
    sizes = problems.config4_sizes(seed=1, count=12, lo=3, hi=10)
    return [problems.hex_elasticity(int(e), E=problems.random_moduli(int(e), 200 + i), name=f"b{i}")
            for i, e in enumerate(sizes)]


def _batch_solver(probs, precision):
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 3, "precision": precision, "relative_tolerance": 1e-8,
                              "absolute_tolerance": 0.0, "max_iter": 5000, "coarse_solver": "cholesky"}})
    s.factorize_batch([p.A for p in probs], coords=[p.coords for p in probs])
    return s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_uneven_batch_each_system_to_tolerance(gpu, uneven_batch, precision):
    s = _batch_solver(uneven_batch, precision)
    xs = [np.zeros(p.n) for p in uneven_batch]
    s.solve_batch([p.b for p in uneven_batch], xs)
    iters = []
    for i, p in enumerate(uneven_batch):
        info = s.batch_info(i)
        assert info["solver_status"] == "Reach relative tolerance", (i, info)
        assert np.linalg.norm(p.b - p.A @ xs[i]) <= 1e-8 * np.linalg.norm(p.b)
        # Each system alone, same kernels: the same solution to solver accuracy.
        one = MASSolver()
        one.set_parameters({"MAS": {"block_dim": 3, "precision": precision, "relative_tolerance": 1e-8,
                                    "absolute_tolerance": 0.0, "max_iter": 5000, "coarse_solver": "cholesky"}})
        one.set_coordinates(p.coords)
        one.factorize(p.A)
        x1 = np.zeros(p.n)
        one.solve(p.b, x1)
        assert np.linalg.norm(xs[i] - x1) <= 1e-6 * np.linalg.norm(x1)
        iters.append(info["solver_iter"])
    # Systems stop at their own convergence checks, not at the slowest one's.
    assert min(iters) < max(iters)


def test_autotune_picks_a_valid_configuration(gpu, uneven_batch):
    s = _batch_solver(uneven_batch, "mixed")
    xs = [np.zeros(p.n) for p in uneven_batch]
    s.solve_batch([p.b for p in uneven_batch], xs)
    rep = s.autotune(reps=5)
    assert set(rep["spmv_us"]) == {"2", "4", "8"} and all(v > 0 for v in rep["spmv_us"].values())
    assert rep["spmv_unroll"] in (2, 4, 8)
    assert len(rep["solve_ms"]) == 6 and rep["best_solve_ms"] > 0
    prm = s.get_parameters()
    assert prm["spmv_unroll"] == rep["spmv_unroll"] and prm["check_interval"] == rep["check_interval"]
    assert prm["coarse_solver"] == rep["coarse_solver"]
    # The tuned solver still solves every system, with the same results to tolerance.
    ys = [np.zeros(p.n) for p in uneven_batch]
    s.solve_batch([p.b for p in uneven_batch], ys)
    for i, p in enumerate(uneven_batch):
        assert s.batch_info(i)["solver_status"] == "Reach relative tolerance"
        assert np.linalg.norm(ys[i] - xs[i]) <= 1e-6 * np.linalg.norm(xs[i])


@pytest.mark.parametrize("unroll", [2, 8])
def test_spmv_unroll_variants_match(gpu, unroll):
    p = problems.hex_elasticity(9, E=problems.random_moduli(9, 5))
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 3, "spmv_unroll": unroll}})
    s.set_coordinates(p.coords)
    s.factorize(p.A)
    x = np.random.default_rng(unroll).standard_normal(p.n)
    y = s.spmv(x)
    ref = p.A @ x
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    y32 = s.spmv(x, fp32=True)
    assert np.linalg.norm(y32 - ref) <= 2e-6 * np.linalg.norm(ref)


@pytest.mark.parametrize("coords", [True, False])
def test_batch_device_build_matches_host_build(gpu, uneven_batch, coords):
    """The batch's matrices built on the GPU system by system and joined (device_setup) against the
    host build of the whole batch: the same layout and, on the same domains, the same solutions."""
    out = []
    for device_setup in (True, False):
        s = MASSolver()
        s.set_parameters({"MAS": {"block_dim": 3, "precision": "fp64", "relative_tolerance": 1e-8,
                                  "absolute_tolerance": 0.0, "max_iter": 5000, "device_setup": device_setup}})
        s.factorize_batch([p.A for p in uneven_batch],
                          coords=[p.coords for p in uneven_batch] if coords else None)
        assert ("device_matrix_build" in s.get_info()["setup_ms"]) == device_setup
        xs = [np.zeros(p.n) for p in uneven_batch]
        s.solve_batch([p.b for p in uneven_batch], xs)
        for i, p in enumerate(uneven_batch):
            assert np.linalg.norm(p.b - p.A @ xs[i]) <= 1e-8 * np.linalg.norm(p.b)
        out.append((s.level_info(), xs, [s.batch_info(i)["solver_iter"] for i in range(len(uneven_batch))]))
    (ld, xd, itd), (lh, xh, ith) = out
    if coords:  # the same domains on both sides
        # (slot counts may differ: the GPU build gives every 64-row slice a slot, empty ones too)
        assert ld["nnzb"] == lh["nnzb"] and ld["npad"] == lh["npad"]
        for i in range(len(xd)):
            assert abs(itd[i] - ith[i]) <= 10
            assert np.linalg.norm(xd[i] - xh[i]) <= 1e-6 * np.linalg.norm(xh[i])
