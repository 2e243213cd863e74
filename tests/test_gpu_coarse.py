"""GPU: the exact FP64 coarsest-level solve (coarse_solver = "cholesky")."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from polysolve_amd import MASSolver, problems
from tests import helpers

pytestmark = pytest.mark.gpu


def _solver(p, **extra):
    s = MASSolver()
    prm = {"block_dim": 3, "aggregation": "graph", "coarse_solver": "cholesky"}
    prm.update(extra)
    s.set_parameters({"MAS": prm})
    s.factorize(p.A)
    return s


def test_dense_coarsest_apply_matches_numpy_construction(gpu):
    p = problems.hex_elasticity(5, E=problems.random_moduli(5, 21))
    s = _solver(p)
    offsets, perm = s.partition()
    li = s.level_info()
    cm = s.coarse_map()
    cm_perm = np.empty_like(cm)
    cm_perm[perm] = cm
    M = helpers.dense_mas(p.A.toarray(), 3, offsets, perm, cm_perm, li["level_num"], coarsest_dense=True)
    rng = np.random.default_rng(5)
    for _ in range(2):
        r = rng.standard_normal(p.n)
        z = s.apply_preconditioner(r)
        assert np.linalg.norm(z - M @ r) <= 1e-9 * np.linalg.norm(M @ r)
    assert "Cholesky" in s.get_info()["coarse_solver"]
    # Symmetric positive definite on this space.
    assert np.linalg.eigvalsh((M + M.T) / 2).min() > 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dense_coarsest_solve_converges(gpu, precision):
    p = problems.hex_elasticity(12, E=problems.random_moduli(12, 3))
    s = _solver(p, precision=precision, relative_tolerance=1e-8, absolute_tolerance=0.0, max_iter=5000)
    s.set_coordinates(p.coords)
    x = np.zeros(p.n)
    s.solve(p.b, x)
    info = s.get_info()
    assert info["solver_status"] == "Reach relative tolerance", info
    assert np.linalg.norm(p.b - p.A @ x) <= 2e-8 * np.linalg.norm(p.b)
    xd = spla.spsolve(p.A.tocsc(), p.b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
