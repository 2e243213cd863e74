"""The CPU direct solver that stands in for CHOLMOD (oracle/sparse_cholesky.py): the factor
reproduces the permuted matrix entry for entry, solutions match dense and SuperLU solves, with
coordinate and graph orderings, disconnected graphs and a non-multiple-of-block size; an
indefinite matrix is refused."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.sparse_cholesky import SparseCholesky
from polysolve_amd import problems
from tests import helpers


def _check_factor(A, S, tol=1e-13):
    L = S.dense_factor()
    Ap = A.toarray()[np.ix_(S.perm, S.perm)]
    assert np.abs(L - np.tril(L)).max() == 0.0
    assert np.abs(L @ L.T - Ap).max() <= tol * np.abs(Ap).max()


@pytest.mark.parametrize("use_coords", [True, False])
def test_factor_reproduces_hex_matrix(use_coords):
    p = problems.hex_elasticity(4, E=problems.random_moduli(4, 9))
    S = SparseCholesky(p.A, 3, p.coords if use_coords else None, leaf=6).factorize()
    assert S.supernodes > 5  # nested dissection actually split the graph
    _check_factor(p.A, S)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(p.n)
    x = S.solve(b)
    xd = np.linalg.solve(p.A.toarray(), b)
    assert np.linalg.norm(x - xd) <= 1e-12 * np.linalg.norm(xd)


def test_graph_ordering_scalar_and_disconnected():
    pat = helpers.grid_laplacian_pattern(9, 7, 3)
    A1 = helpers.prefactor_matrix(pat, seed=3)
    A2 = helpers.prefactor_matrix(helpers.grid_laplacian_pattern(5, 4, 2), seed=4)
    A = sp.block_diag([A1, A2]).tocsc()  # two components
    S = SparseCholesky(A, 1, None, leaf=10).factorize()
    _check_factor(A, S)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x = S.solve(b)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)


def test_block_size_not_dividing_n_falls_back_to_scalar_nodes():
    A = helpers.prefactor_matrix(helpers.grid_laplacian_pattern(11, 7, 3), seed=5)  # n = 231 = 3 * 77
    A = A[:-1, :-1].tocsc()  # 230: not a multiple of 3
    S = SparseCholesky(A, 3, None, leaf=8).factorize()
    assert S.bd == 1
    _check_factor(A, S)


def test_matches_superlu_on_16cubed_elasticity():
    p = problems.config1()
    S = SparseCholesky(p.A, 3, p.coords).factorize()
    x = S.solve(p.b)
    xd = spla.spsolve(p.A.tocsc(), p.b)
    assert np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)
    assert np.linalg.norm(p.b - p.A @ x) <= 1e-12 * np.linalg.norm(p.b)
    assert S.nnz_L > p.A.nnz // 2 and S.flops > 0


def test_indefinite_matrix_refused():
    A = helpers.prefactor_matrix(helpers.grid_laplacian_pattern(6, 6, 1), seed=2).tolil()
    A[7, 7] = -50.0
    S = SparseCholesky(A.tocsc(), 1, None, leaf=4)
    with pytest.raises(np.linalg.LinAlgError, match="not positive definite"):
        S.factorize()
