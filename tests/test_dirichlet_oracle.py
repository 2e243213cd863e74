"""The Dirichlet-elimination oracle (oracle/dirichlet_oracle.py) pinned by what the mathematics
fixes: the constrained solve by index slicing, the generator's own Dirichlet-fixed assembly, and a
planted sign error that these checks must catch.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from oracle import dirichlet_oracle as DO
from polysolve_amd import problems


def _bare(ne, seed=0):
    return problems.hex_elasticity(ne, E=problems.random_moduli(ne, seed), fix_bottom=False)


def _fixed_set(ne, seed=1):
    rng = np.random.default_rng(seed)
    face = problems.bottom_face_dofs(ne)
    n = 3 * (ne + 1) ** 3
    extra = rng.choice(np.arange(face.size, n), size=n // 20, replace=False)  # single components
    return np.concatenate([face, extra, face[:5]])  # duplicates are allowed


def _solution_error(A, nodes, f, g):
    At = DO.eliminate(A, nodes)
    u = sla.spsolve(sp.csc_matrix(At), g)
    ref = DO.reduced_solution(A, nodes, f)
    return np.linalg.norm(u - ref) / np.linalg.norm(ref)


def test_eliminated_matrix_is_generator_dirichlet_assembly():
    ne = 4
    fixed = problems.hex_elasticity(ne)
    At = DO.eliminate(problems.hex_elasticity(ne, fix_bottom=False).A, problems.bottom_face_dofs(ne))
    assert np.array_equal(At.indptr, fixed.A.indptr)
    assert np.array_equal(At.indices, fixed.A.indices)
    assert np.array_equal(At.data, fixed.A.data)


def test_solution_matches_reduced_solve():
    ne = 3
    p = _bare(ne)
    nodes = _fixed_set(ne)
    rng = np.random.default_rng(2)
    f = rng.standard_normal(p.n)
    g = DO.lift(p.A, nodes, f)
    assert _solution_error(p.A, nodes, f, g) < 1e-10
    N = DO.fixed_mask(p.n, nodes)
    assert np.array_equal(g[N], f[N])
    At = DO.eliminate(p.A, nodes)
    assert abs(At - At.T).max() == 0.0
    assert np.array_equal(At[N][:, N].toarray(), np.eye(int(N.sum())))


def test_planted_sign_error_is_caught():
    ne = 3
    p = _bare(ne)
    nodes = _fixed_set(ne)
    f = np.random.default_rng(3).standard_normal(p.n)
    N = DO.fixed_mask(p.n, nodes).astype(float)
    wrong = f + (1.0 - N) * (p.A @ (N * f))  # sign of the lift flipped
    assert _solution_error(p.A, nodes, f, wrong) > 1e-3
    halved = f - 0.5 * (1.0 - N) * (p.A @ (N * f))  # half the coupling
    assert _solution_error(p.A, nodes, f, halved) > 1e-3
    dropped = f.copy()  # no lift at all
    assert _solution_error(p.A, nodes, f, dropped) > 1e-3


def test_lift_equals_sparse_product_form():
    p = _bare(3)
    nodes = _fixed_set(3)
    f = np.random.default_rng(4).standard_normal(p.n)
    N = DO.fixed_mask(p.n, nodes).astype(float)
    dense_form = f - (1.0 - N) * (p.A @ (N * f))
    g = DO.lift(p.A, nodes, f)
    assert np.allclose(g, dense_form, rtol=1e-14, atol=1e-14 * np.abs(f).max())


def test_reference_prefactorized_sequence_lifts_nothing():
    """The reference's call sequence: prefactorize (FEMSolver.cpp:268-308) overwrites the caller's
    A with the eliminated A′, and dirichlet_solve_prefactorized (FEMSolver.cpp:317-342) then lifts
    with the A it is given, i.e. A′, whose free rows have no fixed columns: g = f exactly, and the
    solve ignores the couplings to the prescribed values (u differs from the constrained
    solution).  Lifting with the A before elimination, as psmas_dirichlet_solve does and the
    one-shot dirichlet_solve (FEMSolver.cpp:97-137) does, gives the constrained solution."""
    p = _bare(3)
    nodes = _fixed_set(3)
    f = np.random.default_rng(5).standard_normal(p.n)
    At = DO.eliminate(p.A, nodes)
    g_seq = DO.lift(At, nodes, f)
    assert np.array_equal(g_seq, f)
    u_seq = sla.spsolve(sp.csc_matrix(At), g_seq)
    ref = DO.reduced_solution(p.A, nodes, f)
    assert np.linalg.norm(u_seq - ref) > 1e-3 * np.linalg.norm(ref)
    assert _solution_error(p.A, nodes, f, DO.lift(p.A, nodes, f)) <= 1e-10


def test_missing_diagonal_gets_explicit_zero():
    # 4x4: column 2 has no diagonal; dof 1 fixed.
    dense = np.array([[4.0, 1.0, 2.0, 0.0],
                      [1.0, 5.0, 0.0, 3.0],
                      [2.0, 0.0, 0.0, 1.0],
                      [0.0, 3.0, 1.0, 6.0]])
    A = sp.csc_matrix(dense)
    At = DO.eliminate(A, [1])
    col2 = At.indices[At.indptr[2]:At.indptr[3]]
    assert list(col2) == [0, 2, 3] and At.data[At.indptr[2] + 1] == 0.0
    assert list(At.indices[At.indptr[1]:At.indptr[2]]) == [1] and At.data[At.indptr[1]] == 1.0
    assert At.nnz == A.nnz - 5 + 2  # row and column of dof 1 dropped; two diagonals added
    g = DO.lift(A, [1], np.array([1.0, 2.0, 3.0, 4.0]))
    assert np.array_equal(g, [1.0 - 1.0 * 2.0, 2.0, 3.0, 4.0 - 3.0 * 2.0])


@pytest.mark.parametrize("ne", [2, 3])
def test_fixed_face_makes_bare_stiffness_solvable(ne):
    p = _bare(ne)
    nodes = problems.bottom_face_dofs(ne)
    At = DO.eliminate(p.A, nodes)
    lam_min = np.linalg.eigvalsh(At.toarray()).min()
    assert lam_min > 0.0
