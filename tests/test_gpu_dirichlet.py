"""FEM Dirichlet elimination on the GPU (FEMSolver.cpp:97-342) against the oracle
(oracle/dirichlet_oracle.py): the eliminated matrix, the lifted right-hand side (bit for bit), the
solution, the values-only refresh for later Newton steps (host and device values), and the error
paths."""
import numpy as np
import pytest

import polysolve_amd as pa
from oracle import dirichlet_oracle as DO
from polysolve_amd import problems
from polysolve_amd.linear import DeviceArray, MASError, MASSolver, _as_csc

pytestmark = pytest.mark.gpu

NE = 12


@pytest.fixture(scope="module")
def bare():
    return problems.hex_elasticity(NE, E=problems.random_moduli(NE, 5), fix_bottom=False, name="bare12")


@pytest.fixture(scope="module")
def nodes(bare):
    rng = np.random.default_rng(7)
    face = problems.bottom_face_dofs(NE)
    extra = rng.choice(np.arange(face.size, bare.n), size=bare.n // 50, replace=False)
    return np.concatenate([face, extra, face[:3]]).astype(np.int32)


def _solver(bare, precision="fp64", lazy=False, device_setup=True):
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 3, "precision": precision, "lazy_partitioning": lazy,
                              "device_setup": device_setup,
                              "aggregation": "morton", "relative_tolerance": 1e-10, "absolute_tolerance": 0.0}})
    s.set_coordinates(bare.coords)
    return s


def _f(bare, nodes, seed=0):
    rng = np.random.default_rng(seed)
    f = bare.b.copy()
    f[nodes] = rng.uniform(-1e-3, 1e-3, nodes.size)  # prescribed displacements
    return f


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("device_setup", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dirichlet_factorize_matches_eliminated_matrix(gpu, bare, nodes, precision, device_setup):
    s = _solver(bare, precision, device_setup=device_setup)
    s.dirichlet_factorize(bare.A, nodes)
    info = s.get_info()
    assert info["dirichlet"]["active"] and info["dirichlet"]["fixed_dofs"] == np.unique(nodes).size
    assert ("device_matrix_build" in info["setup_ms"]) == device_setup
    ref = _solver(bare, precision, device_setup=False)
    At = DO.eliminate(bare.A, nodes)
    ref.factorize(At)
    r = np.random.default_rng(1).standard_normal(bare.n)
    assert _rel(s.apply_preconditioner(r), ref.apply_preconditioner(r)) <= 1e-12
    assert _rel(s.spmv(r), At @ r) <= 1e-13
    host = _solver(bare, precision, device_setup=False)
    host.dirichlet_factorize(bare.A, nodes)
    assert np.array_equal(s.spmv(r), host.spmv(r))  # the GPU build gives the host build's matrix
    assert info["dirichlet"] == host.get_info()["dirichlet"]  # the same coupling rows and entries
    assert ("dirichlet_elimination" in info["setup_ms"]) == device_setup


def test_lift_is_bit_exact_and_solution_matches_reduced_solve(gpu, bare, nodes):
    s = _solver(bare)
    s.dirichlet_factorize(bare.A, nodes)
    f = _f(bare, nodes)
    g = f.copy()
    u = np.zeros(bare.n)
    s.dirichlet_solve(g, u)
    assert np.array_equal(g, DO.lift(bare.A, nodes, f))
    assert s.get_info()["solver_status"] == "Reach relative tolerance"
    assert _rel(u, DO.reduced_solution(bare.A, nodes, f)) <= 1e-6
    assert np.abs(u[nodes] - f[nodes]).max() <= 1e-8 * np.abs(u).max()


def test_reference_prefactorized_sequence(gpu, bare, nodes):
    """The reference's prefactorize + dirichlet_solve_prefactorized sequence lifts with the
    overwritten A′, so g = f (tests/test_dirichlet_oracle.py); here that sequence is the Dirichlet
    factorisation followed by a plain solve of f, which gives A′⁻¹ f, while psmas_dirichlet_solve
    lifts with A and gives the constrained solution."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    s = _solver(bare)
    s.dirichlet_factorize(bare.A, nodes)
    f = _f(bare, nodes)
    At = DO.eliminate(bare.A, nodes)
    assert np.array_equal(DO.lift(At, nodes, f), f)
    u_seq = np.zeros(bare.n)
    s.solve(f.copy(), u_seq)
    assert _rel(u_seq, sla.spsolve(sp.csc_matrix(At), f)) <= 1e-6
    g = f.copy()
    u = np.zeros(bare.n)
    s.dirichlet_solve(g, u)
    assert _rel(u, DO.reduced_solution(bare.A, nodes, f)) <= 1e-6
    assert _rel(u_seq, u) > 1e-3


@pytest.mark.parametrize("device_setup", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_newton_steps_refresh_values_only(gpu, bare, nodes, precision, device_setup):
    rng = np.random.default_rng(3)
    d = rng.uniform(0.5, 2.0, bare.n)
    A2 = (bare.A.multiply(d[:, None]).multiply(d[None, :])).tocsc()
    A2.sort_indices()
    s = _solver(bare, precision, lazy=True, device_setup=device_setup)
    s.dirichlet_factorize(bare.A, nodes)
    assert s.get_info()["factorize_mode"] == "full"
    s.dirichlet_factorize(A2, nodes)
    assert s.get_info()["factorize_mode"] == "values"
    fresh = _solver(bare, precision, lazy=True)
    fresh.dirichlet_factorize(A2, nodes)
    r = rng.standard_normal(bare.n)
    assert _rel(s.apply_preconditioner(r), fresh.apply_preconditioner(r)) <= 1e-12
    f = _f(bare, nodes, 1)
    g, u = f.copy(), np.zeros(bare.n)
    s.dirichlet_solve(g, u)
    assert np.array_equal(g, DO.lift(A2, nodes, f))
    assert _rel(u, DO.reduced_solution(A2, nodes, f)) <= 1e-6
    # another fixed set or a plain factorisation takes the full path
    s.dirichlet_factorize(A2, nodes[: nodes.size // 2])
    assert s.get_info()["factorize_mode"] == "full"
    s.factorize(DO.eliminate(A2, nodes))
    assert s.get_info()["factorize_mode"] == "full" and not s.get_info()["dirichlet"]["active"]


def test_device_values_and_device_solve(gpu, bare, nodes):
    rng = np.random.default_rng(4)
    d = rng.uniform(0.5, 2.0, bare.n)
    A2 = (bare.A.multiply(d[:, None]).multiply(d[None, :])).tocsc()
    A2.sort_indices()
    s = _solver(bare, lazy=True)
    s.dirichlet_factorize(bare.A, nodes)
    vals = DeviceArray(A2.nnz, _as_csc(A2)[3])
    s.factorize_values_device(vals.ptr, A2.nnz)
    assert s.get_info()["factorize_mode"] == "values"
    f = _f(bare, nodes, 2)
    g_host, u_host = f.copy(), np.zeros(bare.n)
    s.dirichlet_solve(g_host, u_host)
    assert np.array_equal(g_host, DO.lift(A2, nodes, f))
    df, du = DeviceArray(bare.n, f), DeviceArray(bare.n)
    s.dirichlet_solve_device(df.ptr, du.ptr, bare.n)
    assert np.array_equal(df.download(), g_host)
    assert np.array_equal(du.download(), u_host)
    for a in (vals, df, du):
        a.free()


def test_free_function_helpers(gpu, bare, nodes):
    s = _solver(bare)
    f = _f(bare, nodes, 3)
    f1, u1 = f.copy(), np.zeros(bare.n)
    spec = pa.dirichlet_solve(s, bare.A, f1, nodes, u1)
    assert spec.shape == (4,)
    s2 = _solver(bare)
    pa.prefactorize(s2, bare.A, nodes)
    f2, u2 = f.copy(), np.zeros(bare.n)
    pa.dirichlet_solve_prefactorized(s2, bare.A, f2, nodes, u2)
    assert np.array_equal(f1, f2) and np.array_equal(f1, DO.lift(bare.A, nodes, f))
    assert _rel(u1, u2) <= 1e-8  # two factorisations: Galerkin sums use FP64 atomics
    with pytest.raises(MASError, match="not supported"):
        pa.dirichlet_solve(s, bare.A, f.copy(), nodes, np.zeros(bare.n), remove_zero_cols=True)


def test_dirichlet_errors(gpu, bare, nodes):
    s = _solver(bare)
    with pytest.raises(MASError, match="no Dirichlet factorisation"):
        s.dirichlet_solve(bare.b.copy(), np.zeros(bare.n))
    with pytest.raises(MASError, match="out of range"):
        s.dirichlet_factorize(bare.A, np.array([0, bare.n], dtype=np.int32))
    s.dirichlet_factorize(bare.A, nodes)
    with pytest.raises(MASError, match="size"):
        s.dirichlet_solve(np.zeros(bare.n - 1), np.zeros(bare.n - 1))
    s.factorize(DO.eliminate(bare.A, nodes))  # a plain factorisation ends the Dirichlet mode
    with pytest.raises(MASError, match="no Dirichlet factorisation"):
        s.dirichlet_solve(bare.b.copy(), np.zeros(bare.n))


@pytest.mark.parametrize("device_setup", [True, False])
def test_no_fixed_dofs_and_values_refresh(gpu, device_setup):
    """An empty fixed set (no couplings) on a non-singular system (bottom face fixed in the
    assembly) factorises, refreshes its values and solves like a plain factorisation."""
    p = problems.hex_elasticity(8, E=problems.random_moduli(8, 6), name="fixed8")
    d = np.random.default_rng(6).uniform(0.5, 2.0, p.n)
    A2 = (p.A.multiply(d[:, None]).multiply(d[None, :])).tocsc()
    A2.sort_indices()
    empty = np.zeros(0, dtype=np.int32)
    s = _solver(p, lazy=True, device_setup=device_setup)
    s.dirichlet_factorize(p.A, empty)
    assert s.get_info()["dirichlet"]["couplings"] == 0
    s.dirichlet_factorize(A2, empty)
    assert s.get_info()["factorize_mode"] == "values"
    g, u = p.b.copy(), np.zeros(p.n)
    s.dirichlet_solve(g, u)
    assert np.array_equal(g, p.b)
    assert np.linalg.norm(p.b - A2 @ u) <= 1e-9 * np.linalg.norm(p.b)


def test_device_input_dirichlet_factorisation(gpu, bare, nodes):
    """A's CSC arrays in device memory: the same eliminated matrix as the host-input call, the same
    lift, and Newton values from device memory."""
    n, colptr, rowidx, vals = _as_csc(bare.A)
    dc, dr, dv = DeviceArray(n + 1, colptr, np.int32), DeviceArray(rowidx.size, rowidx, np.int32), \
        DeviceArray(vals.size, vals)
    s = _solver(bare, lazy=True)
    s.dirichlet_factorize_device(dc.ptr, dr.ptr, dv.ptr, n, rowidx.size, nodes)
    host = _solver(bare, lazy=True)
    host.dirichlet_factorize(bare.A, nodes)
    assert s.get_info()["dirichlet"] == host.get_info()["dirichlet"]
    r = np.random.default_rng(8).standard_normal(n)
    assert np.array_equal(s.spmv(r), host.spmv(r))
    f = _f(bare, nodes, 4)
    g, u = f.copy(), np.zeros(n)
    s.dirichlet_solve(g, u)
    assert np.array_equal(g, DO.lift(bare.A, nodes, f))
    assert _rel(u, DO.reduced_solution(bare.A, nodes, f)) <= 1e-6
    d = np.random.default_rng(9).uniform(0.5, 2.0, n)
    A2 = (bare.A.multiply(d[:, None]).multiply(d[None, :])).tocsc()
    A2.sort_indices()
    dv.upload(_as_csc(A2)[3])
    s.factorize_values_device(dv.ptr, rowidx.size)
    assert s.get_info()["factorize_mode"] == "values"
    g, u = f.copy(), np.zeros(n)
    s.dirichlet_solve(g, u)
    assert np.array_equal(g, DO.lift(A2, nodes, f))
    bad = rowidx.copy()
    bad[5] = n
    dr.upload(bad)
    with pytest.raises(MASError, match="row index out of range"):
        s.dirichlet_factorize_device(dc.ptr, dr.ptr, dv.ptr, n, rowidx.size, nodes)
    for a in (dc, dr, dv):
        a.free()
