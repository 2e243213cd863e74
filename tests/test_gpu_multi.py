"""Several right-hand sides of one matrix (factorize_multi / solve_multi): one factorisation (the
single system's bank matrices, coarse map and coarsest inverse) serves every right-hand side;
every column solved to tolerance and against a direct solve; the copies SpMV (matrix read once for
all right-hand sides) gives the iterates of one SpMV launch per copy bit for bit; device arrays
give the host result; each column matches a solve of that right-hand side alone."""
import numpy as np
import pytest

from polysolve_amd import problems
from polysolve_amd.linear import DeviceArray, MASError, MASSolver, _as_csc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def het12():
    return problems.hex_elasticity(12, E=problems.random_moduli(12, 3), name="het12")


def _prm(precision):
    # The copies of a multi-RHS factorisation read the full SELL-64 form (the symmetric form is for
    # one matrix per launch): single solves compared with them use the same form.
    return {"MAS": {"block_dim": 3, "precision": precision, "relative_tolerance": 1e-8, "absolute_tolerance": 0.0,
                    "coarse_solver": "cholesky", "spmv_symmetric": False}}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("k", [3, 8])
def test_multi_rhs(gpu, het12, precision, k):
    rng = np.random.default_rng(k)
    B = rng.standard_normal((het12.n, k))
    B[:, 0] = het12.b
    s = MASSolver()
    s.set_parameters(_prm(precision))
    s.set_coordinates(het12.coords)
    s.factorize_multi(het12.A, k)
    assert s.get_info()["copies"] == k
    X = s.solve_multi(B)
    for c in range(k):
        assert s.batch_info(c)["solver_status"] == "Reach relative tolerance"
        assert np.linalg.norm(B[:, c] - het12.A @ X[:, c]) <= 1e-8 * np.linalg.norm(B[:, c])
    s.set_parameters({"MAS": {"spmv_copies": False}})
    assert np.array_equal(s.solve_multi(B), X)
    s.set_parameters({"MAS": {"spmv_copies": True}})
    dB, dX = DeviceArray(het12.n * k, B.ravel(order="F")), DeviceArray(het12.n * k)
    s.solve_multi_device(dB.ptr, dX.ptr, het12.n, k)
    assert np.array_equal(dX.download().reshape((het12.n, k), order="F"), X)
    dB.free()
    dX.free()
    one = MASSolver()
    one.set_parameters(_prm(precision))
    one.set_coordinates(het12.coords)
    one.factorize(het12.A)
    for c in (0, k - 1):
        x1 = np.zeros(het12.n)
        one.solve(np.ascontiguousarray(B[:, c]), x1)
        assert np.linalg.norm(X[:, c] - x1) <= 1e-6 * np.linalg.norm(x1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("k", [4, 6])
def test_multi_rhs_one_factorisation(gpu, het12, precision, k):
    """factorize_multi keeps the single system's factorisation (bank matrices, coarse map) once,
    not per right-hand side.  Up to 4 right-hand sides the solve does the single system's
    arithmetic per copy, so each column takes exactly the iterations of that right-hand side
    solved alone; from 5 on the mixed path's bank products run on MFMA (inputs split into FP16
    hi + lo parts, FP32 accumulation in another order), within one check interval of them.  Every
    column matches a direct solve."""
    import scipy.sparse.linalg as spla

    B = np.random.default_rng(21).standard_normal((het12.n, k))
    one = MASSolver()
    one.set_parameters(_prm(precision))
    one.set_coordinates(het12.coords)
    one.factorize(het12.A)
    s = MASSolver()
    s.set_parameters(_prm(precision))
    s.set_coordinates(het12.coords)
    s.factorize_multi(het12.A, k)
    li1, lik = one.level_info(), s.level_info()
    assert li1["level_num"] >= 2  # (coarse segments 2.. and their restriction lists are exercised)
    assert lik["bank_matrices"] == li1["bank_matrices"]
    assert np.array_equal(s.coarse_map(), one.coarse_map())
    M1 = one.bank_inverses(3)
    assert np.linalg.norm(s.bank_inverses(3) - M1) <= 1e-12 * np.linalg.norm(M1)
    assert "expand_copies" in s.get_info()["setup_ms"]
    X = s.solve_multi(B)
    lu = spla.splu(het12.A.tocsc())
    for c in range(k):
        info = s.batch_info(c)
        assert info["solver_status"] == "Reach relative tolerance"
        xd = lu.solve(B[:, c])
        assert np.linalg.norm(X[:, c] - xd) <= 1e-6 * np.linalg.norm(xd)
        x1 = np.zeros(het12.n)
        one.solve(np.ascontiguousarray(B[:, c]), x1)
        if k <= 4 or precision == "fp64":
            assert info["solver_iter"] == one.get_info()["solver_iter"], c
            assert np.linalg.norm(X[:, c] - x1) <= 1e-7 * np.linalg.norm(x1)
        else:
            assert abs(info["solver_iter"] - one.get_info()["solver_iter"]) <= 10, c
            assert np.linalg.norm(X[:, c] - x1) <= 1e-6 * np.linalg.norm(x1)


def test_multi_rhs_errors(gpu, het12):
    s = MASSolver()
    s.set_parameters(_prm("mixed"))
    with pytest.raises(MASError, match="right-hand sides"):
        s.factorize_multi(het12.A, 9)
    s.factorize_multi(het12.A, 2)
    with pytest.raises(MASError, match="nrhs"):
        s.solve_multi(np.zeros((het12.n, 3)))


@pytest.mark.parametrize("coords", [True, False])
def test_multi_rhs_replicated_device_build(gpu, het12, coords):
    """factorize_multi builds the matrix on the GPU (or on the host with device_setup off) once for
    all right-hand sides; both builds give the same solutions (to the FP64 Galerkin's rounding)."""
    k = 4
    B = np.random.default_rng(7).standard_normal((het12.n, k))
    out = []
    for device_setup in (True, False):
        s = MASSolver()
        s.set_parameters({"MAS": {**_prm("mixed")["MAS"], "device_setup": device_setup}})
        if coords:
            s.set_coordinates(het12.coords)
        s.factorize_multi(het12.A, k)
        info = s.get_info()
        assert info["copies"] == k
        assert ("device_matrix_build" in info["setup_ms"]) == device_setup
        out.append((s.solve_multi(B), [s.batch_info(c)["solver_iter"] for c in range(k)]))
    (Xd, itd), (Xh, ith) = out
    for c in range(k):
        assert np.linalg.norm(B[:, c] - het12.A @ Xd[:, c]) <= 1e-8 * np.linalg.norm(B[:, c])
        if coords:  # the same domains on both sides
            assert abs(itd[c] - ith[c]) <= 10
            assert np.linalg.norm(Xd[:, c] - Xh[:, c]) <= 1e-6 * np.linalg.norm(Xh[:, c])


def test_multi_rhs_device_input(gpu, het12):
    """factorize_multi from CSC arrays in device memory solves like the host-input call."""
    k = 3
    B = np.random.default_rng(11).standard_normal((het12.n, k))
    n, colptr, rowidx, vals = _as_csc(het12.A)
    dc, dr, dv = DeviceArray(n + 1, colptr, np.int32), DeviceArray(rowidx.size, rowidx, np.int32), \
        DeviceArray(vals.size, vals)
    s = MASSolver()
    s.set_parameters(_prm("mixed"))
    s.set_coordinates(het12.coords)
    s.factorize_multi_device(dc.ptr, dr.ptr, dv.ptr, n, rowidx.size, k)
    assert s.get_info()["copies"] == k
    X = s.solve_multi(B)
    ref = MASSolver()
    ref.set_parameters(_prm("mixed"))
    ref.set_coordinates(het12.coords)
    ref.factorize_multi(het12.A, k)
    Xr = ref.solve_multi(B)
    for c in range(k):
        assert np.linalg.norm(B[:, c] - het12.A @ X[:, c]) <= 1e-8 * np.linalg.norm(B[:, c])
        assert np.linalg.norm(X[:, c] - Xr[:, c]) <= 1e-6 * np.linalg.norm(Xr[:, c])
    with pytest.raises(MASError, match="right-hand sides"):
        s.factorize_multi_device(dc.ptr, dr.ptr, dv.ptr, n, rowidx.size, 0)
    for a in (dc, dr, dv):
        a.free()


@pytest.mark.parametrize("k", [3, 8])
def test_multi_rhs_apply_cap(gpu, het12, k):
    """apply_workgroups = 8: every apply workgroup walks many stored matrices (each applied to
    every copy's bank), so the grid-stride loop runs across many matrices and levels.  The
    iterates equal the uncapped launch's bit for bit, and every column matches a direct solve."""
    import scipy.sparse.linalg as spla

    rng = np.random.default_rng(40 + k)
    B = rng.standard_normal((het12.n, k))
    out = []
    for cap in (0, 8):
        s = MASSolver()
        s.set_parameters({"MAS": {**_prm("mixed")["MAS"], "apply_workgroups": cap}})
        s.set_coordinates(het12.coords)
        s.factorize_multi(het12.A, k)
        out.append((s.solve_multi(B), [s.batch_info(c)["solver_iter"] for c in range(k)]))
    (X0, it0), (X8, it8) = out
    assert it0 == it8
    assert np.array_equal(X0, X8)
    lu = spla.splu(het12.A.tocsc())
    for c in range(k):
        xd = lu.solve(B[:, c])
        assert np.linalg.norm(X8[:, c] - xd) <= 1e-6 * np.linalg.norm(xd)
