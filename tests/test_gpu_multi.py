"""Several right-hand sides of one matrix (factorize_multi / solve_multi): every column solved to
tolerance; the copies SpMV (matrix read once for all right-hand sides) gives the iterates of the
batch's per-system SpMV bit for bit; device arrays give the host result; each column matches a
solve of that right-hand side alone."""
import numpy as np
import pytest

from polysolve_amd import problems
from polysolve_amd.linear import DeviceArray, MASError, MASSolver, _as_csc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def het12():
    return problems.hex_elasticity(12, E=problems.random_moduli(12, 3), name="het12")


def _prm(precision):
    return {"MAS": {"block_dim": 3, "precision": precision, "relative_tolerance": 1e-8, "absolute_tolerance": 0.0,
                    "coarse_solver": "cholesky"}}


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
    """factorize_multi builds system 0's matrix on the GPU once and copies it to the other right-hand
    sides; the host build of every copy gives the same solutions (to the FP64 Galerkin's rounding)."""
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
    """apply_workgroups = 8: every apply workgroup walks many matrices, so the shared fine-bank
    loop (a copy-0 bank stands for every copy's bank, the other copies' fine banks are stepped
    over, coarse banks go one by one) runs its skip and step logic across many matrices.  The
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
