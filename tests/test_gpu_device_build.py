"""The SELL matrix built on the GPU (sell_build.hip) against the host build (host_setup.cpp
build_padded_bsr): the same SpMV bit for bit, partition and coarse map, the same preconditioner to
rounding, and the factorisation from CSC arrays in device memory (validation included)."""
import numpy as np
import pytest
import scipy.sparse as sp

from polysolve_amd import problems
from polysolve_amd.linear import DeviceArray, MASError, MASSolver, _as_csc
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def het10():
    return problems.hex_elasticity(10, E=problems.random_moduli(10, 2), name="het10")


def _solver(bd, device_setup, coords=None, **extra):
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": bd, "device_setup": device_setup, "relative_tolerance": 1e-10,
                              "absolute_tolerance": 0.0, **extra}})
    if coords is not None:
        s.set_coordinates(coords)
    return s


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _compare(A, bd, coords=None, **extra):
    if coords is None:  # the same Morton domains on both sides (block index as a 1-D coordinate)
        coords = np.arange(-(-A.shape[0] // bd), dtype=np.float64)[:, None]
    dev, host = _solver(bd, True, coords, **extra), _solver(bd, False, coords, **extra)
    dev.factorize(A)
    host.factorize(A)
    assert "device_matrix_build" in dev.get_info()["setup_ms"]
    assert "padded_bsr" in host.get_info()["setup_ms"]
    rng = np.random.default_rng(bd)
    x = rng.standard_normal(A.shape[0])
    assert np.array_equal(dev.spmv(x), host.spmv(x))
    assert np.array_equal(dev.spmv(x, fp32=True), host.spmv(x, fp32=True))
    for a, b in zip(dev.partition(), host.partition()):
        assert np.array_equal(a, b)
    assert np.array_equal(dev.coarse_map(), host.coarse_map())
    assert _rel(dev.apply_preconditioner(x), host.apply_preconditioner(x)) <= 1e-12
    return dev, host


def test_hex_morton(gpu, het10):
    dev, host = _compare(het10.A, 3, het10.coords)
    b = het10.b
    xd, xh = np.zeros(het10.n), np.zeros(het10.n)
    dev.solve(b, xd)
    host.solve(b, xh)
    assert abs(dev.get_info()["solver_iter"] - host.get_info()["solver_iter"]) <= 10
    assert _rel(xd, xh) <= 1e-6


@pytest.mark.parametrize("ne", [10, 16])
def test_gpu_graph_aggregation(gpu, ne):
    """Without coordinates the device path aggregates on the GPU: domains of 1..32 nodes covering
    every node once, reproducible, no more banks than the host aggregation and a preconditioner
    as good (iterations) on the same solve."""
    p = problems.hex_elasticity(ne, E=problems.random_moduli(ne, 4))
    nb = p.n // 3
    dev, host = _solver(3, True), _solver(3, False)
    dev.factorize(p.A)
    host.factorize(p.A)
    assert dev.get_info()["aggregation"] == "graph_gpu" and host.get_info()["aggregation"] == "graph"
    offsets, perm = dev.partition()
    sizes = np.diff(offsets)  # banks rounded up to a multiple of 8: the trailing ones are empty
    real = sizes[:np.count_nonzero(sizes)]
    assert real.min() >= 1 and sizes.max() <= 32 and real.sum() == nb and sizes.size - real.size < 8
    assert np.array_equal(np.sort(perm), np.arange(nb))
    again = _solver(3, True)
    again.factorize(p.A)
    for a, b in zip(dev.partition(), again.partition()):
        assert np.array_equal(a, b)
    assert dev.get_info()["parts"] <= host.get_info()["parts"]
    xd, xh = np.zeros(p.n), np.zeros(p.n)
    dev.solve(p.b, xd)
    host.solve(p.b, xh)
    assert dev.get_info()["solver_iter"] <= 1.25 * host.get_info()["solver_iter"] + 10
    assert _rel(xd, xh) <= 1e-6


@pytest.mark.parametrize("bd", [1, 2, 3])
def test_padded_trailing_block_and_block_sizes(gpu, bd):
    A = helpers.prefactor_matrix(helpers.grid_laplacian_pattern(11, 7, 3), seed=5)  # 231 unknowns
    _compare(A, bd)


def test_duplicate_entries_are_summed(gpu):
    A = helpers.prefactor_matrix(helpers.grid_laplacian_pattern(9, 8, 4), seed=3).tocsc()
    A.sort_indices()
    # every stored entry twice (halves), columns unsorted: scipy keeps both copies in CSC form
    half = A.data * 0.5
    n = A.shape[0]
    colptr = A.indptr * 2
    rows = np.empty(2 * A.nnz, dtype=np.int32)
    vals = np.empty(2 * A.nnz)
    for c in range(n):
        lo, hi = A.indptr[c], A.indptr[c + 1]
        rows[2 * lo:2 * hi] = np.concatenate([A.indices[lo:hi][::-1], A.indices[lo:hi]])
        vals[2 * lo:2 * hi] = np.concatenate([half[lo:hi][::-1], half[lo:hi]])
    D = sp.csc_matrix((vals, rows, colptr), shape=(n, n))
    D.has_sorted_indices = False
    x = np.random.default_rng(0).standard_normal(n)
    coords = np.arange(-(-n // 3), dtype=np.float64)[:, None]  # one partition for all three
    ys = []
    for device_setup in (True, False):  # host input: GPU build, host build
        s = _solver(3, device_setup, coords)
        s.factorize(D)
        ys.append(s.spmv(x))
    s = _solver(3, True, coords)  # device input
    dv = [DeviceArray(colptr.size, colptr, np.int32), DeviceArray(rows.size, rows, np.int32),
          DeviceArray(vals.size, vals)]
    s.factorize_device(dv[0].ptr, dv[1].ptr, dv[2].ptr, n, rows.size)
    ys.append(s.spmv(x))
    for d in dv:
        d.free()
    for y in ys:
        assert _rel(y, A @ x) <= 1e-14
        assert np.array_equal(y, ys[1])  # two copies per entry: the sum does not depend on the order


def test_device_input_factorisation_and_values(gpu, het10):
    n, colptr, rowidx, vals = _as_csc(het10.A)
    dc, dr, dv = DeviceArray(n + 1, colptr, np.int32), DeviceArray(rowidx.size, rowidx, np.int32), \
        DeviceArray(vals.size, vals)
    s = _solver(3, True, het10.coords, lazy_partitioning=True, precision="mixed", coarse_solver="cholesky")
    s.factorize_device(dc.ptr, dr.ptr, dv.ptr, n, rowidx.size)
    ref = _solver(3, False, het10.coords, lazy_partitioning=True, precision="mixed", coarse_solver="cholesky")
    ref.factorize(het10.A)
    x = np.random.default_rng(1).standard_normal(n)
    assert np.array_equal(s.spmv(x), ref.spmv(x))
    xs, xr = np.zeros(n), np.zeros(n)
    s.solve(het10.b, xs)
    ref.solve(het10.b, xr)
    assert _rel(xs, xr) <= 1e-6
    # a Newton step with new values on the same pattern, from device memory
    d = np.random.default_rng(2).uniform(0.5, 2.0, n)
    A2 = (het10.A.multiply(d[:, None]).multiply(d[None, :])).tocsc()
    A2.sort_indices()
    dv.upload(_as_csc(A2)[3])
    s.factorize_values_device(dv.ptr, rowidx.size)
    assert s.get_info()["factorize_mode"] == "values"
    fresh = _solver(3, True, het10.coords, lazy_partitioning=True, precision="mixed", coarse_solver="cholesky")
    fresh.factorize(A2)
    assert _rel(s.apply_preconditioner(x), fresh.apply_preconditioner(x)) <= 1e-12
    # no coordinates: the domains come from the device aggregation (the host never sees A)
    g = _solver(3, True)
    g.factorize_device(dc.ptr, dr.ptr, dv.ptr, n, rowidx.size)
    assert g.get_info()["aggregation"] == "graph_gpu"
    for a in (dc, dr, dv):
        a.free()


def test_device_input_validation(gpu):
    A = helpers.prefactor_matrix(helpers.grid_laplacian_pattern(6, 5, 2), seed=1)
    n, colptr, rowidx, vals = _as_csc(A)
    s = _solver(1, True)
    bad_rows = rowidx.copy()
    bad_rows[7] = n
    cases = [(colptr, bad_rows, "row index out of range"),
             (np.concatenate([[1], colptr[1:]]).astype(np.int32), rowidx, "do not span"),
             (np.where(np.arange(n + 1) == 3, colptr[5], colptr).astype(np.int32), rowidx, "not monotone")]
    for cp, ri, msg in cases:
        dc, dr, dv = DeviceArray(n + 1, cp, np.int32), DeviceArray(ri.size, ri, np.int32), DeviceArray(vals.size, vals)
        with pytest.raises(MASError, match=msg):
            s.factorize_device(dc.ptr, dr.ptr, dv.ptr, n, ri.size)
        for a in (dc, dr, dv):
            a.free()
