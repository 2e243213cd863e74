"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and closed forms."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from polysolve_amd import MASError, MASSolver, Solver, problems
from polysolve_amd.linear import DeviceArray
from tests import helpers

pytestmark = pytest.mark.gpu


def _factorized(p, params=None, coords=True):
    s = MASSolver()
    prm = {"MAS": {"block_dim": 3}}
    if params:
        prm["MAS"].update(params)
    s.set_parameters(prm)
    if coords:
        s.set_coordinates(p.coords)
    s.factorize(p.A)
    return s


@pytest.fixture(scope="module")
def cfg1():
    return problems.config1()


@pytest.fixture(scope="module")
def het12():
    return problems.hex_elasticity(12, E=problems.random_moduli(12, 3), name="het12")


def test_spmv_fp64_matches_csc(gpu, cfg1):
    s = _factorized(cfg1)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(cfg1.n)
    y = s.spmv(x)
    ref = cfg1.A @ x
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)


def test_spmv_fp32_tolerance(gpu, cfg1):
    s = _factorized(cfg1)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(cfg1.n)
    y = s.spmv(x, fp32=True)
    ref = cfg1.A @ x
    # FP32 values, vectors and accumulation: relative error a few ulps(fp32) times sqrt(row length).
    assert np.linalg.norm(y - ref) <= 2e-6 * np.linalg.norm(ref)


@pytest.mark.parametrize("bd", [1, 2, 3])
def test_partition_and_coarse_map_match_oracle(gpu, bd, het12):
    p = het12
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": bd, "aggregation": "graph"}})
    s.factorize(p.A)
    offsets, perm = s.partition()
    nb = (p.n + bd - 1) // bd
    assert perm.size == nb and np.array_equal(np.sort(perm), np.arange(nb))
    sizes = np.diff(offsets)
    assert sizes.max() <= 32 and sizes.min() >= 0 and offsets[-1] == nb and (offsets.size - 1) % 2 == 0
    O = oracle.MasOracle(p.A, bd, offsets, perm)
    assert O.status == 0
    li = s.level_info()
    assert li["level_num"] == O.info["level_num"]
    assert [li[f"cco{i}"] for i in range(4)] == [O.info[f"cco{i}"] for i in range(4)]
    om, _ = O.coarse_map()  # over permuted ids
    gm = s.coarse_map()     # over input block ids
    L = O.info["level_num"]
    assert np.array_equal(gm[:, :L], om[perm, :L])


def test_morton_partition_matches_numpy_order(gpu, cfg1):
    s = _factorized(cfg1, {"aggregation": "morton"})
    offsets, perm = s.partition()
    order = helpers.morton_order(cfg1.coords)
    ref_off, ref_perm = helpers.chunk_partition(order)
    assert np.array_equal(perm, ref_perm)
    assert np.array_equal(offsets, ref_off)


@pytest.mark.parametrize("agg", ["morton", "graph"])
def test_bank_inverses_match_oracle(gpu, cfg1, agg):
    s = _factorized(cfg1, {"aggregation": agg})
    offsets, perm = s.partition()
    O = oracle.MasOracle(cfg1.A, 3, offsets, perm)
    _, inv_ref = O.matrices()
    inv = s.bank_inverses(3)
    assert inv.shape == inv_ref.shape
    scale = np.abs(inv_ref).max(axis=1, keepdims=True)
    err = np.abs(inv - inv_ref) / scale
    assert err.max() < 1e-9


@pytest.mark.parametrize("cap", [0, 8])
@pytest.mark.parametrize("agg", ["morton", "graph"])
def test_apply_fp64_matches_oracle(gpu, cfg1, agg, cap):
    """cap = 8 workgroups apply all ~190 bank matrices, so each walks ~24 of them through the
    grid-stride loop with two matrices in flight (the path every workgroup takes on 64^3)."""
    s = _factorized(cfg1, {"aggregation": agg, "apply_workgroups": cap})
    offsets, perm = s.partition()
    O = oracle.MasOracle(cfg1.A, 3, offsets, perm)
    assert O.info["total_mats"] > 20 * 8
    rng = np.random.default_rng(2)
    for _ in range(3):
        r = rng.standard_normal(cfg1.n)
        z = s.apply_preconditioner(r)
        zr = O.apply(r)
        assert np.linalg.norm(z - zr) <= 1e-10 * np.linalg.norm(zr)
    if cap:
        # Which workgroup applies a matrix changes nothing in its arithmetic.
        z16 = s.apply_preconditioner(r, mixed=True)
        s.set_parameters({"MAS": {"apply_workgroups": 0}})
        assert np.array_equal(s.apply_preconditioner(r), z)
        assert np.array_equal(s.apply_preconditioner(r, mixed=True), z16)


def test_split_restriction_lists_match(gpu, het12):
    """Coarse levels 2.. are restricted straight from level 1; with restrict_piece = 9 every
    descendant list longer than 9 is cut into pieces whose sums the second apply launch adds.
    Both applies must equal the whole-list ones (FP64 to rounding) and the oracle."""
    base = _factorized(het12, {"coarse_solver": "cholesky"})
    cut = _factorized(het12, {"coarse_solver": "cholesky", "restrict_piece": 9})
    assert base.level_info()["level_num"] >= 2
    assert base.get_info()["coarse_restriction"]["split_nodes"] == 0
    info = cut.get_info()["coarse_restriction"]
    assert info["split_nodes"] > 0 and info["wave_pieces"] > info["split_nodes"], info
    offsets, perm = cut.partition()
    O = oracle.MasOracle(het12.A, 3, offsets, perm)
    rng = np.random.default_rng(5)
    for _ in range(2):
        r = rng.standard_normal(het12.n)
        z0, z1 = base.apply_preconditioner(r), cut.apply_preconditioner(r)
        assert np.linalg.norm(z1 - z0) <= 1e-13 * np.linalg.norm(z0)
        m0, m1 = base.apply_preconditioner(r, mixed=True), cut.apply_preconditioner(r, mixed=True)
        assert np.linalg.norm(m1 - m0) <= 1e-6 * np.linalg.norm(m0)
    # Without the dense coarsest solve the FP64 apply is the reference operator.
    cut.set_parameters({"MAS": {"coarse_solver": "bank"}})
    r = rng.standard_normal(het12.n)
    zr = O.apply(r)
    assert np.linalg.norm(cut.apply_preconditioner(r) - zr) <= 1e-10 * np.linalg.norm(zr)


@pytest.mark.parametrize("cap", [0, 8])
def test_apply_mixed_close_to_fp64(gpu, het12, cap):
    s = _factorized(het12, {"apply_workgroups": cap})
    rng = np.random.default_rng(3)
    r = rng.standard_normal(het12.n)
    z64 = s.apply_preconditioner(r)
    z16 = s.apply_preconditioner(r, mixed=True)
    # FP16 fine-level inverses (11-bit significand, unit-diagonal scaling): ~1e-3 relative.
    assert np.linalg.norm(z16 - z64) <= 5e-3 * np.linalg.norm(z64)
    # Still a symmetric positive operator on this vector.
    assert r @ z16 > 0
    # Symmetric to FP32 rounding: every off-diagonal FP16 entry lands at (i, j) and (j, i).
    v = rng.standard_normal(het12.n)
    zv = s.apply_preconditioner(v, mixed=True)
    assert abs(v @ z16 - r @ zv) <= 1e-5 * np.linalg.norm(v) * np.linalg.norm(z16)
    # Linear in the input (no state kept between calls).
    zs = s.apply_preconditioner(r + 2.0 * v, mixed=True)
    assert np.linalg.norm(zs - (z16 + 2.0 * zv)) <= 1e-5 * np.linalg.norm(zs)


@pytest.mark.parametrize("bd", [1, 2, 3])
@pytest.mark.parametrize("cap", [0, 3])
def test_wave_apply_matches_persistent_apply(gpu, bd, cap):
    """The wave-per-matrix FP16 apply (registers, DPP transposed products, slot partials) and the
    persistent LDS-unpacked one compute the same operator from the same FP16 matrices: equal to
    FP32 rounding (different summation orders), at block sizes 1-3, with the grid-stride loop
    (cap = 3 workgroups walk every matrix), split level-2.. inputs (restrict_piece = 9) and the
    dense coarsest rows of the second launch."""
    if bd == 3:
        p = problems.hex_elasticity(10, E=problems.random_moduli(10, 11))
        A, coords = p.A, p.coords
    else:
        A = helpers.prefactor_matrix(helpers.grid_laplacian_pattern(17, 13, 5), seed=7)
        coords = None
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": bd, "coarse_solver": "cholesky", "restrict_piece": 9,
                              "apply_workgroups": cap}})
    if coords is not None:
        s.set_coordinates(coords)
    s.factorize(A)
    rng = np.random.default_rng(21 + bd)
    for _ in range(2):
        r = rng.standard_normal(A.shape[0])
        s.set_parameters({"MAS": {"apply_wave": True}})
        zw = s.apply_preconditioner(r, mixed=True)
        s.set_parameters({"MAS": {"apply_wave": False}})
        zp = s.apply_preconditioner(r, mixed=True)
        z64 = s.apply_preconditioner(r)
        assert np.all(np.isfinite(zw))
        assert np.linalg.norm(zw - zp) <= 2e-6 * np.linalg.norm(zp)
        assert np.linalg.norm(zw - z64) <= 5e-3 * np.linalg.norm(z64)


def test_apply_matches_dense_construction(gpu):
    p = problems.hex_elasticity(4, E=problems.random_moduli(4, 7))
    s = _factorized(p, {"aggregation": "graph"})
    offsets, perm = s.partition()
    li = s.level_info()
    cm = s.coarse_map()
    cm_perm = np.empty_like(cm)
    cm_perm[perm] = cm
    M = helpers.dense_mas(p.A.toarray(), 3, offsets, perm, cm_perm, li["level_num"])
    rng = np.random.default_rng(4)
    r = rng.standard_normal(p.n)
    z = s.apply_preconditioner(r)
    assert np.linalg.norm(z - M @ r) <= 1e-10 * np.linalg.norm(M @ r)


def test_pcg_fp64_matches_oracle_schedule(gpu, het12):
    p = het12
    s = _factorized(p, {"aggregation": "graph", "relative_tolerance": 1e-8, "absolute_tolerance": 0.0,
                        "max_iter": 3000})
    x = np.zeros(p.n)
    s.solve(p.b, x)
    info = s.get_info()
    offsets, perm = s.partition()
    O = oracle.MasOracle(p.A, 3, offsets, perm)
    ref = O.pcg(p.b, max_iter=3000, rel_tol=1e-8, abs_tol=0.0)
    assert info["solver_status"] == "Reach relative tolerance"
    assert ref["status"] == 1
    assert abs(info["solver_iter"] - ref["iters"]) <= 10
    assert info["solver_iter"] % 10 == 0
    assert np.linalg.norm(x - ref["x"]) <= 1e-6 * np.linalg.norm(ref["x"])
    assert np.linalg.norm(p.b - p.A @ x) <= 2e-8 * np.linalg.norm(p.b)


def test_fp64_solution_matches_direct(gpu, cfg1):
    s = _factorized(cfg1, {"relative_tolerance": 1e-10, "absolute_tolerance": 0.0, "max_iter": 5000})
    x = np.zeros(cfg1.n)
    s.solve(cfg1.b, x)
    xd = spla.spsolve(cfg1.A.tocsc(), cfg1.b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


@pytest.mark.parametrize("cfg", ["cfg1", "het12"])
def test_mixed_reaches_fp64_tolerance(gpu, cfg, request):
    p = request.getfixturevalue(cfg)
    s = _factorized(p, {"precision": "mixed", "relative_tolerance": 1e-8, "absolute_tolerance": 0.0,
                        "max_iter": 5000})
    x = np.zeros(p.n)
    s.solve(p.b, x)
    info = s.get_info()
    assert info["solver_status"] == "Reach relative tolerance", info
    rel = np.linalg.norm(p.b - p.A @ x) / np.linalg.norm(p.b)
    assert rel <= 1e-8
    xd = spla.spsolve(p.A.tocsc(), p.b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


def test_api_semantics(gpu, cfg1):
    s = Solver.create("MAS", "")
    assert s.name() == "MAS"
    x = np.zeros(cfg1.n)
    with pytest.raises(MASError, match="Size mismatch"):
        s.solve(cfg1.b, x)
    s.set_parameters({"MAS": {"block_dim": 3, "max_iter": 1}})
    s.factorize(cfg1.A)
    s.solve(cfg1.b, x)
    info = s.get_info()
    assert info["solver_status"] == "Reach max iterations" and info["solver_iter"] == 1
    with pytest.raises(MASError, match="Size mismatch"):
        s.solve(cfg1.b[:-3], x[:-3])
    bad = cfg1.b.copy()
    bad[5] = np.nan
    with pytest.raises(MASError, match="Invalid initial residual"):
        s.solve(bad, x)


def test_parameters_ignored_without_mas_key(gpu):
    s = MASSolver()
    s.set_parameters({"Eigen::ConjugateGradient": {"max_iter": 3}})
    assert s.get_parameters()["max_iter"] == 5000


@pytest.mark.parametrize("bd", [1, 2, 3])
def test_prefactor_sequence(gpu, bd):
    """polysolve's "pre_factor" test shape: one solver, ten matrices on one pattern, each solved
    to ||Ax - b|| < 1e-8 (tests/test_linear_solver.cpp:241-307)."""
    pattern = helpers.grid_laplacian_pattern(17, 13, 5)
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": bd, "relative_tolerance": 0.0, "absolute_tolerance": 1e-10,
                              "max_iter": 5000}})
    rng = np.random.default_rng(42)
    for i in range(10):
        A = helpers.prefactor_matrix(pattern, seed=42 + i)
        b = rng.uniform(-1, 1, A.shape[0])
        x = np.zeros_like(b)
        s.factorize(A)
        s.solve(b, x)
        assert np.linalg.norm(A @ x - b) < 1e-8


@pytest.mark.parametrize("bd", [1, 2])
def test_mixed_small_block_sizes(gpu, bd):
    """The tiled FP16 apply and the side work of both apply launches at block sizes 1 and 2 (one
    and two 16-byte chunks per thread, 4 and 8 tiles per side): close to the FP64 apply,
    symmetric, and the mixed solve reaches the FP64 tolerance."""
    pattern = helpers.grid_laplacian_pattern(17, 13, 5)  # 1,105 unknowns: odd for bd = 2
    A = helpers.prefactor_matrix(pattern, seed=7)
    rng = np.random.default_rng(8)
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": bd, "precision": "mixed", "coarse_solver": "cholesky",
                              "relative_tolerance": 1e-8, "absolute_tolerance": 0.0, "max_iter": 5000}})
    s.factorize(A)
    r, v = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    z64 = s.apply_preconditioner(r)
    z16, zv = s.apply_preconditioner(r, mixed=True), s.apply_preconditioner(v, mixed=True)
    assert np.linalg.norm(z16 - z64) <= 5e-3 * np.linalg.norm(z64)
    assert abs(v @ z16 - r @ zv) <= 1e-5 * np.linalg.norm(v) * np.linalg.norm(z16)
    b = rng.uniform(-1, 1, A.shape[0])
    x = np.zeros_like(b)
    s.solve(b, x)
    assert s.get_info()["solver_status"] == "Reach relative tolerance"
    assert np.linalg.norm(b - A @ x) <= 1e-8 * np.linalg.norm(b)


def test_odd_sizes_trailing_block_padding(gpu):
    pattern = helpers.grid_laplacian_pattern(11, 7, 3)  # 231 unknowns
    A = helpers.prefactor_matrix(pattern, seed=5)
    b = np.random.default_rng(5).uniform(-1, 1, A.shape[0])
    for bd in (2, 3):  # 231 = 3*77, not a multiple of 2; exercise both
        s = MASSolver()
        s.set_parameters({"MAS": {"block_dim": bd, "relative_tolerance": 0.0, "absolute_tolerance": 1e-10}})
        s.factorize(A)
        x = np.zeros_like(b)
        s.solve(b, x)
        assert np.linalg.norm(A @ x - b) < 1e-8
    A2 = helpers.prefactor_matrix(helpers.grid_laplacian_pattern(10, 10, 1)[:-1, :-1], seed=6)  # 99 unknowns
    b2 = np.ones(A2.shape[0])
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 2, "relative_tolerance": 0.0, "absolute_tolerance": 1e-10}})
    s.factorize(A2)
    x2 = np.zeros_like(b2)
    s.solve(b2, x2)
    assert np.linalg.norm(A2 @ x2 - b2) < 1e-8


def test_lazy_partitioning_reuses_partition(gpu, het12):
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 3, "lazy_partitioning": True, "aggregation": "graph"}})
    s.factorize(het12.A)
    off1, perm1 = s.partition()
    A2 = het12.A * 2.0
    s.factorize(A2)
    off2, perm2 = s.partition()
    assert np.array_equal(off1, off2) and np.array_equal(perm1, perm2)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_values_only_refactorisation_matches_full(gpu, het12, precision):
    """With lazy_partitioning, a matrix with the last one's sparsity pattern only has its values
    scattered into the device arrays and the numeric MAS parts rebuilt: its solve must equal a
    fresh factorisation's (same Morton partition).  A changed pattern takes the full path."""
    prm = {"block_dim": 3, "lazy_partitioning": True, "aggregation": "morton", "precision": precision,
           "coarse_solver": "cholesky", "relative_tolerance": 1e-8, "absolute_tolerance": 0.0}
    rng = np.random.default_rng(11)
    d = sp.diags(rng.uniform(0.5, 2.0, het12.n))
    A2 = (d @ het12.A @ d).tocsc()  # same pattern, other values, still SPD
    s = MASSolver()
    s.set_parameters({"MAS": prm})
    s.set_coordinates(het12.coords)
    s.factorize(het12.A)
    assert s.get_info()["factorize_mode"] == "full"
    x1 = np.zeros(het12.n)
    s.solve(het12.b, x1)
    s.factorize(A2)
    assert s.get_info()["factorize_mode"] == "values"
    x2 = np.zeros(het12.n)
    s.solve(het12.b, x2)
    ref = MASSolver()
    ref.set_parameters({"MAS": prm})
    ref.set_coordinates(het12.coords)
    ref.factorize(A2)
    xr = np.zeros(het12.n)
    ref.solve(het12.b, xr)
    # The same operator (the Galerkin fill adds with FP64 atomics, so two factorisations agree to
    # rounding, and a solve's iteration count can move by one check interval).
    r = rng.standard_normal(het12.n)
    z, zr = s.apply_preconditioner(r), ref.apply_preconditioner(r)
    assert np.linalg.norm(z - zr) <= 1e-12 * np.linalg.norm(zr)
    assert abs(s.get_info()["solver_iter"] - ref.get_info()["solver_iter"]) <= 10
    assert np.linalg.norm(het12.b - A2 @ x2) <= 1e-8 * np.linalg.norm(het12.b)
    assert np.linalg.norm(x2 - xr) <= 1e-6 * np.linalg.norm(xr)
    # One more off-diagonal pair: a new pattern, so a full factorisation.
    A3 = A2.tolil()
    A3[0, het12.n - 1] = A3[het12.n - 1, 0] = 1e-3
    s.factorize(A3.tocsc())
    assert s.get_info()["factorize_mode"] == "full"


def test_values_from_device_memory(gpu, het12):
    """psmas_factorize_values_device: the values of a same-pattern matrix handed over in device
    memory give the same operator as the host values-only path; without a recorded pattern or
    with the wrong count it fails."""
    prm = {"block_dim": 3, "lazy_partitioning": True, "aggregation": "morton", "coarse_solver": "cholesky",
           "relative_tolerance": 1e-8, "absolute_tolerance": 0.0}
    rng = np.random.default_rng(12)
    d = sp.diags(rng.uniform(0.5, 2.0, het12.n))
    A2 = (d @ het12.A @ d).tocsc()
    A2.sort_indices()
    dev = MASSolver()
    dev.set_parameters({"MAS": prm})
    dev.set_coordinates(het12.coords)
    dev.factorize(het12.A)
    vals = DeviceArray(A2.nnz, A2.data)
    with pytest.raises(MASError, match="recorded pattern"):
        dev.factorize_values_device(vals.ptr, A2.nnz - 1)
    dev.factorize_values_device(vals.ptr, A2.nnz)
    assert dev.get_info()["factorize_mode"] == "values"
    host = MASSolver()
    host.set_parameters({"MAS": prm})
    host.set_coordinates(het12.coords)
    host.factorize(het12.A)
    host.factorize(A2)
    r = rng.standard_normal(het12.n)
    z, zh = dev.apply_preconditioner(r), host.apply_preconditioner(r)
    assert np.linalg.norm(z - zh) <= 1e-12 * np.linalg.norm(zh)
    x = np.zeros(het12.n)
    dev.solve(het12.b, x)
    assert np.linalg.norm(het12.b - A2 @ x) <= 1e-8 * np.linalg.norm(het12.b)
    plain = _factorized(het12)  # no lazy_partitioning: no pattern recorded
    with pytest.raises(MASError, match="recorded pattern"):
        plain.factorize_values_device(vals.ptr, A2.nnz)
    vals.free()


def test_inverse_failure_raises(gpu):
    pattern = helpers.grid_laplacian_pattern(8, 8, 1)
    A = helpers.prefactor_matrix(pattern, seed=1).tolil()
    A[5, 5] = np.nan
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 1}})
    with pytest.raises(MASError, match="inverse failed"):
        s.factorize(A.tocsc())


def test_solve_device_pointers(gpu, cfg1):
    s = _factorized(cfg1, {"relative_tolerance": 1e-8, "absolute_tolerance": 0.0})
    b = DeviceArray(cfg1.n, cfg1.b)
    x = DeviceArray(cfg1.n, np.zeros(cfg1.n))
    s.solve_device(b.ptr, x.ptr, cfg1.n)
    xh = x.download()
    x_host = np.zeros(cfg1.n)
    s.solve(cfg1.b, x_host)
    assert np.array_equal(xh, x_host)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_individual_solves(gpu, precision):
    """Uneven systems solved in one batch: each meets its tolerance and agrees with solving it
    alone (same kernels; the batch's coarse levels span all systems, so agreement is to tolerance)."""
    sizes = [3, 6, 4, 8]
    probs = [problems.hex_elasticity(e, E=problems.random_moduli(e, 100 + i)) for i, e in enumerate(sizes)]
    prm = {"MAS": {"block_dim": 3, "precision": precision, "relative_tolerance": 1e-9, "absolute_tolerance": 0.0,
                   "max_iter": 5000}}
    s = MASSolver()
    s.set_parameters(prm)
    s.factorize_batch([p.A for p in probs], coords=[p.coords for p in probs])
    xs = [np.zeros(p.n) for p in probs]
    s.solve_batch([p.b for p in probs], xs)
    for i, p in enumerate(probs):
        info = s.batch_info(i)
        assert info["solver_status"] == "Reach relative tolerance", (i, info)
        assert np.linalg.norm(p.b - p.A @ xs[i]) <= 1.5e-9 * np.linalg.norm(p.b)
        solo = _factorized(p, prm["MAS"])
        x1 = np.zeros(p.n)
        solo.solve(p.b, x1)
        assert np.linalg.norm(xs[i] - x1) <= 1e-6 * np.linalg.norm(x1)


def test_reference_division_guard_reproduces_stall(gpu):
    """division_guard = 1e-20 is the reference rule (MASSolver.cu:43-56): on a small-magnitude
    system p^T A p falls below 1e-20 and the iteration freezes, in the oracle and on the GPU."""
    p = problems.hex_elasticity(5, E=problems.random_moduli(5, 11))
    s = _factorized(p, {"aggregation": "graph", "relative_tolerance": 1e-12, "absolute_tolerance": 0.0,
                        "max_iter": 300, "division_guard": 1e-20})
    x = np.zeros(p.n)
    s.solve(p.b, x)
    info = s.get_info()
    offsets, perm = s.partition()
    ref = oracle.MasOracle(p.A, 3, offsets, perm).pcg(p.b, max_iter=300, rel_tol=1e-12, abs_tol=0.0,
                                                       division_guard=1e-20)
    assert info["solver_status"] == "Reach max iterations" and ref["status"] == 3
    rel_gpu = np.linalg.norm(p.b - p.A @ x) / np.linalg.norm(p.b)
    rel_ref = np.linalg.norm(p.b - p.A @ ref["x"]) / np.linalg.norm(p.b)
    assert 1e-9 < rel_gpu < 1e-6 and 1e-9 < rel_ref < 1e-6
    assert abs(np.log10(rel_gpu) - np.log10(rel_ref)) < 0.5
    # The default rule converges on the same system.
    s.set_parameters({"MAS": {"division_guard": 0.0}})
    s.solve(p.b, x)
    assert s.get_info()["solver_status"] == "Reach relative tolerance"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spmv_halo_matches_global_gathers(gpu, het12, precision):
    """x staged per workgroup in LDS (16-bit local columns) gives the same products, bit for bit,
    as gathering x through the 32-bit columns: a whole solve (its A p and b - A x kernels) ends
    with the same iterate."""
    # (The single-reduction iteration and the symmetric form need the halo lists: both solves use
    # the classic iteration over the full matrix.)
    s = _factorized(het12, {"precision": precision, "relative_tolerance": 1e-8, "absolute_tolerance": 0.0,
                            "single_reduction": False, "spmv_symmetric": False})
    info = s.get_info()["spmv_halo"]
    assert info["active"] and 0 < info["max_nodes"] <= 2048 and info["total_nodes"] >= het12.n // 3
    xs_halo = np.zeros(het12.n)
    s.solve(het12.b, xs_halo)
    it_halo = s.get_info()["solver_iter"]
    s.set_parameters({"MAS": {"spmv_halo": False}})
    assert not s.get_info()["spmv_halo"]["active"]
    xs = np.zeros(het12.n)
    s.solve(het12.b, xs)
    assert s.get_info()["solver_iter"] == it_halo
    assert np.array_equal(xs, xs_halo)
