"""The search-direction update folded into the SpMV (spmv_fuse_direction) against the separate
update kernel: the same arithmetic on the same values, so the iterates are bit-identical — the
FP64 and mixed paths, graph-captured or launched directly, the flexible PCG of the learned
correction (with its fall-back restart to plain MAS), and a batch of unequal systems."""
from __future__ import annotations

import numpy as np
import pytest

from polysolve_amd import MASSolver, problems

pytestmark = pytest.mark.gpu

# The fused update is a form of the classic iteration (two reductions per iteration) over the full
# matrix: the separate update is compared on the same matrix form.
BASE = {"block_dim": 3, "coarse_solver": "cholesky", "relative_tolerance": 1e-8, "absolute_tolerance": 0.0,
        "max_iter": 5000, "single_reduction": False, "spmv_symmetric": False}


def _solver(p, **extra):
    s = MASSolver()
    s.set_parameters({"MAS": dict(BASE, **extra)})
    s.set_coordinates(p.coords)
    s.factorize(p.A)
    return s


def _solve(s, p, fuse):
    """One solve on the solver's factorisation (the same bank inverses for both forms)."""
    s.set_parameters({"MAS": {"spmv_fuse_direction": fuse}})
    x = np.zeros(p.n)
    s.solve(p.b, x)
    return s.get_info(), x


@pytest.fixture(scope="module")
def hex16():
    return problems.hex_elasticity(16)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("graphs", [True, False])
def test_fused_equals_separate_update(gpu, hex16, precision, graphs):
    p = hex16
    s = _solver(p, precision=precision, use_graphs=graphs)
    i0, x0 = _solve(s, p, False)
    i1, x1 = _solve(s, p, True)
    assert i0["solver_status"] == i1["solver_status"] == "Reach relative tolerance"
    assert i0["solver_iter"] == i1["solver_iter"]
    assert np.array_equal(x0, x1)
    assert np.linalg.norm(p.b - p.A @ x1) <= 1e-8 * np.linalg.norm(p.b)
    # The fused kernel is the one that ran.
    s.set_kernel_timing(True)
    s.solve(p.b, np.zeros(p.n))
    names = set(s.kernel_timing())
    s.set_kernel_timing(False)
    assert ("spmv_dir_fp64" if precision == "fp64" else "spmv_dir_fp32") in names, names
    assert not {"spmv_fp64", "spmv_fp32"} & names, names


def test_repeated_solves_alternate_direction_buffers(gpu, hex16):
    """Graph segments replayed from the cache start on either direction buffer (7 iterations per
    segment): three solves in a row give the same iterates as the separate update."""
    p = hex16
    s = _solver(p, precision="mixed", check_interval=7)
    xs = [_solve(s, p, True)[1] for _ in range(3)]
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[1], xs[2])
    assert np.array_equal(xs[0], _solve(s, p, False)[1])


@pytest.mark.parametrize("scale,stall", [(1.0, 3), (-1e3, 1)])
def test_fused_flexible_pcg_of_the_learned_correction(gpu, hex16, scale, stall):
    """The exact lattice coarse solve installed as the correction (flexible PCG, z_prev kept by
    the SpMV).  Second case: a negative scale makes the preconditioner indefinite and one check
    interval (2 iterations) without a new residual minimum switches the correction off, so plain
    PCG restarts from the current residual."""
    p = hex16
    s = _solver(p, precision="mixed", lattice_points=5, learned_output_scale=scale, learned_stall_segments=stall,
                check_interval=2)
    s.learned_exact()
    s.set_parameters({"MAS": {"learned_correction": True}})
    out = []
    for fuse in (False, True):
        info, x = _solve(s, p, fuse)
        assert info["learned"]["used"]
        assert info["solver_status"] == "Reach relative tolerance"
        out.append((info["solver_iter"], info["learned"]["fallback"], x))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert out[1][1] == (scale < 0.0)
    assert np.array_equal(out[0][2], out[1][2])


def test_fused_batch_of_unequal_systems(gpu):
    probs = [problems.hex_elasticity(ne, name=f"b{ne}") for ne in (6, 11, 9)]
    s = MASSolver()
    s.set_parameters({"MAS": dict(BASE, precision="mixed")})
    s.factorize_batch([q.A for q in probs], coords=[q.coords for q in probs])
    res = []
    for fuse in (False, True):
        s.set_parameters({"MAS": {"spmv_fuse_direction": fuse}})
        xs = [np.zeros(q.n) for q in probs]
        s.solve_batch([q.b for q in probs], xs)
        res.append(([s.batch_info(i)["solver_iter"] for i in range(len(probs))], xs))
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1], res[1][1]):
        assert np.array_equal(a, b)
