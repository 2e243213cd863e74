"""psmas_solve_device_noalloc (no device allocation; partials and scalars in the library's static
per-GPU scratch) against psmas_solve_device, through the C ABI: bit-identical solutions and the
same iteration counts on two calls in a row, the second on a smaller system, for the FP64 and
the mixed path; a factorisation whose first-solve parts are missing is refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from polysolve_amd import MASError, MASSolver, _native, problems
from polysolve_amd.linear import DeviceArray

pytestmark = pytest.mark.gpu


def _noalloc(s, b, x, n):
    rc = _native.lib().psmas_solve_device_noalloc(s._h, int(n), C.c_void_p(b.ptr), C.c_void_p(x.ptr))
    if rc != _native.PSMAS_OK:
        raise MASError(_native.lib().psmas_last_error(s._h).decode())


def _solver(p, precision):
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 3, "precision": precision, "coarse_solver": "cholesky",
                              "relative_tolerance": 1e-8, "absolute_tolerance": 0.0, "max_iter": 5000}})
    s.set_coordinates(p.coords)
    s.factorize(p.A)
    return s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_solve_matches_solve_device(gpu, precision):
    big = problems.hex_elasticity(12, E=problems.random_moduli(12, 31), name="big")
    small = problems.hex_elasticity(6, E=problems.random_moduli(6, 32), name="small")
    cases = []
    for p in (big, small):
        s = _solver(p, precision)
        b, x_ref, x = DeviceArray(p.n, p.b), DeviceArray(p.n, np.zeros(p.n)), DeviceArray(p.n, np.zeros(p.n))
        s.solve_device(b.ptr, x_ref.ptr, p.n)  # also builds the parts made at a factorisation's first solve
        cases.append((p, s, b, x_ref, x, s.get_info()["solver_iter"]))
    # Two calls in a row, the second on the smaller system (the scratch keeps the first's contents).
    for p, s, b, x_ref, x, it_ref in cases:
        _noalloc(s, b, x, p.n)
        info = s.get_info()
        assert info["solver_status"] == "Reach relative tolerance"
        assert info["solver_iter"] == it_ref
        xr, xn = x_ref.download(), x.download()
        assert np.array_equal(xn, xr)
        assert np.linalg.norm(p.b - p.A @ xn) <= 1e-8 * np.linalg.norm(p.b)
    # The ordinary entry point still gives the same answer after the scratch solves.
    p, s, b, x_ref, x, it_ref = cases[0]
    s.solve_device(b.ptr, x.ptr, p.n)
    assert np.array_equal(x.download(), x_ref.download())
    for _, _, b, x_ref, x, _ in cases:
        for a in (b, x_ref, x):
            a.free()


def test_noalloc_refuses_what_would_allocate(gpu):
    p = problems.hex_elasticity(5, E=problems.random_moduli(5, 33))
    s = _solver(p, "mixed")
    b, x = DeviceArray(p.n, p.b), DeviceArray(p.n, np.zeros(p.n))
    with pytest.raises(MASError, match="first solve"):
        _noalloc(s, b, x, p.n)
    s.solve_device(b.ptr, x.ptr, p.n)
    s.set_parameters({"MAS": {"learned_collect": True}})
    with pytest.raises(MASError, match="learned_collect"):
        _noalloc(s, b, x, p.n)
    s.set_parameters({"MAS": {"learned_collect": False}})
    _noalloc(s, b, x, p.n)
    assert s.get_info()["solver_status"] == "Reach relative tolerance"
    with pytest.raises(MASError, match="Size mismatch"):
        _noalloc(s, b, x, p.n - 3)
    b.free()
    x.free()
