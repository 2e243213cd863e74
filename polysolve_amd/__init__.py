"""polysolve_amd — MI355X-native MAS-preconditioned PCG solver behind polysolve's Solver API.

The solver itself is the HIP/C++ library ``libpsmas.so`` (C ABI in ``include/psmas.h``);
:mod:`polysolve_amd.linear` mirrors ``polysolve::linear::Solver`` on top of it and
:mod:`polysolve_amd.problems` manufactures the Q1-hex elasticity benchmark systems.
"""
from .linear import (  # noqa: F401
    MASError,
    MASSolver,
    OutOfMemoryError,
    Solver,
    device_count,
    dirichlet_solve,
    dirichlet_solve_prefactorized,
    parse_parameters,
    prefactorize,
)

__all__ = ["Solver", "MASSolver", "MASError", "OutOfMemoryError", "device_count", "parse_parameters",
           "prefactorize", "dirichlet_solve", "dirichlet_solve_prefactorized"]
