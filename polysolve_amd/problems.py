"""Benchmark and test systems: Q1-hex linear elasticity on the unit cube.

The five workloads of the MAS solver's benchmark:

1. ``config1``: 16^3 elements, E = 1e5, nu = 0.3, bottom face fixed, gravity RHS (14,739 DOF).
2. ``config2``: 64^3 elements, per-element E = 10^U(3,6) from a seed (823,875 DOF).
3. ``config3``: 48^3 elements (E = 1e5) plus 20k random node pairs within 2 grid spacings
   (Chebyshev offsets in {-2..2}^3), each adding kappa g g^T, g ~ N(0,1)^6,
   kappa = 10^U(2,6); RHS ~ N(0,1).
4. ``config4``: 64 independent systems with edge e ~ U{8..32}, moduli as in 2.
5. ``config5_moduli``: 200-step sequence at 64^3 with E_{t+1} = clip(E_t (1 + 0.05 xi), 1e3, 1e6).

All randomness is numpy ``default_rng(seed)``; the assembly itself is deterministic C++
(``csrc/problems.cpp``).  Dirichlet DOFs follow polysolve's FEM helper convention
(src/polysolve/linear/FEMSolver.cpp:97-160): identity rows/columns, zero RHS.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


@dataclass
class Problem:
    A: object  # scipy.sparse.csc_matrix, int32 indices
    b: np.ndarray
    coords: np.ndarray  # (nodes, 3)
    ne: int
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.A.shape[0])


def grid_coords(ne: int) -> np.ndarray:
    N1 = ne + 1
    k, j, i = np.meshgrid(np.arange(N1), np.arange(N1), np.arange(N1), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) / ne


def unit_element_stiffness(nu: float = 0.3) -> np.ndarray:
    out = np.zeros(24 * 24, dtype=np.float64)
    N.problems_lib().psgen_unit_element_stiffness(nu, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out.reshape(24, 24)


def hex_elasticity(ne: int, E=1e5, nu: float = 0.3, body_force_z: float = -9.81,
                   pairs: tuple | None = None, rhs: np.ndarray | None = None, name: str = "",
                 fix_bottom: bool = True) -> Problem:
    """Assemble the system.  ``E`` is a scalar or ne^3 per-element moduli (element order
    i + ne*(j + ne*k)); ``pairs`` = (i, j, g (p,6), kappa (p,)) adds kappa g g^T blocks.
    ``fix_bottom=False`` returns the bare stiffness matrix (singular: rigid motions) and load,
    whose bottom-face DOFs are bottom_face_dofs(ne)."""
    import scipy.sparse as sp

    L = N.problems_lib()
    E_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(E, dtype=np.float64), (ne ** 3,)))
    if pairs is not None:
        pi, pj, pg, pk = pairs
        pi = np.ascontiguousarray(pi, dtype=np.int32)
        pj = np.ascontiguousarray(pj, dtype=np.int32)
        pg = np.ascontiguousarray(pg, dtype=np.float64).reshape(-1, 6)
        pk = np.ascontiguousarray(pk, dtype=np.float64)
        npairs = int(pi.size)
        args = (npairs, pi.ctypes.data_as(C.POINTER(C.c_int)), pj.ctypes.data_as(C.POINTER(C.c_int)),
                pg.ctypes.data_as(C.POINTER(C.c_double)), pk.ctypes.data_as(C.POINTER(C.c_double)))
    else:
        args = (0, None, None, None, None)
    n = C.c_int(0)
    nnz = C.c_longlong(0)
    h = L.psgen_hex_assemble(int(ne), E_arr.ctypes.data_as(C.POINTER(C.c_double)), float(nu),
                             float(body_force_z), *args, int(bool(fix_bottom)), C.byref(n), C.byref(nnz))
    if not h:
        raise MemoryError("hex assembly failed")
    try:
        colptr = np.empty(n.value + 1, dtype=np.int32)
        rowidx = np.empty(nnz.value, dtype=np.int32)
        vals = np.empty(nnz.value, dtype=np.float64)
        b = np.empty(n.value, dtype=np.float64)
        L.psgen_fetch(h, colptr.ctypes.data_as(C.POINTER(C.c_int)), rowidx.ctypes.data_as(C.POINTER(C.c_int)),
                      vals.ctypes.data_as(C.POINTER(C.c_double)), b.ctypes.data_as(C.POINTER(C.c_double)))
    finally:
        L.psgen_free(h)
    A = sp.csc_matrix((vals, rowidx, colptr), shape=(n.value, n.value))
    A.has_sorted_indices = True
    if rhs is not None:
        b = np.ascontiguousarray(rhs, dtype=np.float64).copy()
        if fix_bottom:
            b[bottom_face_dofs(ne)] = 0.0  # Dirichlet DOFs (bottom face) carry their zero value
    return Problem(A=A, b=b, coords=grid_coords(ne), ne=ne, name=name, meta={"nu": nu})


def bottom_face_dofs(ne: int) -> np.ndarray:
    """DOF ids of the bottom face z = 0 (nodes k = 0, all three components)."""
    return np.arange(3 * (ne + 1) ** 2, dtype=np.int32)


def random_moduli(ne: int, seed: int = 0, lo: float = 3.0, hi: float = 6.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return 10.0 ** rng.uniform(lo, hi, ne ** 3)


def config1() -> Problem:
    return hex_elasticity(16, E=1e5, name="hex16_E1e5")


def config2(seed: int = 0, ne: int = 64) -> Problem:
    p = hex_elasticity(ne, E=random_moduli(ne, seed), name=f"hex{ne}_logE_U3_6_seed{seed}")
    p.meta["seed"] = seed
    return p


def contact_pairs(ne: int, npairs: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    N1 = ne + 1
    nn = N1 ** 3
    ii, jj = [], []
    while len(ii) < npairs:
        m = 2 * (npairs - len(ii)) + 16
        a = rng.integers(0, nn, m)
        off = rng.integers(-2, 3, (m, 3))
        ai, aj, ak = a % N1, (a // N1) % N1, a // (N1 * N1)
        bi, bj, bk = ai + off[:, 0], aj + off[:, 1], ak + off[:, 2]
        ok = (np.abs(off).sum(axis=1) > 0) & (bi >= 0) & (bi < N1) & (bj >= 0) & (bj < N1) & (bk >= 0) & (bk < N1)
        b = bi + N1 * (bj + N1 * bk)
        ii.extend(a[ok].tolist())
        jj.extend(b[ok].tolist())
    ii = np.array(ii[:npairs], dtype=np.int32)
    jj = np.array(jj[:npairs], dtype=np.int32)
    g = rng.standard_normal((npairs, 6))
    kappa = 10.0 ** rng.uniform(2.0, 6.0, npairs)
    return ii, jj, g, kappa


def config3(seed: int = 0, ne: int = 48, npairs: int = 20000) -> Problem:
    pairs = contact_pairs(ne, npairs, seed)
    rng = np.random.default_rng(seed + 1)
    n = 3 * (ne + 1) ** 3
    rhs = rng.standard_normal(n)
    p = hex_elasticity(ne, E=1e5, pairs=pairs, rhs=rhs, name=f"hex{ne}_pairs{npairs}_seed{seed}")
    p.meta["seed"] = seed
    return p


def config4_sizes(seed: int = 0, count: int = 64, lo: int = 8, hi: int = 32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, count)


def config4(seed: int = 0, count: int = 64) -> list[Problem]:
    sizes = config4_sizes(seed, count)
    return [hex_elasticity(int(e), E=random_moduli(int(e), seed + 1000 + s), name=f"batch{s}_hex{e}")
            for s, e in enumerate(sizes)]


def config5_moduli(ne: int = 64, steps: int = 200, seed: int = 0):
    """Yield the per-element moduli of the drifting sequence, step 0 first."""
    E = random_moduli(ne, seed)
    rng = np.random.default_rng(seed + 5)
    for _ in range(steps):
        yield E
        E = np.clip(E * (1.0 + 0.05 * rng.standard_normal(E.size)), 1e3, 1e6)
