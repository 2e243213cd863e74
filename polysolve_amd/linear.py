"""Host-side mirror of polysolve's linear-solver interface for the MAS path.

Same names, argument meaning and error behaviour as ``polysolve::linear::Solver`` /
``MASSolver`` (src/polysolve/linear/Solver.hpp:54-131, MASSolver.hpp:35-72):

* ``Solver.create("MAS", precond)`` returns a :class:`MASSolver` (Solver.cpp:403-408);
* ``set_parameters(params)`` reads ``params["MAS"]`` (MASSolver.cu:605-614);
* ``get_info(info)`` writes ``solver_iter``, ``solver_error``, ``solver_status``;
* ``analyze_pattern`` is a no-op; ``factorize(A)`` copies A; ``solve(b, x)`` writes x in place;
* failures raise :class:`MASError` (a ``RuntimeError``, as the reference throws
  ``std::runtime_error``), allocation failures :class:`OutOfMemoryError`.

Every call goes through the C ABI of ``include/psmas.h`` into the HIP library.
"""
from __future__ import annotations

import ctypes as C
import json as _json
from typing import Any

import numpy as np

from . import _native as N


class MASError(RuntimeError):
    """std::runtime_error raised by the MAS solver."""


class OutOfMemoryError(MASError):
    """CudaOutOfMemoryError counterpart (device allocation failed)."""


def _check(rc: int, handle) -> None:
    if rc == N.PSMAS_OK:
        return
    msg = N.lib().psmas_last_error(handle)
    text = msg.decode() if msg else f"psmas error {rc}"
    if rc == N.PSMAS_ERROR_OUT_OF_MEMORY:
        raise OutOfMemoryError(text)
    raise MASError(text)


def _as_csc(A):
    """scipy sparse (any format) -> (n, colptr int32, rowidx int32, vals float64), sorted."""
    import scipy.sparse as sp

    if not sp.issparse(A):
        raise MASError("[MAS] invalid A: expected a scipy sparse matrix")
    if A.shape[0] != A.shape[1]:
        raise MASError("[MAS] invalid A: matrix must be square")
    csc = A.tocsc()
    if not csc.has_sorted_indices:
        csc = csc.sorted_indices()
    colptr = np.ascontiguousarray(csc.indptr, dtype=np.int32)
    rowidx = np.ascontiguousarray(csc.indices, dtype=np.int32)
    vals = np.ascontiguousarray(csc.data, dtype=np.float64)
    return int(A.shape[0]), colptr, rowidx, vals


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Solver:
    """Abstract linear solver (polysolve::linear::Solver)."""

    @staticmethod
    def create(solver: str, precond: str = "") -> "Solver":
        if solver == "MAS":
            return MASSolver()
        raise MASError(f"Solver {solver} is not available")

    @staticmethod
    def available_solvers() -> list[str]:
        return ["MAS"]

    @staticmethod
    def default_solver() -> str:
        return "MAS"

    def set_parameters(self, params) -> None:  # pragma: no cover - interface
        pass

    def get_info(self, info: dict | None = None) -> dict:  # pragma: no cover - interface
        return {} if info is None else info

    def analyze_pattern(self, A, precond_num: int) -> None:  # pragma: no cover - interface
        pass

    def factorize(self, A) -> None:  # pragma: no cover - interface
        pass

    def solve(self, b, x) -> None:  # pragma: no cover - interface
        raise NotImplementedError

    def name(self) -> str:
        return ""


class MASSolver(Solver):
    """MI355X MAS-preconditioned PCG behind the polysolve Solver interface."""

    def __init__(self) -> None:
        L = N.lib()
        h = C.c_void_p()
        rc = L.psmas_create(C.byref(h))
        if rc != N.PSMAS_OK:
            msg = L.psmas_last_error(None)
            text = msg.decode() if msg else f"psmas_create failed ({rc})"
            if rc == N.PSMAS_ERROR_OUT_OF_MEMORY:
                raise OutOfMemoryError(text)
            raise MASError(text)
        self._h = h
        self._n = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().psmas_destroy(h)
            self._h = None

    # -- Solver interface -------------------------------------------------------------------
    def name(self) -> str:
        return N.lib().psmas_name().decode()

    def set_parameters(self, params) -> None:
        text = params if isinstance(params, str) else _json.dumps(params)
        _check(N.lib().psmas_set_parameters(self._h, text.encode()), self._h)

    def get_parameters(self) -> dict:
        rc, s = N.read_string(N.lib().psmas_get_parameters, self._h)
        _check(rc, self._h)
        return _json.loads(s)

    def get_info(self, info: dict | None = None) -> dict:
        rc, s = N.read_string(N.lib().psmas_get_info, self._h)
        _check(rc, self._h)
        d = _json.loads(s)
        if info is not None:
            info.update(d)
            return info
        return d

    def analyze_pattern(self, A, precond_num: int = 0) -> None:
        _check(N.lib().psmas_analyze_pattern(self._h, int(A.shape[0]), None, None, int(precond_num)), self._h)

    def factorize(self, A) -> None:
        n, colptr, rowidx, vals = _as_csc(A)
        _check(N.lib().psmas_factorize_csc(self._h, n, int(rowidx.size), _ptr(colptr, C.c_int),
                                           _ptr(rowidx, C.c_int), _ptr(vals, C.c_double)), self._h)
        self._n = n

    def solve(self, b, x) -> None:
        """Solve A x = b; x (float64, contiguous) is written in place."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
            raise MASError("[MAS] x must be a contiguous float64 numpy array")
        if b.size != x.size:
            raise MASError("[MAS] Size mismatch. Did you forget to call factorize?")
        _check(N.lib().psmas_solve(self._h, int(b.size), _ptr(b, C.c_double), _ptr(x, C.c_double)), self._h)

    # -- extensions ---------------------------------------------------------------------------
    def solve_device(self, b_ptr: int, x_ptr: int, n: int) -> None:
        """Solve with b and x already in device memory (raw HIP pointers)."""
        _check(N.lib().psmas_solve_device(self._h, int(n), C.c_void_p(b_ptr), C.c_void_p(x_ptr)), self._h)

    def set_coordinates(self, xyz) -> None:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        if xyz.ndim != 2:
            raise MASError("[MAS] coordinates must be an (nb, dim) array")
        _check(N.lib().psmas_set_coordinates(self._h, int(xyz.shape[0]), int(xyz.shape[1]),
                                             _ptr(xyz, C.c_double)), self._h)

    def level_info(self) -> dict:
        info = np.zeros(16, dtype=np.int32)
        _check(N.lib().psmas_get_level_info(self._h, _ptr(info, C.c_int), 16), self._h)
        keys = ["level_num", "cco0", "cco1", "cco2", "cco3", "parts", "npad", "bank_matrices", "nnzb",
                "sell_slots", "nb", "n"]
        return {k: int(v) for k, v in zip(keys, info)}

    def partition(self) -> tuple[np.ndarray, np.ndarray]:
        li = self.level_info()
        nparts = C.c_int(0)
        offsets = np.zeros(li["parts"] + 1, dtype=np.int32)
        perm = np.zeros(li["nb"], dtype=np.int32)
        _check(N.lib().psmas_get_partition(self._h, C.byref(nparts), _ptr(offsets, C.c_int), offsets.size,
                                           _ptr(perm, C.c_int), perm.size), self._h)
        return offsets, perm

    def coarse_map(self) -> np.ndarray:
        li = self.level_info()
        m = np.zeros(li["nb"] * 4, dtype=np.int32)
        _check(N.lib().psmas_get_coarse_map(self._h, _ptr(m, C.c_int), m.size), self._h)
        return m.reshape(-1, 4)

    def bank_inverses(self, block_dim: int) -> np.ndarray:
        li = self.level_info()
        dim = 32 * block_dim
        L = dim * (dim + 1) // 2
        out = np.zeros(li["bank_matrices"] * L, dtype=np.float64)
        _check(N.lib().psmas_get_bank_inverses(self._h, _ptr(out, C.c_double), out.size), self._h)
        return out.reshape(li["bank_matrices"], L)

    def apply_preconditioner(self, r, mixed: bool = False) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        _check(N.lib().psmas_apply_preconditioner(self._h, int(r.size), _ptr(r, C.c_double),
                                                  _ptr(z, C.c_double), int(mixed)), self._h)
        return z

    def spmv(self, x, fp32: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _check(N.lib().psmas_spmv(self._h, int(x.size), _ptr(x, C.c_double), _ptr(y, C.c_double), int(fp32)),
               self._h)
        return y

    def set_kernel_timing(self, on: bool) -> None:
        _check(N.lib().psmas_set_kernel_timing(self._h, int(on)), self._h)

    def kernel_timing(self) -> dict[str, Any]:
        rc, s = N.read_string(N.lib().psmas_get_kernel_timing, self._h)
        _check(rc, self._h)
        return _json.loads(s)


def parse_parameters(params) -> dict:
    """Parameter set that ``set_parameters(params)`` would produce on a fresh solver (no GPU)."""
    text = params if isinstance(params, str) else _json.dumps(params)
    rc, s = N.read_string(N.lib().psmas_parse_parameters, text.encode())
    if rc != N.PSMAS_OK:
        raise MASError(s)
    return _json.loads(s)


def device_count() -> int:
    return int(N.lib().psmas_device_count())
