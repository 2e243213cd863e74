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

    def solve_device_noalloc(self, b_ptr: int, x_ptr: int, n: int) -> None:
        """solve_device without any device allocation (psmas_solve_device_noalloc): the solve's
        partials and scalars live in the library's static per-GPU scratch; run one solve_device
        after factorize first so that the parts built at the first solve exist."""
        _check(N.lib().psmas_solve_device_noalloc(self._h, int(n), C.c_void_p(b_ptr), C.c_void_p(x_ptr)), self._h)

    def dirichlet_factorize(self, A, dirichlet_nodes) -> None:
        """prefactorize (FEMSolver.cpp:269-315): factorise A with the rows and columns of the fixed
        dofs replaced by identity; A itself is not modified."""
        n, colptr, rowidx, vals = _as_csc(A)
        nodes = np.ascontiguousarray(dirichlet_nodes, dtype=np.int32).ravel()
        _check(N.lib().psmas_dirichlet_factorize_csc(self._h, n, int(rowidx.size), _ptr(colptr, C.c_int),
                                                     _ptr(rowidx, C.c_int), _ptr(vals, C.c_double),
                                                     int(nodes.size), _ptr(nodes, C.c_int)), self._h)
        self._n = n

    def dirichlet_factorize_device(self, colptr_ptr: int, rowidx_ptr: int, vals_ptr: int, n: int, nnz: int,
                                   dirichlet_nodes) -> None:
        """dirichlet_factorize with A's CSC arrays already in device memory (raw HIP pointers:
        int32 colptr of n+1, int32 rowidx and float64 vals of nnz entries); the fixed dofs are a
        host array."""
        nodes = np.ascontiguousarray(dirichlet_nodes, dtype=np.int32).ravel()
        _check(N.lib().psmas_dirichlet_factorize_csc_device(self._h, int(n), int(nnz), C.c_void_p(colptr_ptr),
                                                            C.c_void_p(rowidx_ptr), C.c_void_p(vals_ptr),
                                                            int(nodes.size), _ptr(nodes, C.c_int)), self._h)
        self._n = int(n)

    def dirichlet_solve(self, f, u) -> None:
        """dirichlet_solve_prefactorized (FEMSolver.cpp:318-342): f := g = f - (I - N) A N f, then
        u = solve(g); f and u are contiguous float64 arrays written in place."""
        for name, a in (("f", f), ("u", u)):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
                raise MASError(f"[MAS] {name} must be a contiguous float64 numpy array")
        if f.size != u.size:
            raise MASError("[MAS] Size mismatch between f and u.")
        _check(N.lib().psmas_dirichlet_solve(self._h, int(f.size), _ptr(f, C.c_double), _ptr(u, C.c_double)),
               self._h)

    def dirichlet_solve_device(self, f_ptr: int, u_ptr: int, n: int) -> None:
        """dirichlet_solve with f and u in device memory (raw HIP pointers); f becomes g."""
        _check(N.lib().psmas_dirichlet_solve_device(self._h, int(n), C.c_void_p(f_ptr), C.c_void_p(u_ptr)),
               self._h)

    def factorize_device(self, colptr_ptr: int, rowidx_ptr: int, vals_ptr: int, n: int, nnz: int) -> None:
        """Full factorisation from CSC arrays already in device memory (raw HIP pointers: int32
        colptr of n+1, int32 rowidx and float64 vals of nnz entries)."""
        _check(N.lib().psmas_factorize_csc_device(self._h, int(n), int(nnz), C.c_void_p(colptr_ptr),
                                                  C.c_void_p(rowidx_ptr), C.c_void_p(vals_ptr)), self._h)
        self._n = int(n)

    def factorize_multi(self, A, nrhs: int) -> None:
        """Factorise A for nrhs (1..8) right-hand sides solved together (solve_multi)."""
        n, colptr, rowidx, vals = _as_csc(A)
        _check(N.lib().psmas_factorize_multi_csc(self._h, int(nrhs), n, int(rowidx.size), _ptr(colptr, C.c_int),
                                                 _ptr(rowidx, C.c_int), _ptr(vals, C.c_double)), self._h)
        self._n = n


    def factorize_multi_device(self, colptr_ptr: int, rowidx_ptr: int, vals_ptr: int, n: int, nnz: int,
                               nrhs: int) -> None:
        """factorize_multi with A's CSC arrays already in device memory (raw HIP pointers)."""
        _check(N.lib().psmas_factorize_multi_csc_device(self._h, int(nrhs), int(n), int(nnz), C.c_void_p(colptr_ptr),
                                                        C.c_void_p(rowidx_ptr), C.c_void_p(vals_ptr)), self._h)
        self._n = int(n)
    def solve_multi(self, B) -> np.ndarray:
        """Solve A X = B for the nrhs columns of B (n x nrhs) after factorize_multi."""
        B = np.asfortranarray(B, dtype=np.float64)
        if B.ndim != 2:
            raise MASError("[MAS] B must be n x nrhs")
        n, k = B.shape
        X = np.zeros((n, k), dtype=np.float64, order="F")
        _check(N.lib().psmas_solve_multi(self._h, n, k, B.ctypes.data_as(C.POINTER(C.c_double)),
                                         X.ctypes.data_as(C.POINTER(C.c_double))), self._h)
        return X

    def solve_multi_device(self, b_ptr: int, x_ptr: int, n: int, nrhs: int) -> None:
        """solve_multi with B and X (n x nrhs, column-major) in device memory."""
        _check(N.lib().psmas_solve_multi_device(self._h, int(n), int(nrhs), C.c_void_p(b_ptr), C.c_void_p(x_ptr)),
               self._h)

    def factorize_values_device(self, vals_ptr: int, nnz: int) -> None:
        """Refactorise from nnz values already in device memory (raw HIP pointer), in the CSC order
        of the last full factorisation, which must have used lazy_partitioning."""
        _check(N.lib().psmas_factorize_values_device(self._h, int(nnz), C.c_void_p(vals_ptr)), self._h)

    def set_coordinates(self, xyz) -> None:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        if xyz.ndim != 2:
            raise MASError("[MAS] coordinates must be an (nb, dim) array")
        _check(N.lib().psmas_set_coordinates(self._h, int(xyz.shape[0]), int(xyz.shape[1]),
                                             _ptr(xyz, C.c_double)), self._h)

    # -- batches of independent systems -------------------------------------------------------
    def factorize_batch(self, As, coords=None) -> None:
        """Factorise several systems together (one launch sequence per PCG phase for all)."""
        L = N.lib()
        if coords is not None:
            for s, xyz in enumerate(coords):
                xyz = np.ascontiguousarray(xyz, dtype=np.float64)
                _check(L.psmas_batch_set_coordinates(self._h, s, int(xyz.shape[0]), int(xyz.shape[1]),
                                                     _ptr(xyz, C.c_double)), self._h)
        parts = [_as_csc(A) for A in As]
        k = len(parts)
        ns = np.array([p[0] for p in parts], dtype=np.int32)
        nnz = (C.c_longlong * k)(*[int(p[2].size) for p in parts])
        cp = (C.POINTER(C.c_int) * k)(*[_ptr(p[1], C.c_int) for p in parts])
        ri = (C.POINTER(C.c_int) * k)(*[_ptr(p[2], C.c_int) for p in parts])
        va = (C.POINTER(C.c_double) * k)(*[_ptr(p[3], C.c_double) for p in parts])
        self._batch_keep = parts
        _check(L.psmas_batch_factorize_csc(self._h, k, _ptr(ns, C.c_int), nnz, cp, ri, va), self._h)
        self._batch_n = [int(v) for v in ns]

    def solve_batch(self, bs, xs) -> None:
        """Solve every system of the batch; xs (contiguous float64 arrays) are written in place."""
        k = len(bs)
        bs = [np.ascontiguousarray(b, dtype=np.float64) for b in bs]
        for x in xs:
            if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
                raise MASError("[MAS] x must be a contiguous float64 numpy array")
        bp = (C.POINTER(C.c_double) * k)(*[_ptr(b, C.c_double) for b in bs])
        xp = (C.POINTER(C.c_double) * k)(*[_ptr(x, C.c_double) for x in xs])
        _check(N.lib().psmas_batch_solve(self._h, k, bp, xp, 0), self._h)

    def solve_batch_device(self, b_ptrs, x_ptrs) -> None:
        k = len(b_ptrs)
        bp = (C.POINTER(C.c_double) * k)(*[C.cast(C.c_void_p(p), C.POINTER(C.c_double)) for p in b_ptrs])
        xp = (C.POINTER(C.c_double) * k)(*[C.cast(C.c_void_p(p), C.POINTER(C.c_double)) for p in x_ptrs])
        _check(N.lib().psmas_batch_solve(self._h, k, bp, xp, 1), self._h)

    def batch_info(self, s: int) -> dict:
        rc, txt = N.read_string(N.lib().psmas_batch_get_info, self._h, int(s))
        _check(rc, self._h)
        return _json.loads(txt)

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

    # -- learned coarse correction (mixed path) ------------------------------------------------
    def learned_train(self, epochs: int = 200, batch: int = 64, lr: float = 3e-3, seed: int = 0) -> float:
        """Fit the coarse-correction network to the pairs collected by solves run with
        ``learned_collect``; returns the final mean training loss."""
        loss = C.c_double()
        _check(N.lib().psmas_learned_train(self._h, int(epochs), int(batch), float(lr), int(seed), C.byref(loss)),
               self._h)
        return float(loss.value)

    def autotune(self, reps: int = 20) -> dict:
        """Time the SpMV batch depth, check interval and coarsest solver on this factorisation and
        the last right-hand side(s); keeps the fastest; returns the timing report."""
        cap = 65536
        buf = C.create_string_buffer(cap)
        needed = C.c_size_t(0)
        _check(N.lib().psmas_autotune(self._h, int(reps), buf, cap, C.byref(needed)), self._h)
        return _json.loads(buf.value.decode())

    def lattice_restrict(self, v) -> np.ndarray:
        """u = P^T v on the learned correction's lattice (GPU kernels)."""
        m = C.c_int()
        _check(N.lib().psmas_lattice_size(self._h, C.byref(m)), self._h)
        v = np.ascontiguousarray(v, dtype=np.float64)
        u = np.empty(m.value)
        _check(N.lib().psmas_lattice_restrict(self._h, int(v.size), _ptr(v, C.c_double), _ptr(u, C.c_double)),
               self._h)
        return u

    def lattice_prolong(self, w) -> np.ndarray:
        """e = P w on the learned correction's lattice (GPU kernels)."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        e = np.empty(self._n)
        _check(N.lib().psmas_lattice_prolong(self._h, int(self._n), _ptr(w, C.c_double), _ptr(e, C.c_double)),
               self._h)
        return e

    def learned_reset(self, drop_pairs: bool = False) -> None:
        _check(N.lib().psmas_learned_reset(self._h, int(bool(drop_pairs))), self._h)

    def learned_exact(self) -> float:
        """Replace the network by the exact lattice coarse solve (P^T A P)^-1 P^T of the current
        factorisation (ablation); returns the ms it took."""
        ms = C.c_double(0.0)
        _check(N.lib().psmas_learned_exact(self._h, C.byref(ms)), self._h)
        return ms.value

    def set_kernel_timing(self, on: bool) -> None:
        _check(N.lib().psmas_set_kernel_timing(self._h, int(on)), self._h)

    def kernel_timing(self) -> dict[str, Any]:
        rc, s = N.read_string(N.lib().psmas_get_kernel_timing, self._h)
        _check(rc, self._h)
        return _json.loads(s)


def _mas_only(solver) -> "MASSolver":
    if not isinstance(solver, MASSolver):
        raise MASError("[MAS] the Dirichlet helpers here drive a MASSolver")
    return solver


def prefactorize(solver, A, dirichlet_nodes, precond_num: int = 0, save_path: str = "") -> None:
    """FEMSolver.hpp:42-44.  The reference overwrites A with the eliminated matrix; A is kept
    here (its couplings to the fixed dofs live on the device for the right-hand side lift)."""
    if save_path:
        raise MASError("[MAS] save_path is not supported: the eliminated matrix is built on the solver side")
    s = _mas_only(solver)
    s.analyze_pattern(A, precond_num)
    s.dirichlet_factorize(A, dirichlet_nodes)


def dirichlet_solve_prefactorized(solver, A, f, dirichlet_nodes, u) -> None:
    """FEMSolver.hpp:46-47: f := g = f - (I - N) A N f and u = solve(g).  The lift uses the A
    and the fixed dofs given to prefactorize; A and dirichlet_nodes are accepted for the
    reference's signature."""
    _mas_only(solver).dirichlet_solve(f, u)


def dirichlet_solve(solver, A, f, dirichlet_nodes, u, precond_num: int = 0, save_path: str = "",
                    compute_spectrum: bool = False, remove_zero_cols: bool = False,
                    skip_last_cols: bool = False) -> np.ndarray:
    """FEMSolver.hpp:35-40 (FEMSolver.cpp:97-266): prefactorize + dirichlet_solve_prefactorized.
    Returns the spectrum slot of the reference (zeros; compute_spectrum needs Spectra there and is
    refused here, as are remove_zero_cols and skip_last_cols)."""
    if compute_spectrum or remove_zero_cols or skip_last_cols:
        raise MASError("[MAS] compute_spectrum / remove_zero_cols / skip_last_cols are not supported")
    prefactorize(solver, A, dirichlet_nodes, precond_num, save_path)
    dirichlet_solve_prefactorized(solver, A, f, dirichlet_nodes, u)
    return np.zeros(4)


def parse_parameters(params) -> dict:
    """Parameter set that ``set_parameters(params)`` would produce on a fresh solver (no GPU)."""
    text = params if isinstance(params, str) else _json.dumps(params)
    rc, s = N.read_string(N.lib().psmas_parse_parameters, text.encode())
    if rc != N.PSMAS_OK:
        raise MASError(s)
    return _json.loads(s)


def block_graph(A, block_dim: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The block graph the host aggregation partitions (psmas_block_graph): CSR (ptr, adj) of the
    distinct block pairs joined by a stored entry, both directions, neighbours ascending, and the
    L1 weight of every edge (sum of |a| over both blocks)."""
    n, colptr, rowidx, vals = _as_csc(A)
    nb = -(-n // int(block_dim))
    ptr = np.zeros(nb + 1, dtype=np.int32)
    ne = C.c_longlong(0)
    lib = N.lib()
    args = (n, rowidx.size, _ptr(colptr, C.c_int), _ptr(rowidx, C.c_int), _ptr(vals, C.c_double), int(block_dim),
            _ptr(ptr, C.c_int), ptr.size)
    rc = lib.psmas_block_graph(*args, None, None, 0, C.byref(ne))
    if rc not in (N.PSMAS_OK, N.PSMAS_ERROR_BUFFER_TOO_SMALL):
        _check(rc, None)
    adj = np.zeros(max(ne.value, 1), dtype=np.int32)
    w = np.zeros(max(ne.value, 1), dtype=np.float64)
    _check(lib.psmas_block_graph(*args, _ptr(adj, C.c_int), _ptr(w, C.c_double), adj.size, C.byref(ne)), None)
    return ptr, adj[:ne.value], w[:ne.value]


def device_count() -> int:
    return int(N.lib().psmas_device_count())


def synchronize() -> None:
    _check(N.lib().psmas_device_synchronize(), None)


class DeviceArray:
    """A vector in device memory (float64 unless ``dtype`` says otherwise) owned through the C ABI
    (psmas_device_alloc).

    For callers without their own GPU allocator: ``solve_device`` and ``factorize_device`` take
    ``.ptr``.
    """

    def __init__(self, n: int, host: np.ndarray | None = None, dtype=np.float64) -> None:
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        _check(N.lib().psmas_device_alloc(max(1, self.n) * self.dtype.itemsize, C.byref(p)), None)
        self.ptr = int(p.value)
        if host is not None:
            self.upload(host)

    def upload(self, host) -> None:
        h = np.ascontiguousarray(host, dtype=self.dtype)
        if h.size != self.n:
            raise MASError("[MAS] invalid host array size")
        _check(N.lib().psmas_memcpy(C.c_void_p(self.ptr), h.ctypes.data_as(C.c_void_p), h.nbytes, 1), None)

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _check(N.lib().psmas_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes, 2), None)
        return out

    def free(self) -> None:
        if self.ptr:
            N.lib().psmas_device_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """GPU MLP d_in -> hidden -> hidden -> d_out (ReLU) of the learned coarse correction.

    BF16-operand MFMA GEMMs with FP32 accumulation, FP32 master weights, Adam on the GPU.
    """

    def __init__(self, d_in: int, d_out: int, hidden: int = 128, seed: int = 1234, skip: bool = False) -> None:
        h = C.c_void_p()
        rc = N.lib().psmas_mlp_create(int(d_in), int(d_out), int(hidden), int(bool(skip)), int(seed), C.byref(h))
        if rc != N.PSMAS_OK:
            raise MASError(N.lib().psmas_mlp_last_error(None).decode())
        self._h = h
        self.d_in, self.d_out, self.hidden, self.skip = int(d_in), int(d_out), int(hidden), bool(skip)

    def _check(self, rc: int) -> None:
        if rc != N.PSMAS_OK:
            msg = N.lib().psmas_mlp_last_error(self._h).decode()
            raise (OutOfMemoryError if rc == N.PSMAS_ERROR_OUT_OF_MEMORY else MASError)(msg)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().psmas_mlp_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def num_parameters(self) -> int:
        return int(N.lib().psmas_mlp_num_parameters(self._h))

    def set_parameters(self, theta) -> None:
        t = np.ascontiguousarray(theta, dtype=np.float32)
        if t.size != self.num_parameters:
            raise MASError("[MAS] invalid MLP parameter vector size")
        self._check(N.lib().psmas_mlp_set_parameters(self._h, _fptr(t)))

    def get_parameters(self) -> np.ndarray:
        t = np.empty(self.num_parameters, dtype=np.float32)
        self._check(N.lib().psmas_mlp_get_parameters(self._h, _fptr(t)))
        return t

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.d_in)
        Y = np.empty((X.shape[0], self.d_out), dtype=np.float32)
        self._check(N.lib().psmas_mlp_forward(self._h, X.shape[0], _fptr(X), _fptr(Y)))
        return Y

    def loss_and_grad(self, X, T) -> tuple[float, np.ndarray]:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.d_in)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, self.d_out)
        if X.shape[0] != T.shape[0]:
            raise MASError("[MAS] invalid MLP batch")
        loss = C.c_double()
        g = np.empty(self.num_parameters, dtype=np.float32)
        self._check(N.lib().psmas_mlp_loss_grad(self._h, X.shape[0], _fptr(X), _fptr(T), C.byref(loss), _fptr(g)))
        return float(loss.value), g

    def train(self, X, T, epochs: int = 10, batch: int = 256, lr: float = 1e-3, seed: int = 0) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, self.d_in)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, self.d_out)
        loss = C.c_double()
        self._check(N.lib().psmas_mlp_train(self._h, X.shape[0], _fptr(X), _fptr(T), int(epochs), int(batch),
                                            float(lr), int(seed), C.byref(loss)))
        return float(loss.value)


def gemm_bf16(A, B, trans_a: bool = False) -> np.ndarray:
    """C = op(A) B on the GPU with BF16-rounded operands and FP32 accumulation (the MLP's GEMM)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    K, Nn = B.shape
    M = A.shape[1] if trans_a else A.shape[0]
    if (A.shape[0] if trans_a else A.shape[1]) != K:
        raise MASError("[MAS] invalid GEMM shapes")
    Cm = np.empty((M, Nn), dtype=np.float32)
    rc = N.lib().psmas_gemm_bf16(M, Nn, K, int(trans_a), _fptr(A), _fptr(B), _fptr(Cm))
    if rc != N.PSMAS_OK:
        raise MASError(f"psmas_gemm_bf16 failed ({rc})")
    return Cm
