"""CPU oracle for the MAS-preconditioned PCG path — TEST INFRASTRUCTURE ONLY.

Python bindings of ``liborc.so`` (``mas_oracle.c``: a plain-C restatement of the reference MAS
algorithm, file:line citations in its header).  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this package, and only as the checker / the CPU
baseline — never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"
_lib = None

_i, _d, _vp = C.c_int, C.c_double, C.c_void_p
_pi, _pd = C.POINTER(C.c_int), C.POINTER(C.c_double)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from . import build as _b

            _b.build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_mas_create.restype = _vp
        L.orc_mas_create.argtypes = [_i, _pi, _pi, _pd, _i, _i, _pi, _pi]
        L.orc_mas_free.restype = None
        L.orc_mas_free.argtypes = [_vp]
        L.orc_mas_status.restype = _i
        L.orc_mas_status.argtypes = [_vp]
        L.orc_mas_info.restype = None
        L.orc_mas_info.argtypes = [_vp, _pi]
        L.orc_mas_get_bsr.restype = None
        L.orc_mas_get_bsr.argtypes = [_vp, _pi, _pi, _pd]
        L.orc_mas_get_map.restype = None
        L.orc_mas_get_map.argtypes = [_vp, _pi, _pi]
        L.orc_mas_get_matrices.restype = None
        L.orc_mas_get_matrices.argtypes = [_vp, _pd, _pd]
        L.orc_mas_apply.restype = None
        L.orc_mas_apply.argtypes = [_vp, _pd, _pd]
        L.orc_mas_spmv.restype = None
        L.orc_mas_spmv.argtypes = [_vp, _pd, _pd]
        L.orc_mas_pcg.restype = _i
        L.orc_mas_pcg.argtypes = [_vp, _pd, _pd, _i, _d, _d, _i, _i, _d, _pi, _pd, _pi]
        L.orc_jacobi_cg.restype = _i
        L.orc_jacobi_cg.argtypes = [_i, _pi, _pi, _pd, _pd, _pd, _d, _i, _pd]
        L.orc_jacobi_cg_omp.restype = _i
        L.orc_jacobi_cg_omp.argtypes = [_i, _pi, _pi, _pd, _pd, _pd, _d, _i, _i, _pd]
        L.orc_omp_max_threads.restype = _i
        L.orc_omp_max_threads.argtypes = []
        L.orc_extend_add.restype = None
        L.orc_extend_add.argtypes = [_i, _pd, _pi, _i, _i, _pd, _pd, _pd]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _csc(A):
    csc = A.tocsc()
    if not csc.has_sorted_indices:
        csc = csc.sorted_indices()
    return (int(A.shape[0]), np.ascontiguousarray(csc.indptr, dtype=np.int32),
            np.ascontiguousarray(csc.indices, dtype=np.int32), np.ascontiguousarray(csc.data, dtype=np.float64))


class MasOracle:
    """FP64 MAS preconditioner + reference-schedule PCG for a given domain partition."""

    INFO_KEYS = ["nb", "nnzb", "npad", "level_num", "cco0", "cco1", "cco2", "cco3", "total_mats", "mat_dim",
                 "mat_storage", "mc0", "mc1", "mc2", "mc3", "mc4", "total_level_nodes"]

    def __init__(self, A, block_dim: int, offsets, perm):
        self.n, colptr, rowidx, vals = _csc(A)
        self._keep = (colptr, rowidx, vals)
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        perm = np.ascontiguousarray(perm, dtype=np.int32)
        self.bd = block_dim
        self._h = lib().orc_mas_create(self.n, _p(colptr, C.c_int), _p(rowidx, C.c_int), _p(vals, C.c_double),
                                       block_dim, int(offsets.size - 1), _p(offsets, C.c_int), _p(perm, C.c_int))
        if not self._h:
            raise MemoryError("oracle allocation failed")
        self.status = lib().orc_mas_status(self._h)
        info = np.zeros(17, dtype=np.int32)
        lib().orc_mas_info(self._h, _p(info, C.c_int))
        self.info = {k: int(v) for k, v in zip(self.INFO_KEYS, info)}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_mas_free(h)
            self._h = None

    def bsr(self):
        nb, nnzb, bd = self.info["nb"], self.info["nnzb"], self.bd
        rows = np.zeros(nb + 1, dtype=np.int32)
        cols = np.zeros(nnzb, dtype=np.int32)
        vals = np.zeros(nnzb * bd * bd, dtype=np.float64)
        lib().orc_mas_get_bsr(self._h, _p(rows, C.c_int), _p(cols, C.c_int), _p(vals, C.c_double))
        return rows, cols, vals.reshape(nnzb, bd, bd)

    def coarse_map(self):
        """(map over permuted block ids (nb, 4), real_to_padded (nb,))."""
        nb = self.info["nb"]
        m = np.zeros(nb * 4, dtype=np.int32)
        r2p = np.zeros(nb, dtype=np.int32)
        lib().orc_mas_get_map(self._h, _p(m, C.c_int), _p(r2p, C.c_int))
        return m.reshape(nb, 4), r2p

    def matrices(self):
        tot, L = self.info["total_mats"], self.info["mat_storage"]
        fwd = np.zeros(tot * L, dtype=np.float64)
        inv = np.zeros(tot * L, dtype=np.float64)
        lib().orc_mas_get_matrices(self._h, _p(fwd, C.c_double), _p(inv, C.c_double))
        return fwd.reshape(tot, L), inv.reshape(tot, L)

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        lib().orc_mas_apply(self._h, _p(r, C.c_double), _p(z, C.c_double))
        return z

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().orc_mas_spmv(self._h, _p(x, C.c_double), _p(y, C.c_double))
        return y

    def pcg(self, b, max_iter=5000, rel_tol=1e-5, abs_tol=1e-8, use_precond_norm=False, true_residual_period=4,
            division_guard=0.0):
        """division_guard=1e-20 is the reference's scalar_division rule; 0 the product default."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        it = C.c_int(0)
        res = C.c_double(0.0)
        st = C.c_int(0)
        rc = lib().orc_mas_pcg(self._h, _p(b, C.c_double), _p(x, C.c_double), int(max_iter), float(rel_tol),
                               float(abs_tol), int(use_precond_norm), int(true_residual_period),
                               float(division_guard), C.byref(it),
                               C.byref(res), C.byref(st))
        return {"x": x, "iters": it.value, "residual_norm": res.value, "status": st.value, "rc": rc}


def jacobi_cg(A, b, tol=1e-8, max_iter=1000, x0=None):
    """Eigen::ConjugateGradient<Lower|Upper, DiagonalPreconditioner> semantics (single thread)."""
    n, colptr, rowidx, vals = _csc(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    err = C.c_double(0.0)
    it = lib().orc_jacobi_cg(n, _p(colptr, C.c_int), _p(rowidx, C.c_int), _p(vals, C.c_double), _p(b, C.c_double),
                             _p(x, C.c_double), float(tol), int(max_iter), C.byref(err))
    return x, it, err.value


def jacobi_cg_omp(A, b, tol=1e-8, max_iter=1000, threads=0, x0=None):
    """The same Eigen CG semantics with OpenMP over ``threads`` threads (0: all); A symmetric.
    Returns (x, iterations, ||r|| / ||b||)."""
    n, colptr, rowidx, vals = _csc(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    err = C.c_double(0.0)
    it = lib().orc_jacobi_cg_omp(n, _p(colptr, C.c_int), _p(rowidx, C.c_int), _p(vals, C.c_double),
                                 _p(b, C.c_double), _p(x, C.c_double), float(tol), int(max_iter), int(threads),
                                 C.byref(err))
    return x, it, err.value


def omp_max_threads() -> int:
    return int(lib().orc_omp_max_threads())


STATUS = {0: "Running", 1: "Reach relative tolerance", 2: "Reach absolute tolerance", 3: "Reach max iterations"}
