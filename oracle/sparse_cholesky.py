"""FP64 sparse Cholesky for the accuracy oracle and the CPU direct-solver baseline — TEST
INFRASTRUCTURE ONLY (used by ``tests/`` and the CPU-baseline leg of ``bench.py``, never by the
product).

The reference's direct CPU solver is CHOLMOD's supernodal LLᵀ (``Solver.cpp:335-339`` creates
``CholmodSolver``; SuiteSparse, fetched by CPM, absent from this image).  This module is the
stand-in the measurement plan names: a multifrontal supernodal Cholesky A = P L Lᵀ Pᵀ with a
nested-dissection ordering, dense fronts factorised by LAPACK/BLAS (``dpotrf``, ``dtrsm``,
``dsyrk`` through scipy, threaded by its OpenBLAS).

* **Ordering** (``analyze``): nested dissection on the node graph of the bd×bd block structure.
  With node coordinates each step splits the node set at the median along its widest axis, and
  the nodes of the upper half that touch the lower half form the separator; without coordinates
  the split is the median level of a breadth-first search.  Sets of at most ``leaf`` nodes are
  not split.  Every separator (or leaf) is one supernode, ordered after its two subtrees.
* **Symbolic**: a supernode's front holds its own dofs and its boundary, i.e. the later dofs
  adjacent to it in A or in the boundary of a child (nested dissection puts them all in
  ancestors).
* **Numeric** (``factorize``): fronts in postorder; each assembles its columns of A and the
  children's update matrices (extend-add, a C loop in ``jacobi_cg_omp.c``), then L11 = chol(F11), L21 = F21 L11⁻ᵀ and the update
  F22 − L21 L21ᵀ for its parent.  Only lower triangles are used.
* **Solve**: forward and backward substitution over the supernodes.

Pinned by ``tests/test_sparse_cholesky.py``: P L Lᵀ Pᵀ = A entry for entry on small systems, the
solution against ``numpy.linalg.solve`` / ``scipy.sparse.linalg.spsolve``, an indefinite matrix
refused.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg.blas as blas
import scipy.linalg.lapack as lapack
import scipy.sparse as sp
from scipy.linalg import solve_triangular


def _orc():
    from . import lib

    return lib()


def _dp(a):
    assert a.flags.f_contiguous and a.dtype == np.float64
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


@dataclass
class _Super:
    nodes: np.ndarray                  # block nodes eliminated here
    children: list = field(default_factory=list)
    bnd_nodes: np.ndarray | None = None
    piv: tuple = (0, 0)                # dof positions [a, b)
    bnd: np.ndarray | None = None      # dof positions of the boundary, ascending
    L11: np.ndarray | None = None
    L21: np.ndarray | None = None


class SparseCholesky:
    """A (symmetric positive definite, both triangles stored) = P L Lᵀ Pᵀ."""

    def __init__(self, A, block_dim: int = 1, coords=None, leaf: int = 64):
        A = sp.csc_matrix(A, dtype=np.float64)
        n = A.shape[0]
        if A.shape != (n, n):
            raise ValueError("square matrix expected")
        bd = block_dim if (block_dim >= 1 and n % block_dim == 0) else 1
        self.n, self.bd, self.leaf = n, bd, max(int(leaf), 1)
        self.A = A
        self.coords = None if coords is None else np.asarray(coords, dtype=np.float64).reshape(n // bd, -1)
        self.times = {}
        t = time.perf_counter()
        self._analyze()
        self.times["analyze_s"] = time.perf_counter() - t

    # ------------------------------------------------------------------ ordering + symbolic
    def _node_graph(self):
        bd, nb = self.bd, self.n // self.bd
        coo = self.A.tocoo()
        G = sp.csr_matrix((np.ones(coo.nnz), (coo.row // bd, coo.col // bd)), shape=(nb, nb))
        G.sum_duplicates()
        G.setdiag(0)
        G.eliminate_zeros()
        return G

    def _split(self, nodes):
        """(lower half, separator, upper half without the separator) or None."""
        G = self.G
        if self.coords is not None:
            c = self.coords[nodes]
            axis = int(np.argmax(np.ptp(c, axis=0)))
            v = c[:, axis]
            order = np.argsort(v, kind="stable")
            cut = v[order[len(order) // 2]]
            lower = v < cut
            if not lower.any():
                lower = v <= cut
                if lower.all():
                    return None
        else:
            sub = G[nodes][:, nodes]
            from scipy.sparse.csgraph import breadth_first_order
            start = int(np.argmin(np.diff(sub.indptr)))
            levels = np.full(len(nodes), -1)
            order = breadth_first_order(sub, start, directed=False, return_predecessors=False)
            # BFS levels from the order: level(v) = level(parent) + 1
            _, pred = breadth_first_order(sub, start, directed=False, return_predecessors=True)
            levels[order[0]] = 0
            for v in order[1:]:
                levels[v] = levels[pred[v]] + 1
            if (levels < 0).any():  # disconnected: split off the unreached part
                lower = levels >= 0
            else:
                cut = np.median(levels)
                lower = levels < cut
                if not lower.any() or lower.all():
                    return None
        mark = np.zeros(G.shape[0], dtype=np.float64)
        mark[nodes[lower]] = 1.0
        rest = nodes[~lower]
        touch = G[rest] @ mark > 0
        return nodes[lower], rest[touch], rest[~touch]

    def _build(self, nodes):
        if len(nodes) <= self.leaf:
            return _Super(nodes=np.sort(nodes))
        parts = self._split(nodes)
        if parts is None:
            return _Super(nodes=np.sort(nodes))
        lo, sep, hi = parts
        kids = [self._build(x) for x in (lo, hi) if len(x)]
        if len(sep) == 0 and len(kids) == 2:
            # no separator: the two halves are independent; keep them under an empty supernode
            return _Super(nodes=np.zeros(0, dtype=np.int64), children=kids)
        return _Super(nodes=np.sort(sep), children=kids)

    def _analyze(self):
        bd, nb = self.bd, self.n // self.bd
        self.G = self._node_graph()
        root = self._build(np.arange(nb))
        post = []
        stack = [(root, False)]
        while stack:
            s, done = stack.pop()
            if done:
                post.append(s)
                continue
            stack.append((s, True))
            for c in reversed(s.children):
                stack.append((c, False))
        self.post = post
        node_pos = np.empty(nb, dtype=np.int64)
        p = 0
        for s in post:
            node_pos[s.nodes] = np.arange(p, p + len(s.nodes))
            p += len(s.nodes)
        assert p == nb
        self.node_pos = node_pos
        # dof permutation: new position -> old dof
        order_nodes = np.empty(nb, dtype=np.int64)
        order_nodes[node_pos] = np.arange(nb)
        self.perm = (order_nodes[:, None] * bd + np.arange(bd)[None, :]).ravel()
        self.iperm = np.empty(self.n, dtype=np.int64)
        self.iperm[self.perm] = np.arange(self.n)
        G = self.G
        nnz_l = 0
        flops = 0.0
        for s in post:
            end = node_pos[s.nodes].max() if len(s.nodes) else -1
            cand = [G[s.nodes].indices] if len(s.nodes) else []
            cand += [c.bnd_nodes for c in s.children]
            cand = np.unique(np.concatenate(cand)) if cand else np.zeros(0, dtype=np.int64)
            if len(s.nodes):
                cand = cand[node_pos[cand] > end]
            s.bnd_nodes = cand[np.argsort(node_pos[cand])]
            a = int(node_pos[s.nodes].min()) * bd if len(s.nodes) else 0
            k = len(s.nodes) * bd
            s.piv = (a, a + k)
            s.bnd = (node_pos[s.bnd_nodes][:, None] * bd + np.arange(bd)[None, :]).ravel()
            mb = len(s.bnd)
            nnz_l += k * (k + 1) // 2 + mb * k
            flops += k ** 3 / 3.0 + mb * k * k + mb * mb * k
        self.nnz_L = nnz_l
        self.flops = flops
        self.supernodes = len(post)
        del self.G

    # ----------------------------------------------------------------------------- numeric
    def factorize(self, progress=None, every_s: float = 20.0):
        """Numeric factorisation; progress(done_flops_fraction, seconds) is called about every
        ``every_s`` seconds (long factorisations print from it)."""
        t = time.perf_counter()
        last = t
        done = 0.0
        Ap = sp.csc_matrix(self.A[self.perm][:, self.perm])
        Ap.sort_indices()
        ptr, ind, val = Ap.indptr, Ap.indices, Ap.data
        n = self.n
        loc = np.full(n, -1, dtype=np.int64)
        updates = {}
        for s in self.post:
            a, b = s.piv
            k, mb = b - a, len(s.bnd)
            F11 = np.zeros((k, k), order="F")
            F21 = np.zeros((mb, k), order="F")
            F22 = np.zeros((mb, mb), order="F")
            loc[a:b] = np.arange(k)
            loc[s.bnd] = k + np.arange(mb)
            # columns a..b of A, lower part
            lo, hi = ptr[a], ptr[b]
            rows = ind[lo:hi]
            cols = np.repeat(np.arange(a, b), np.diff(ptr[a:b + 1]))
            v = val[lo:hi]
            m = rows >= cols
            rows, cols, v = rows[m], cols[m], v[m]
            lr = loc[rows]
            if (lr < 0).any():
                raise RuntimeError("sparse Cholesky: structure error (row outside the front)")
            top = lr < k
            F11[lr[top], cols[top] - a] = v[top]
            F21[lr[~top] - k, cols[~top] - a] = v[~top]
            # extend-add the children's updates (lower triangles; a C loop of the oracle library)
            for c in s.children:
                U = updates.pop(id(c), None)
                if U is None:
                    continue
                li = np.ascontiguousarray(loc[c.bnd], dtype=np.int32)
                if (li < 0).any():
                    raise RuntimeError("sparse Cholesky: structure error (child boundary outside the front)")
                _orc().orc_extend_add(len(li), _dp(U), _ip(li), k, mb, _dp(F11), _dp(F21), _dp(F22))
            if k:
                L11, info = lapack.dpotrf(F11, lower=1, clean=1, overwrite_a=1)
                if info != 0:
                    raise np.linalg.LinAlgError(f"sparse Cholesky: matrix not positive definite (pivot {a + info})")
                if mb:
                    F21 = blas.dtrsm(1.0, L11, F21, side=1, lower=1, trans_a=1, diag=0, overwrite_b=1)
                    F22 = blas.dsyrk(-1.0, F21, beta=1.0, c=F22, trans=0, lower=1, overwrite_c=1)
                s.L11, s.L21 = L11, F21
            else:
                s.L11, s.L21 = np.zeros((0, 0)), np.zeros((mb, 0))
            if mb:
                updates[id(s)] = F22
            loc[a:b] = -1
            loc[s.bnd] = -1
            done += k ** 3 / 3.0 + mb * k * k + mb * mb * k
            if progress is not None and time.perf_counter() - last >= every_s:
                last = time.perf_counter()
                progress(done / max(self.flops, 1.0), last - t)
        self.times["factorize_s"] = time.perf_counter() - t
        return self

    def solve(self, b):
        t = time.perf_counter()
        b = np.asarray(b, dtype=np.float64)
        y = b[self.perm].copy()
        for s in self.post:
            a, c = s.piv
            if c > a:
                y[a:c] = solve_triangular(s.L11, y[a:c], lower=True, check_finite=False)
                if len(s.bnd):
                    y[s.bnd] -= s.L21 @ y[a:c]
        for s in reversed(self.post):
            a, c = s.piv
            if c > a:
                rhs = y[a:c] - (s.L21.T @ y[s.bnd] if len(s.bnd) else 0.0)
                y[a:c] = solve_triangular(s.L11, rhs, lower=True, trans="T", check_finite=False)
        x = y[self.iperm]
        self.times["solve_s"] = time.perf_counter() - t
        return x

    def dense_factor(self):
        """L (dense, permuted ordering) for small-system checks."""
        L = np.zeros((self.n, self.n))
        for s in self.post:
            a, c = s.piv
            L[a:c, a:c] = np.tril(s.L11)
            if len(s.bnd):
                L[np.ix_(s.bnd, np.arange(a, c))] = s.L21
        return L
