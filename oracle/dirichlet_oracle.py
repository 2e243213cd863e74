"""CPU oracle of the FEM Dirichlet elimination — TEST INFRASTRUCTURE ONLY.

Restates polysolve's FEM helper (src/polysolve/linear/FEMSolver.cpp) on scipy / plain Python:

* ``eliminate(A, nodes)`` — the matrix of ``prefactorize`` / ``dirichlet_solve``
  (FEMSolver.cpp:139-160, 283-306): with N(k) = 1 for a fixed dof k, every entry a_ij with N(i) = 1
  or N(j) = 1 is dropped and the triplet (k, k, N(k)) is added for every k, so a fixed dof gets a
  lone 1 on its diagonal and a free dof keeps a_kk (a free column without a diagonal gets an
  explicit stored zero, as setFromTriplets keeps one).  Built column by column straight from the
  CSC arrays, row order preserved.
* ``lift(A, nodes, f)`` — the right-hand side g = f - (I - N) A N f (FEMSolver.cpp:123 and 332):
  g_i = f_i for a fixed i; for a free i, g_i = f_i - s_i with s_i the sum of a_ij f_j over the fixed
  columns j of row i, added in ascending column order with a separate rounding of every product
  and sum (pure Python floats; meant for test-sized matrices).  The reference's sparse product
  A (N f) visits the same terms in the same column order; its terms at free columns are exact
  zeros.

Parity is pinned in ``tests/test_dirichlet_oracle.py`` by what the mathematics fixes: the solution
of eliminate(A) u = lift(f) equals the reduced solve u_F = A_FF^-1 (f_F - A_FG f_G), u_G = f_G
built by index slicing (a dropped term, a wrong sign or a transposed coupling fails it), the
eliminated matrix equals the generator's own Dirichlet-fixed assembly entry for entry, and a
planted sign error in the lift is caught.  The reference holds no test vectors for this file.

Only ``tests/`` may import this module, and only as the checker.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def fixed_mask(n: int, nodes) -> np.ndarray:
    N = np.zeros(n, dtype=bool)
    N[np.asarray(nodes, dtype=np.int64)] = True
    return N


def eliminate(A, nodes) -> sp.csc_matrix:
    """FEMSolver.cpp:139-160 (prefactorize: 283-306) on a CSC matrix; explicit zeros are kept."""
    A = sp.csc_matrix(A)
    n = A.shape[0]
    N = fixed_mask(n, nodes)
    colptr, rowidx, vals = A.indptr, A.indices, A.data
    out_ptr = [0]
    out_rows: list[np.ndarray] = []
    out_vals: list[np.ndarray] = []
    for c in range(n):
        if N[c]:
            rows, v = np.array([c]), np.array([1.0])
        else:
            r = rowidx[colptr[c]:colptr[c + 1]]
            x = vals[colptr[c]:colptr[c + 1]]
            keep = ~N[r]
            rows, v = r[keep], x[keep]
            if not np.any(rows == c):
                below = np.flatnonzero(rows > c)  # before the first row below the diagonal
                at = int(below[0]) if below.size else rows.size
                rows = np.insert(rows, at, c)
                v = np.insert(v, at, 0.0)
        out_rows.append(rows)
        out_vals.append(v)
        out_ptr.append(out_ptr[-1] + rows.size)
    out = sp.csc_matrix((np.concatenate(out_vals), np.concatenate(out_rows).astype(np.int32),
                         np.asarray(out_ptr, dtype=np.int32)), shape=(n, n))
    return out


def lift(A, nodes, f) -> np.ndarray:
    """g = f - (I - N) A N f (FEMSolver.cpp:123, 332), sequential per row, fixed columns ascending."""
    A = sp.csc_matrix(A)
    n = A.shape[0]
    N = fixed_mask(n, nodes)
    f = np.asarray(f, dtype=np.float64)
    terms: dict[int, list[tuple[int, float]]] = {}
    for j in np.flatnonzero(N):
        for k in range(A.indptr[j], A.indptr[j + 1]):
            i = int(A.indices[k])
            if not N[i]:
                terms.setdefault(i, []).append((int(j), float(A.data[k])))
    g = f.copy()
    for i, lst in terms.items():
        s = 0.0
        for j, a in sorted(lst, key=lambda t: t[0]):
            s = s + a * float(f[j])
        g[i] = float(f[i]) - s
    return g


def reduced_solution(A, nodes, f) -> np.ndarray:
    """u_G = f_G, u_F = A_FF^-1 (f_F - A_FG f_G): the constrained problem solved by slicing."""
    import scipy.sparse.linalg as sla

    A = sp.csr_matrix(A)
    n = A.shape[0]
    N = fixed_mask(n, nodes)
    F, G = np.flatnonzero(~N), np.flatnonzero(N)
    u = np.zeros(n)
    u[G] = f[G]
    rhs = f[F] - A[F][:, G] @ f[G]
    u[F] = sla.spsolve(sp.csc_matrix(A[F][:, F]), rhs)
    return u
