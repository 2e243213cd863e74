"""Shared helpers of the test suite (CPU-side constructions used as independent checks)."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def morton_order(coords: np.ndarray) -> np.ndarray:
    """Node order by 3-D Morton code of coordinates quantised to 21 bits over the bounding box
    (same quantisation contract as the solver's Morton aggregation)."""
    lo = coords.min(axis=0)
    ext = coords.max(axis=0) - lo
    inv = np.where(ext > 0, 1.0 / np.where(ext > 0, ext, 1.0), 0.0)
    u = np.clip((coords - lo) * inv, 0.0, 1.0)
    q = (u * 2097151.0).astype(np.uint64)
    key = np.zeros(coords.shape[0], dtype=np.uint64)
    for d in range(coords.shape[1]):
        v = q[:, d]
        for bit in range(21):
            key |= ((v >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + d)
    return np.argsort(key, kind="stable")


def chunk_partition(order: np.ndarray, even: bool = True):
    """Consecutive runs of 32 nodes; with ``even`` the part count is padded with empty parts to a
    multiple of 8 (the solver's layout: whole 256-node blocks)."""
    nb = order.size
    P = max(1, (nb + 31) // 32)
    offsets = np.minimum(np.arange(P + 1) * 32, nb)
    if even and P % 8:
        offsets = np.append(offsets, np.full(8 - P % 8, nb))
    perm = np.empty(nb, dtype=np.int32)
    perm[order] = np.arange(nb, dtype=np.int32)
    return offsets.astype(np.int32), perm


def prefactor_matrix(pattern: sp.spmatrix, seed: int = 42) -> sp.csc_matrix:
    """Random SPD matrix on a symmetric pattern, as polysolve's "pre_factor" test builds them
    (tests/test_linear_solver.cpp:262-284): diagonal U(0.1,5)*100, symmetric off-diagonal
    -U(0.1,5).  numpy's generator replaces std::default_random_engine."""
    rng = np.random.default_rng(seed)
    coo = sp.triu(pattern, k=1).tocoo()
    n = pattern.shape[0]
    off = -rng.uniform(0.1, 5.0, coo.nnz)
    diag = rng.uniform(0.1, 5.0, n) * 100.0
    U = sp.coo_matrix((off, (coo.row, coo.col)), shape=(n, n))
    A = U + U.T + sp.diags(diag)
    return A.tocsc()


def grid_laplacian_pattern(nx: int, ny: int, nz: int = 1) -> sp.csc_matrix:
    """7-point (or 5-point) grid pattern."""
    def lap1(m):
        return sp.diags([np.ones(m - 1), 2 * np.ones(m), np.ones(m - 1)], [-1, 0, 1])

    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    L = sp.kron(Iz, sp.kron(Iy, lap1(nx))) + sp.kron(Iz, sp.kron(lap1(ny), Ix))
    if nz > 1:
        L = L + sp.kron(lap1(nz), sp.kron(Iy, Ix))
    return (L != 0).astype(np.float64).tocsc()


def dense_mas(A: np.ndarray, bd: int, offsets, perm, cmap_perm, level_num, coarsest_dense: bool = False):
    """Dense additive MAS operator sum_l R_l^T blockdiag(inv(R_l A R_l^T)|banks) R_l built from
    scratch with numpy, given the partition and the coarse map (over permuted block ids).
    coarsest_dense: on the last level, the coarse nodes that contain a node with an off-diagonal
    block get the inverse of their full Galerkin matrix instead of the bank blocks."""
    n = A.shape[0]
    nb = (n + bd - 1) // bd
    Ap = np.zeros((nb * bd, nb * bd))
    Ap[:n, :n] = A
    for t in range(n, nb * bd):
        Ap[t, t] = 1.0
    # permuted block -> input block
    iperm = np.empty(nb, dtype=np.int64)
    iperm[np.asarray(perm)] = np.arange(nb)
    P = np.zeros((nb * bd, nb * bd))
    for new in range(nb):
        for c in range(bd):
            P[new * bd + c, iperm[new] * bd + c] = 1.0
    Aq = P @ Ap @ P.T  # permuted
    offsets = np.asarray(offsets)
    M = np.zeros_like(Aq)
    for p in range(offsets.size - 1):
        ids = np.arange(offsets[p], offsets[p + 1])
        if ids.size == 0:
            continue
        idx = (ids[:, None] * bd + np.arange(bd)[None, :]).ravel()
        M[np.ix_(idx, idx)] += np.linalg.inv(Aq[np.ix_(idx, idx)])
    for l in range(level_num):
        cm = cmap_perm[:, l]
        nc = int(cm.max()) + 1
        R = np.zeros((nc * bd, nb * bd))
        for blk in range(nb):
            for c in range(bd):
                R[cm[blk] * bd + c, blk * bd + c] = 1.0
        Ac = R @ Aq @ R.T
        Ci = np.zeros_like(Ac)
        for bank in range((nc + 31) // 32):
            ids = np.arange(bank * 32, min(nc, bank * 32 + 32))
            idx = (ids[:, None] * bd + np.arange(bd)[None, :]).ravel()
            Ci[np.ix_(idx, idx)] = np.linalg.inv(Ac[np.ix_(idx, idx)])
        if coarsest_dense and l == level_num - 1:
            blocks = np.abs(Aq).reshape(nb, bd, nb, bd).sum(axis=(1, 3))
            coupled = (blocks != 0).sum(axis=1) > 1  # has an off-diagonal block
            active = np.unique(cm[coupled])
            idx = (active[:, None] * bd + np.arange(bd)[None, :]).ravel()
            Ci[idx, :] = 0.0
            Ci[:, idx] = 0.0
            Ci[np.ix_(idx, idx)] = np.linalg.inv(Ac[np.ix_(idx, idx)])
        M += R.T @ Ci @ R
    Mo = P.T @ M @ P
    return Mo[:n, :n]


def unpack_upper(v: np.ndarray, N: int) -> np.ndarray:
    M = np.zeros((N, N))
    iu = np.triu_indices(N)
    M[iu] = v
    M[(iu[1], iu[0])] = v
    return M


def lattice_prolongation(A, coords, k, bd=3):
    """Trilinear lattice prolongation P (n x bd*k^3) of the learned coarse correction, built the
    way the solver builds it: bounding box of the coordinates, t = (x - lo) / (hi - lo) (k - 1),
    cell c = clamp(floor(t), 0, k - 2), fraction clamp(t - c, 0, 1) in float32, weights
    wx*wy*wz in float32; block rows without off-diagonal blocks (Dirichlet rows) are left out."""
    import scipy.sparse as sp

    X = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    nb = X.shape[0]
    n = A.shape[0]
    lo, hi = X.min(0), X.max(0)
    ext = hi - lo
    t = np.where(ext > 0, (X - lo) / np.where(ext > 0, ext, 1.0) * (k - 1), 0.0)
    c = np.clip(np.floor(t).astype(np.int64), 0, k - 2)
    f = np.clip(t - c, 0.0, 1.0).astype(np.float32)
    Ab = sp.csr_matrix(A)
    rows = np.repeat(np.arange(n), np.diff(Ab.indptr)) // bd
    cols = Ab.indices // bd
    coupled = np.zeros(nb, dtype=bool)
    coupled[rows[rows != cols]] = True
    R, Cc, V = [], [], []
    one = np.float32(1.0)
    for q in range(8):
        dx, dy, dz = q & 1, (q >> 1) & 1, (q >> 2) & 1
        w = ((f[:, 0] if dx else one - f[:, 0]) * (f[:, 1] if dy else one - f[:, 1])).astype(np.float32)
        w = (w * (f[:, 2] if dz else one - f[:, 2])).astype(np.float32)
        j = (c[:, 0] + dx) + k * ((c[:, 1] + dy) + k * (c[:, 2] + dz))
        for d in range(bd):
            idx = np.nonzero(coupled & (np.arange(nb) * bd + d < n))[0]
            R.append(idx * bd + d)
            Cc.append(j[idx] * bd + d)
            V.append(w[idx].astype(np.float64))
    return sp.csr_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(Cc))), shape=(n, bd * k ** 3))
