"""The Q1-hex elasticity assembler: closed forms and invariants (CPU)."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from polysolve_amd import problems


def rigid_modes(coords):
    """Six rigid-body displacement fields (3 translations, 3 rotations) at the given nodes."""
    n = coords.shape[0]
    modes = []
    for d in range(3):
        u = np.zeros((n, 3))
        u[:, d] = 1.0
        modes.append(u.ravel())
    x, y, z = coords.T
    for (a, b) in [(0, 1), (1, 2), (0, 2)]:
        u = np.zeros((n, 3))
        u[:, a] = -coords[:, b]
        u[:, b] = coords[:, a]
        modes.append(u.ravel())
    return np.stack(modes, axis=1)


def test_element_stiffness_kernel_is_rigid_motions():
    K = problems.unit_element_stiffness(0.3)
    assert np.array_equal(K, K.T)
    corners = np.array([[a & 1, (a >> 1) & 1, (a >> 2) & 1] for a in range(8)], dtype=float)
    R = rigid_modes(corners)
    assert np.abs(K @ R).max() < 1e-12 * np.abs(K).max()
    ev = np.linalg.eigvalsh(K)
    assert np.sum(ev < 1e-10 * ev.max()) == 6 and ev.min() > -1e-12 * ev.max()


def test_patch_test_linear_field_is_equilibrated_in_the_interior():
    ne = 6
    p = problems.hex_elasticity(ne, E=1e5)
    rng = np.random.default_rng(0)
    G = rng.standard_normal((3, 3))
    u = (p.coords @ G.T + rng.standard_normal(3)).ravel()
    f = p.A @ u
    N1 = ne + 1
    ijk = np.rint(p.coords * ne).astype(int)
    interior = np.all((ijk > 0) & (ijk < ne), axis=1) & (ijk[:, 2] > 1)
    fi = f.reshape(-1, 3)[interior]
    assert np.abs(fi).max() < 1e-9 * np.abs(p.A).max()


@pytest.mark.parametrize("ne,nnzb", [(16, 110735), (4, None)])
def test_sizes_symmetry_and_dirichlet(ne, nnzb):
    p = problems.hex_elasticity(ne, E=problems.random_moduli(ne, 1))
    N1 = ne + 1
    assert p.n == 3 * N1 ** 3
    assert (p.A - p.A.T).count_nonzero() == 0
    S = sp.kron(sp.identity(N1 ** 3), np.ones((3, 1))).tocsc()  # sums the 3x3 blocks
    Bn = S.T @ abs(p.A) @ S
    if nnzb is not None:
        assert (Bn != 0).sum() == nnzb
    # Dirichlet rows are identity, their RHS is zero.
    nd = 3 * N1 * N1
    Ad = p.A[:nd, :].toarray()
    assert np.array_equal(Ad, np.eye(nd, p.n))
    assert np.all(p.b[:nd] == 0.0)


def test_gravity_load_total():
    ne = 8
    p = problems.hex_elasticity(ne, E=1e5)
    h = 1.0 / ne
    # rho g V lumped to nodes; the bottom layer (Dirichlet) carries h/2 of the unit-volume load.
    assert np.isclose(p.b.reshape(-1, 3)[:, 2].sum(), -9.81 * (1.0 - h / 2.0), rtol=1e-13)
    assert np.all(p.b.reshape(-1, 3)[:, :2] == 0.0)


def test_spd_small():
    p = problems.hex_elasticity(3, E=problems.random_moduli(3, 2))
    assert np.linalg.eigvalsh(p.A.toarray()).min() > 0


def test_contact_pairs_add_spd_blocks():
    ne = 4
    pairs = problems.contact_pairs(ne, 30, seed=1)
    base = problems.hex_elasticity(ne, E=1e5)
    p = problems.hex_elasticity(ne, E=1e5, pairs=pairs)
    D = (p.A - base.A).toarray()
    # Expected: sum of kappa g g^T on the (i, j) DOFs, with Dirichlet rows/cols removed.
    ref = np.zeros_like(D)
    ii, jj, g, kappa = pairs
    for a, b, gg, k in zip(ii, jj, g, kappa):
        idx = np.r_[3 * a:3 * a + 3, 3 * b:3 * b + 3]
        ref[np.ix_(idx, idx)] += k * np.outer(gg, gg)
    nd = 3 * (ne + 1) ** 2
    ref[:nd, :] = 0.0
    ref[:, :nd] = 0.0
    assert np.allclose(D, ref, rtol=1e-12, atol=1e-9 * np.abs(ref).max())
    off = np.abs(np.stack([(jj % 5) - (ii % 5), (jj // 5) % 5 - (ii // 5) % 5, jj // 25 - ii // 25]))
    assert off.max() <= 2 and np.all(off.sum(axis=0) > 0)
    assert np.linalg.eigvalsh(p.A.toarray()).min() > 0


def test_config_sizes_match_workload_table():
    sizes = problems.config4_sizes(0)
    assert sizes.size == 64 and sizes.min() >= 8 and sizes.max() <= 32
    seq = list(problems.config5_moduli(ne=4, steps=3, seed=0))
    assert len(seq) == 3 and np.all(seq[1] >= 1e3) and np.all(seq[1] <= 1e6)
    ratio = seq[1] / seq[0]
    assert 0.5 < np.median(ratio) < 1.5
