"""The HIP path against the committed golden fixture (tests/golden/mas_hex5.npz)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from polysolve_amd import MASSolver

pytestmark = pytest.mark.gpu
FIX = Path(__file__).with_name("golden") / "mas_hex5.npz"


@pytest.fixture(scope="module")
def golden():
    d = np.load(FIX)
    n = d["b"].size
    A = sp.csc_matrix((d["vals"], d["rowidx"], d["colptr"]), shape=(n, n))
    s = MASSolver()
    s.set_parameters({"MAS": {"block_dim": 3, "aggregation": "morton", "relative_tolerance": 1e-10,
                              "absolute_tolerance": 0.0, "max_iter": 2000}})
    s.set_coordinates(d["coords"])
    s.factorize(A)
    return d, A, s


def test_partition_levels_and_coarse_map(gpu, golden):
    d, A, s = golden
    offsets, perm = s.partition()
    assert np.array_equal(offsets, d["offsets"]) and np.array_equal(perm, d["perm"])
    li = s.level_info()
    assert li["level_num"] == int(d["level_num"])
    assert [li[f"cco{i}"] for i in range(4)] == list(d["cco"])
    L = int(d["level_num"])
    assert np.array_equal(s.coarse_map()[:, :L], d["coarse_map"][d["perm"], :L])


def test_bank_inverses_and_apply(gpu, golden):
    d, A, s = golden
    inv = s.bank_inverses(3)
    ref = d["bank_inverses"]
    assert inv.shape == ref.shape
    assert np.max(np.abs(inv - ref) / np.abs(ref).max(axis=1, keepdims=True)) < 1e-9
    for r, z in zip(d["apply_in"], d["apply_out"]):
        y = s.apply_preconditioner(r)
        assert np.linalg.norm(y - z) <= 1e-10 * np.linalg.norm(z)


def test_pcg_solution(gpu, golden):
    d, A, s = golden
    x = np.zeros(d["b"].size)
    s.solve(d["b"], x)
    info = s.get_info()
    assert info["solver_status"] == "Reach relative tolerance"
    assert abs(info["solver_iter"] - int(d["pcg_iters"])) <= 10
    assert np.linalg.norm(x - d["pcg_x"]) <= 1e-6 * np.linalg.norm(d["pcg_x"])
