"""The committed golden fixture (tests/golden/mas_hex5.npz, written by make_fixtures.py) against
the oracle (CPU) and against itself (the system it holds is consistent)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

import oracle

FIX = Path(__file__).with_name("golden") / "mas_hex5.npz"


@pytest.fixture(scope="module")
def golden():
    d = np.load(FIX)
    n = d["b"].size
    A = sp.csc_matrix((d["vals"], d["rowidx"], d["colptr"]), shape=(n, n))
    return d, A


def test_fixture_is_consistent(golden):
    d, A = golden
    assert abs(A - A.T).max() <= 1e-12 * abs(A).max()
    x = d["pcg_x"]
    assert np.linalg.norm(d["b"] - A @ x) <= 1e-10 * np.linalg.norm(d["b"])
    perm = d["perm"]
    assert np.array_equal(np.sort(perm), np.arange(perm.size))


def test_oracle_reproduces_fixture(golden):
    d, A = golden
    O = oracle.MasOracle(A, 3, d["offsets"], d["perm"])
    assert O.info["level_num"] == int(d["level_num"])
    assert [O.info[f"cco{i}"] for i in range(4)] == list(d["cco"])
    cm, _ = O.coarse_map()
    assert np.array_equal(cm, d["coarse_map"])
    _, inv = O.matrices()
    assert np.array_equal(inv, d["bank_inverses"])
    for r, z in zip(d["apply_in"], d["apply_out"]):
        assert np.array_equal(O.apply(r), z)
    res = O.pcg(d["b"], max_iter=2000, rel_tol=1e-10, abs_tol=0.0)
    assert res["iters"] == int(d["pcg_iters"]) and res["status"] == int(d["pcg_status"])
    assert np.array_equal(res["x"], d["pcg_x"])
