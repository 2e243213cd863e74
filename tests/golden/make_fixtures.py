"""Write the golden fixtures of the MAS path: a small seeded heterogeneous elasticity system and
the oracle's outputs on it.  Run from the repository root:  python tests/golden/make_fixtures.py

The reference (CUDA C++ with external libraries) does not build in this image and its tests hold
no vectors for this path, so the expected outputs come from the CPU oracle (oracle/), whose
functions are pinned against closed forms and brute force in tests/test_oracle.py."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from polysolve_amd import problems  # noqa: E402
from tests import helpers  # noqa: E402


def main() -> None:
    p = problems.hex_elasticity(5, E=problems.random_moduli(5, 7), name="golden_hex5")
    A = p.A.tocsc()
    A.sort_indices()
    order = helpers.morton_order(p.coords)
    offsets, perm = helpers.chunk_partition(order)
    O = oracle.MasOracle(p.A, 3, offsets, perm)
    cmap, _ = O.coarse_map()
    _, inv = O.matrices()
    rng = np.random.default_rng(42)
    R = rng.standard_normal((3, p.n))
    Z = np.stack([O.apply(r) for r in R])
    res = O.pcg(p.b, max_iter=2000, rel_tol=1e-10, abs_tol=0.0)
    info = O.info
    np.savez_compressed(
        Path(__file__).with_name("mas_hex5.npz"),
        colptr=A.indptr.astype(np.int32), rowidx=A.indices.astype(np.int32), vals=A.data, b=p.b,
        coords=p.coords, offsets=offsets.astype(np.int32), perm=perm.astype(np.int32),
        level_num=np.int32(info["level_num"]), cco=np.array([info[f"cco{i}"] for i in range(4)], dtype=np.int32),
        coarse_map=cmap.astype(np.int32), bank_inverses=inv, apply_in=R, apply_out=Z,
        pcg_x=res["x"], pcg_iters=np.int32(res["iters"]), pcg_status=np.int32(res["status"]))


if __name__ == "__main__":
    main()
