"""The block graph (psmas_block_graph, the graph the host aggregation partitions) against the
oracle's restatement of the reference's adjacency build (oracle/block_graph.py), and the oracle
pinned by what the definitions fix: a hand-built matrix, dense block sums of |A|, and the
quantisation's end points.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import block_graph as BG
from polysolve_amd import MASError, problems
from polysolve_amd.linear import block_graph


def _rows(ptr, adj, w):
    return [dict(zip(adj[ptr[u]:ptr[u + 1]].tolist(), w[ptr[u]:ptr[u + 1]].tolist())) for u in range(len(ptr) - 1)]


def _random(n, bd, density, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format="csc")
    A.data = rng.uniform(-2.0, 2.0, A.nnz)
    A.data[A.data == 0.0] = 0.5
    return (A + sp.diags(np.full(n, 4.0))).tocsc()


def test_oracle_hand_example():
    # n = 5, bd = 2: block nodes {0,1}, {2,3}, {4} (trailing partial block).
    dense = np.zeros((5, 5))
    dense[0, 1] = 7.0     # inside block 0: not an edge
    dense[2, 0] = -1.5    # (1, 0)
    dense[0, 3] = 2.0     # (0, 1)
    dense[1, 2] = -0.25   # (0, 1)
    dense[4, 3] = 3.0     # (2, 1)
    for i in range(5):
        dense[i, i] = 10.0
    ptr, adj, w = BG.reference_adjacency(sp.csc_matrix(dense), 2)
    assert _rows(ptr, adj, w) == [{1: 3.75}, {0: 3.75, 2: 3.0}, {1: 3.0}]
    assert BG.quantize(np.array([3.75, 3.0])).tolist() == [1000000, 1]


def test_oracle_equals_dense_block_sums():
    n, bd = 37, 3
    A = _random(n, bd, 0.08, 1)
    nb = -(-n // bd)
    D = np.abs(A.toarray())
    P = (A.toarray() != 0)
    ptr, adj, w = BG.reference_adjacency(A, bd)
    rows = _rows(ptr, adj, w)
    for u in range(nb):
        exp = {}
        for v in range(nb):
            if u == v:
                continue
            bu, bv = slice(u * bd, min(n, u * bd + bd)), slice(v * bd, min(n, v * bd + bd))
            if P[bu, bv].any() or P[bv, bu].any():
                exp[v] = D[bu, bv].sum() + D[bv, bu].sum()
        assert rows[u].keys() == exp.keys()
        for v in exp:
            assert abs(rows[u][v] - exp[v]) <= 1e-13 * exp[v]
            assert rows[u][v] == rows[v][u]  # symmetric


def test_quantize_end_points():
    q = BG.quantize(np.array([2.0, 5.0, 3.5, 2.0 + 1e-9]))
    assert q.tolist() == [1, 1000000, 500000, 1]
    assert BG.quantize(np.full(4, 3.0)).tolist() == [1, 1, 1, 1]


@pytest.mark.parametrize("case", ["hex3", "random2"])
def test_host_block_graph_matches_reference(case):
    if case == "hex3":
        p = problems.hex_elasticity(4, E=problems.random_moduli(4, 2))
        A, bd = p.A, 3
    else:
        A, bd = _random(53, 2, 0.06, 3), 2
    ptr, adj, w = block_graph(A, bd)
    rptr, radj, rw = BG.reference_adjacency(A, bd)
    assert np.array_equal(ptr, rptr) and np.array_equal(adj, radj)
    assert np.allclose(w, rw, rtol=1e-14, atol=0.0)
    # Quantised as the reference quantises its undirected edges (each appears twice here).
    assert np.array_equal(BG.quantize(w), BG.quantize(rw))


def test_block_graph_errors():
    A = _random(10, 2, 0.2, 4)
    with pytest.raises(MASError, match="block size"):
        block_graph(A, 4)
    ptr, adj, w = block_graph(sp.diags(np.ones(6)).tocsc(), 3)  # no edges
    assert ptr.tolist() == [0, 0, 0] and adj.size == 0 and w.size == 0
