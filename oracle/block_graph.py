"""CPU oracle of the reference's block adjacency — TEST INFRASTRUCTURE ONLY.

Restates ``build_adjacency_from_csc`` of src/polysolve/linear/mas_utils/BSRAdjacency.cu in numpy:

* every stored scalar entry (row i, column j) maps to the canonical undirected pair of block nodes
  (min, max) of (i // bd, j // bd) with weight |a_ij|; entries inside a diagonal block are not
  edges (``build_keys_abs_flags``, BSRAdjacency.cu:60-84; filtered at :237-283);
* the entries are sorted by pair key and their |a| summed per pair (radix sort and reduce-by-key,
  :288-350): the L1 edge weight;
* the symmetric CSR holds every pair in both directions (histogram, exclusive scan and scatter,
  :352-460; the scatter's order within a row comes from atomics, so rows are compared as sorted
  sets);
* ``quantize`` restates ``quantize_weights`` (:131-162): (w - min) / (max - min) * 10^6 truncated
  to int and clamped to [1, 10^6], all ones when every weight is equal.

Pinned in ``tests/test_block_graph.py`` by what the definitions fix: a hand-built matrix whose
pairs and weights are written out, block sums of |A| formed densely, symmetry, and the
quantisation's end points.  Only ``tests/`` may import this module, and only as the checker.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

MAX_QUANT = 1000000


def reference_adjacency(A, bd: int):
    """(ptr, adj, w): CSR of the undirected block graph, both directions, rows sorted by neighbour,
    w the L1 weight of each edge (BSRAdjacency.cu:184-460 before quantisation)."""
    A = sp.coo_matrix(A)
    n = A.shape[0]
    nb = -(-n // bd)
    rb, cb = A.row // bd, A.col // bd
    off = rb != cb
    u = np.minimum(rb, cb)[off].astype(np.int64)
    v = np.maximum(rb, cb)[off].astype(np.int64)
    a = np.abs(A.data[off])
    key = u * nb + v
    order = np.argsort(key, kind="stable")
    key, a = key[order], a[order]
    uniq, start = np.unique(key, return_index=True)
    w = np.add.reduceat(a, start) if a.size else np.zeros(0)
    eu, ev = uniq // nb, uniq % nb
    src = np.concatenate([eu, ev])
    dst = np.concatenate([ev, eu])
    ww = np.concatenate([w, w])
    order = np.lexsort((dst, src))
    src, dst, ww = src[order], dst[order], ww[order]
    ptr = np.zeros(nb + 1, dtype=np.int64)
    np.add.at(ptr, src + 1, 1)
    return np.cumsum(ptr), dst, ww


def quantize(w: np.ndarray) -> np.ndarray:
    """quantize_weights (BSRAdjacency.cu:131-162) over every undirected edge weight."""
    w = np.asarray(w, dtype=np.float64)
    if w.size == 0:
        return np.zeros(0, dtype=np.int64)
    lo, hi = float(w.min()), float(w.max())
    rng = hi - lo
    if rng <= 0.0:
        return np.ones(w.size, dtype=np.int64)
    q = ((w - lo) / rng * MAX_QUANT).astype(np.int64)  # static_cast<int>: truncation
    return np.clip(q, 1, MAX_QUANT)
