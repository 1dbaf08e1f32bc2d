"""Similarity-graph oracle -- TEST INFRASTRUCTURE ONLY (only tests/ and
bench.py's CPU comparison use it).

``reference_graph`` is the construction of clustkit/graph.py:7-39 itself:
stack (i, j, w) and (j, i, w) into a float32 COO matrix and convert it with
scipy's tocsr() (the library call the reference makes; scipy is the
reference's own dependency and present here).  ``restated_graph`` restates
that conversion with NumPy (stable sort by (row, column), duplicates summed
in entry order, rows counted) and is pinned against ``reference_graph`` in
tests/test_graph_oracle.py.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse


def reference_graph(n, pairs, sims):
    if len(pairs) == 0:
        return sparse.csr_matrix((n, n), dtype=np.float32)
    pairs = np.asarray(pairs)
    rows, cols = pairs[:, 0], pairs[:, 1]
    coo = sparse.coo_matrix((np.concatenate([sims, sims]),
                             (np.concatenate([rows, cols]), np.concatenate([cols, rows]))),
                            shape=(n, n), dtype=np.float32)
    return coo.tocsr()


def restated_graph(n, pairs, sims):
    """(indptr, indices, data) with int32 indices and float32 data."""
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    sims = np.asarray(sims, np.float32)
    r = np.concatenate([pairs[:, 0], pairs[:, 1]])
    c = np.concatenate([pairs[:, 1], pairs[:, 0]])
    w = np.concatenate([sims, sims])
    order = np.lexsort((c, r))                 # stable: entry order within equal keys
    r, c, w = r[order], c[order], w[order]
    if len(r):
        head = np.ones(len(r), bool)
        head[1:] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
        starts = np.flatnonzero(head)
        ends = np.append(starts[1:], len(r))
        data = np.empty(len(starts), np.float32)
        for q, (a, b) in enumerate(zip(starts, ends)):
            s = w[a]
            for x in range(a + 1, b):
                s = np.float32(s + w[x])
            data[q] = s
        ur, uc = r[starts], c[starts]
    else:
        data = np.zeros(0, np.float32)
        ur = uc = np.zeros(0, np.int64)
    indptr = np.searchsorted(ur, np.arange(n + 1), side="left").astype(np.int32)
    return indptr, uc.astype(np.int32), data
