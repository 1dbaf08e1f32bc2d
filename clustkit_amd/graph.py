"""Drop-in for clustkit/graph.py:7-39 (``build_similarity_graph``) on the GPU.

The symmetric (N, N) float32 CSR of the passing pairs is built by
ck_graph_csr_host (entries (i, j, w) and (j, i, w), stable radix sort by
(row, column), duplicate entries summed, explicit zeros kept), with the
layout scipy's coo -> csr conversion gives the reference: int32 indptr and
indices, sorted column indices, canonical format.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse

from . import _lib


def build_similarity_graph(num_sequences: int, filtered_pairs: np.ndarray,
                           similarities: np.ndarray) -> sparse.csr_matrix:
    """graph.py:7-39: (N, N) symmetric CSR weighted by similarity."""
    n = int(num_sequences)
    if len(filtered_pairs) == 0:
        return sparse.csr_matrix((n, n), dtype=np.float32)
    lib = _lib.require_gpu()
    pairs = np.ascontiguousarray(filtered_pairs, np.int32).reshape(-1, 2)
    sims = np.ascontiguousarray(similarities, np.float32).reshape(-1)
    if len(sims) != len(pairs):
        raise ValueError("similarities and filtered_pairs differ in length")
    k = len(pairs)
    indptr = np.zeros(n + 1, np.int32)
    indices = np.empty(2 * k, np.int32)
    data = np.empty(2 * k, np.float32)
    nnz = np.zeros(1, np.int64)
    _lib.check(lib.ck_graph_csr_host(pairs.ctypes.data, k, sims.ctypes.data, n,
                                     indptr.ctypes.data, indices.ctypes.data, data.ctypes.data,
                                     nnz.ctypes.data))
    m = int(nnz[0])
    g = sparse.csr_matrix((data[:m].copy(), indices[:m].copy(), indptr), shape=(n, n))
    g.has_sorted_indices = True
    return g
