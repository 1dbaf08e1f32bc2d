"""The NumPy restatement of the similarity-graph construction against the
reference's own construction (clustkit/graph.py:7-39: scipy coo -> csr)."""

import numpy as np

from oracle.graph_oracle import reference_graph, restated_graph


def _cases():
    rng = np.random.default_rng(3)
    n = 300
    pairs = rng.integers(0, n, (2000, 2)).astype(np.int32)
    pairs = pairs[pairs[:, 0] < pairs[:, 1]]
    pairs = np.unique(pairs, axis=0)
    rng.shuffle(pairs)
    sims = rng.random(len(pairs)).astype(np.float32)
    yield n, pairs, sims                                     # unique i < j (the pipeline case)
    flip = pairs[:50, ::-1].copy()                           # both orientations of 50 pairs
    yield n, np.concatenate([pairs, flip]), np.concatenate([sims, rng.random(50).astype(np.float32)])
    self_p = np.array([[5, 5], [7, 7]], np.int32)            # self pairs -> 2 entries each
    yield n, np.concatenate([pairs, self_p]), np.concatenate([sims, np.float32([0.5, 0.25])])
    z = sims.copy()
    z[::7] = 0.0                                             # explicit zeros kept
    yield n, pairs, z
    yield 10, np.zeros((0, 2), np.int32), np.zeros(0, np.float32)
    yield 1000, pairs, sims                                  # isolated rows at the end


def test_restated_graph_matches_scipy():
    for n, pairs, sims in _cases():
        ref = reference_graph(n, pairs, sims)
        ip, ix, d = restated_graph(n, pairs, sims)
        assert ref.dtype == np.float32
        assert np.array_equal(ref.indptr, ip)
        assert np.array_equal(ref.indices, ix)
        assert np.array_equal(ref.data.view(np.uint32), d.view(np.uint32))
