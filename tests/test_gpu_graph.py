"""GPU similarity-graph construction (ck_graph_csr_*) against the
reference's own construction (clustkit/graph.py:7-39: scipy coo -> csr):
identical indptr, indices and data bits."""

import numpy as np
import pytest

from oracle.graph_oracle import reference_graph
from test_graph_oracle import _cases

pytestmark = pytest.mark.gpu


def _same(g, ref):
    assert g.shape == ref.shape and g.dtype == np.float32
    assert np.array_equal(g.indptr, ref.indptr)
    assert np.array_equal(g.indices, ref.indices)
    assert np.array_equal(g.data.view(np.uint32), ref.data.view(np.uint32))


def test_gpu_graph_small_cases():
    from clustkit_amd.graph import build_similarity_graph
    for q, (n, pairs, sims) in enumerate(_cases()):
        print(f"graph case {q}: n={n} k={len(pairs)}", flush=True)
        _same(build_similarity_graph(n, pairs, sims), reference_graph(n, pairs, sims))


def test_gpu_graph_large():
    from clustkit_amd.graph import build_similarity_graph
    rng = np.random.default_rng(8)
    n = 200_000
    pairs = rng.integers(0, n, (1_500_000, 2)).astype(np.int32)
    pairs = np.unique(np.sort(pairs, axis=1), axis=0)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    rng.shuffle(pairs)
    sims = rng.random(len(pairs)).astype(np.float32)
    print(f"graph large: n={n} k={len(pairs)}", flush=True)
    _same(build_similarity_graph(n, pairs, sims), reference_graph(n, pairs, sims))


def test_gpu_graph_rejects_bad_index():
    from clustkit_amd import _lib
    from clustkit_amd.graph import build_similarity_graph
    with pytest.raises(_lib.HipLibraryError):
        build_similarity_graph(5, np.array([[0, 5]], np.int32), np.ones(1, np.float32))


def test_gpu_graph_device_entry():
    """ck_graph_csr_device on HBM tensors equals the reference construction."""
    import torch

    from clustkit_amd.device import graph_csr_device
    n, pairs, sims = next(iter(_cases()))
    ip, ix, d, nnz = graph_csr_device(torch.from_numpy(pairs).cuda(),
                                      torch.from_numpy(sims).cuda(), n)
    torch.cuda.synchronize()
    ref = reference_graph(n, pairs, sims)
    m = int(nnz.item())
    assert m == ref.nnz
    assert np.array_equal(ip.cpu().numpy(), ref.indptr)
    assert np.array_equal(ix[:m].cpu().numpy(), ref.indices)
    assert np.array_equal(d[:m].cpu().numpy().view(np.uint32), ref.data.view(np.uint32))
