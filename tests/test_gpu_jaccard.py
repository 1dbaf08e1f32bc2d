"""GPU parity of the Jaccard pre-filter (jaccard.hip through ck_jaccard_host
and ck_jaccard_device; pairwise._batch_jaccard, _jaccard_prefilter,
compute_pairwise_jaccard, jaccard_from_sketches) against the oracle
(oracle.jaccard_batch, oracle.jaccard_prefilter: pairwise.py:949-1001,
1564-1612).  Similarities are compared as float32 bit patterns, masks and
pair lists exactly."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX = np.iinfo(np.uint64).max


def _device(pairs, sk):
    import torch
    from clustkit_amd.device import jaccard_device
    dev = torch.device("cuda:0")
    d_p = torch.from_numpy(np.ascontiguousarray(pairs, np.int32)).to(dev)
    d_s = torch.from_numpy(np.ascontiguousarray(sk, np.uint64).view(np.int64)).to(dev)
    out = jaccard_device(d_p, d_s)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _host(pairs, sk):
    from clustkit_amd.pairwise import _batch_jaccard
    return _batch_jaccard(pairs, sk, 0.0)[0]


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


def test_gpu_jaccard_reference_known_answers(oracle_mod):
    """The reference test suite's cases (tests/test_pairwise.py:14-68),
    through the drop-in functions, and equal to the oracle bit for bit."""
    from clustkit_amd.pairwise import compute_pairwise_jaccard, jaccard_from_sketches
    s5 = np.array([1, 2, 3, 4, 5], np.uint64)
    assert jaccard_from_sketches(s5, s5) == 1.0
    assert jaccard_from_sketches(s5, np.array([6, 7, 8, 9, 10], np.uint64)) == 0.0
    part = jaccard_from_sketches(s5, np.array([1, 2, 6, 7, 8], np.uint64))
    assert 0.0 < part < 1.0
    pa = np.array([1, 2, 3, MAX, MAX], np.uint64)
    pb = np.array([1, 2, 4, MAX, MAX], np.uint64)
    padded = jaccard_from_sketches(pa, pb)
    assert 0.0 < padded < 1.0
    sk = np.stack([s5, np.array([1, 2, 6, 7, 8], np.uint64), pa, pb])
    want = oracle_mod.jaccard_batch(np.array([[0, 1], [2, 3]], np.int32), sk)
    assert np.float32(part) == want[0] and np.float32(padded) == want[1]

    sketches = np.array([[1, 2, 3, 4, 5], [1, 2, 3, 4, 5], [6, 7, 8, 9, 10]], np.uint64)
    f, s = compute_pairwise_jaccard(np.array([[0, 1], [0, 2]], np.int32), sketches, 0.5)
    assert f.tolist() == [[0, 1]] and s.tolist() == [1.0]
    f, s = compute_pairwise_jaccard(np.empty((0, 2), np.int32), sketches[:1, :3], 0.5)
    assert len(f) == 0 and len(s) == 0
    f, s = compute_pairwise_jaccard(np.array([[0, 1]], np.int32), sketches[[0, 2]], 0.5)
    assert len(f) == 0


def _edge_rows(rng, n, s):
    """Sorted bottom-s style rows: values from a small universe (overlaps),
    duplicates (a multiset sketch), a random all-ones padding suffix, rows
    of only padding, all-zero rows (sequences shorter than k), one-value
    rows, and values at the top of the 64-bit range."""
    rows = []
    for i in range(n):
        kind = i % 7
        if kind == 0:
            r = np.full(s, MAX, np.uint64)
        elif kind == 1:
            r = np.zeros(s, np.uint64)
        elif kind == 2:
            r = np.full(s, rng.integers(0, 4), np.uint64)
        else:
            v = rng.integers(0, 3 * s, s).astype(np.uint64)
            if kind == 3:
                v[: s // 4] = MAX - rng.integers(1, 8, s // 4).astype(np.uint64)
            if kind == 4:
                v = np.repeat(v[: s // 3 + 1], 3)[:s]
            v.sort()
            pad = int(rng.integers(0, s + 1)) if kind >= 5 else 0
            if pad:
                v[s - pad:] = MAX
            r = v
        rows.append(r)
    return np.stack(rows)


@pytest.mark.parametrize("s", [16, 128, 1024])
def test_gpu_jaccard_host_and_device_match_oracle(oracle_mod, s):
    print(f"jaccard edge rows s={s}", flush=True)
    rng = np.random.default_rng(s)
    sk = _edge_rows(rng, 70, s)
    n = len(sk)
    grid = np.stack(np.meshgrid(np.arange(n), np.arange(n)), -1).reshape(-1, 2)
    rand = rng.integers(0, n, (3000, 2))
    pairs = np.concatenate([grid, rand]).astype(np.int32)
    want = oracle_mod.jaccard_batch(pairs, sk)
    # every distinct outcome class occurs: 0, 1, strictly between
    assert (want == 0).any() and (want == 1).any() and ((want > 0) & (want < 1)).any()
    assert _bits_equal(_host(pairs, sk), want)
    assert _bits_equal(_device(pairs, sk), want)


def test_gpu_jaccard_on_device_sketches(oracle_mod):
    """Sketches written by ck_sketch_device (duplicates from repeats, rows
    of padding for sequences with fewer than s windows, zero rows for
    sequences shorter than k), then Jaccard on the device."""
    import torch
    from clustkit_amd.device import DeviceSequences
    from clustkit_amd.sketch import compute_sketches_device
    rng = np.random.default_rng(21)
    seqs = [rng.integers(0, 20, int(rng.integers(0, 600))).astype(np.uint8) for _ in range(200)]
    seqs += [np.tile(np.arange(4, dtype=np.uint8), 150), np.zeros(3, np.uint8),
             np.zeros(0, np.uint8), rng.integers(0, 20, 40).astype(np.uint8)]
    base = rng.integers(0, 20, 500).astype(np.uint8)
    for _ in range(20):                      # near copies: large Jaccard values
        v = base.copy()
        hit = rng.random(500) < 0.05
        v[hit] = rng.integers(0, 20, int(hit.sum()))
        seqs.append(v)
    ln = np.array([len(x) for x in seqs], np.int32)
    off = np.zeros(len(seqs), np.int64)
    off[1:] = np.cumsum(ln.astype(np.int64))[:-1]
    flat = np.concatenate(seqs)
    d_sk = compute_sketches_device(DeviceSequences.upload(flat, off, ln), 5, 128, "protein", 42)
    torch.cuda.synchronize()
    sk = d_sk.cpu().numpy().view(np.uint64)
    n = len(seqs)
    pairs = np.concatenate([rng.integers(0, n, (20000, 2)),
                            np.stack(np.triu_indices(n, 1), 1)[-3000:]]).astype(np.int32)
    want = oracle_mod.jaccard_batch(pairs, sk)
    assert (want >= 0.5).sum() > 50
    from clustkit_amd.device import jaccard_device
    got = jaccard_device(torch.from_numpy(pairs).cuda(), d_sk)
    torch.cuda.synchronize()
    assert _bits_equal(got.cpu().numpy(), want)
    assert _bits_equal(_host(pairs, sk), want)


@pytest.mark.parametrize("case", ["sampled_applied", "sampled_skipped", "unsampled"])
def test_gpu_jaccard_prefilter_decision_and_mask(oracle_mod, case):
    """_jaccard_prefilter: the RandomState(0) 2000-pair sample decides
    whether the filter applies (None when the estimated removal is below
    5%), then mask = sims >= floor -- the same decision and the same mask as
    the oracle's restatement, for thresholds 0.5 and 0.3 (floors 0.009375
    and 0.001)."""
    from clustkit_amd.pairwise import _jaccard_prefilter
    rng = np.random.default_rng({"sampled_applied": 1, "sampled_skipped": 2,
                                 "unsampled": 3}[case])
    s = 128
    n = 400
    sk = np.sort(rng.integers(0, 1 << 40, (n, s)).astype(np.uint64), axis=1)
    sk[n // 2:] = sk[: n // 2]                  # half the rows duplicated: Jaccard 1
    if case != "sampled_skipped":
        sk[n // 2::3, : s // 2] = sk[0, : s // 2]   # partial overlaps
        sk.sort(axis=1)
    m = {"sampled_applied": 30000, "sampled_skipped": 30000, "unsampled": 1500}[case]
    pairs = rng.integers(0, n, (m, 2)).astype(np.int32)
    if case == "sampled_skipped":
        # 97% of the pairs are (i, i + n/2): identical sketches
        same = rng.random(m) < 0.97
        pairs[same, 1] = (pairs[same, 0] + n // 2) % n
    for t in (0.5, 0.3):
        got = _jaccard_prefilter(pairs, sk, t)
        want = oracle_mod.jaccard_prefilter(pairs, sk, t)
        if case == "sampled_skipped":
            assert got is None and want is None
        else:
            assert got is not None and want is not None
            assert got.dtype == np.bool_ and np.array_equal(got, want)
            assert 0 < got.sum() < m
