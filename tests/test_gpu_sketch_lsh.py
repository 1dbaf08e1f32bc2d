"""GPU parity of the device MinHash sketches and LSH candidate pairs
(ck_sketch_device, ck_lsh_candidates_device) against the CPU oracle
(oracle/sketch_lsh.py).  Bit-exact: every sketch word, every pair."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import sketch_lsh as O

pytestmark = pytest.mark.gpu


def _pack(seqs):
    ln = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(len(seqs), np.int64)
    if len(seqs):
        off[1:] = np.cumsum(ln.astype(np.int64))[:-1]
    flat = np.concatenate(seqs) if len(seqs) else np.zeros(0, np.uint8)
    return flat, off, ln


def _dev_sketch(flat, off, ln, k, s, mode, seed=42):
    import torch
    from clustkit_amd.device import DeviceSequences
    from clustkit_amd.sketch import compute_sketches_device
    d = compute_sketches_device(DeviceSequences.upload(flat, off, ln), k, s, mode, seed)
    torch.cuda.synchronize()
    return d.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("k,s,mode", [(5, 128, "protein"), (3, 32, "nucleotide"),
                                      (4, 100, "protein"), (20, 16, "protein"),
                                      (2, 1024, "protein"), (1, 7, "nucleotide")])
def test_gpu_sketches_match_oracle(k, s, mode):
    """Lengths 0 to 7000 (several 2048-word LDS chunks), codes up to 24
    (unknown residues enter the code), repeated k-mers, fewer windows than
    s, s not a power of two, the largest supported s, k=20 (wrapping code)."""
    print(f"sketch k={k} s={s} {mode}", flush=True)
    rng = np.random.default_rng(k * 1000 + s)
    hi = 20 if mode == "protein" else 4
    seqs = [rng.integers(0, hi, int(rng.integers(0, 700))).astype(np.uint8) for _ in range(300)]
    seqs += [rng.integers(0, 25, 400).astype(np.uint8), np.zeros(0, np.uint8),
             np.full(3, 1, np.uint8), rng.integers(0, hi, 7000).astype(np.uint8),
             np.tile(np.arange(3, dtype=np.uint8), 900)]
    flat, off, ln = _pack(seqs)
    want = O.compute_sketches(flat, off, ln, k, s, mode, seed=42)
    got = _dev_sketch(flat, off, ln, k, s, mode)
    bad = np.flatnonzero((got != want).any(axis=1))
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}"


def test_gpu_sketch_seed_and_signature():
    """Reference signature (padded matrix input) and a seed above 2^31."""
    from clustkit_amd.sketch import compute_sketches
    rng = np.random.default_rng(3)
    seqs = [rng.integers(0, 20, int(rng.integers(1, 90))).astype(np.uint8) for _ in range(64)]
    flat, off, ln = _pack(seqs)
    enc = np.zeros((len(seqs), int(ln.max())), np.uint8)
    for i, q in enumerate(seqs):
        enc[i, :len(q)] = q
    seed = (1 << 40) + 17
    want = O.compute_sketches(flat, off, ln, 5, 64, "protein", seed=seed)
    got = compute_sketches(enc, ln, 5, 64, "protein", seed=seed)
    assert got.dtype == np.uint64 and np.array_equal(got, want)
    got2 = compute_sketches(None, ln, 5, 64, "protein", seed=seed, flat_sequences=flat,
                            offsets=off, device="0")
    assert np.array_equal(got2, want)


def _dev_lsh(sk, T, b, seed=42, max_bucket=1000):
    import torch
    from clustkit_amd.lsh import lsh_candidates_device
    d = torch.from_numpy(np.ascontiguousarray(sk).view(np.int64)).to("cuda:0")
    out = lsh_candidates_device(d, T, b, seed, max_bucket)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_gpu_lsh_matches_oracle_on_families(small_family):
    """The family fixture (300 proteins, k=5, s=128) with the automatic
    parameters of t=0.5 (146 tables x 1 band) and of t=0.9 (3 bands), and
    2 bands x 40 tables."""
    from clustkit_amd.lsh import auto_lsh_params
    fam = small_family
    sk = O.compute_sketches(fam.flat, fam.offsets, fam.lengths, 5, 128, "protein", seed=42)
    for t in (0.5, 0.9):
        p = auto_lsh_params(t, "medium", 5)
        print(f"lsh families t={t} {p}", flush=True)
        want = O.lsh_candidates(sk, p["num_tables"], p["num_bands"], seed=42)
        got = _dev_lsh(sk, p["num_tables"], p["num_bands"])
        assert got.shape == want.shape and np.array_equal(got, want), t
    want = O.lsh_candidates(sk, 40, 2, seed=7)
    assert np.array_equal(_dev_lsh(sk, 40, 2, seed=7), want)


def test_gpu_lsh_bucket_cap_groups_and_edges(monkeypatch):
    """Buckets above the cap (identical rows, max_bucket 3 and 1000 with a
    1,500-member bucket), tables split into several groups and merged
    (CK_LSH_GROUP_ELEMS), fewer than two sequences, no tables."""
    rng = np.random.default_rng(9)
    n, s = 3000, 32
    sk = rng.integers(0, 2 ** 63, size=(n, s), dtype=np.int64).astype(np.uint64)
    sk[100:1600] = sk[100]                                    # one bucket of 1,500
    for g in range(50):
        rows = rng.choice(n, size=int(rng.integers(2, 8)), replace=False)
        sk[rows, :8] = sk[rows[0], :8]
    sk[rng.choice(n, 40, replace=False), 5] = 99
    for cap in (1000, 3):
        print(f"lsh cap={cap}", flush=True)
        want = O.lsh_candidates(sk, 24, 1, seed=42, max_bucket=cap)
        got = _dev_lsh(sk, 24, 1, max_bucket=cap)
        assert np.array_equal(got, want), cap
    monkeypatch.setenv("CK_LSH_GROUP_ELEMS", str(5 * n))        # 5 tables per group
    for b in (2, 1):
        want = O.lsh_candidates(sk, 24, b, seed=3, max_bucket=1000)
        got = _dev_lsh(sk, 24, b, seed=3)
        assert np.array_equal(got, want), b
    monkeypatch.delenv("CK_LSH_GROUP_ELEMS")
    assert _dev_lsh(sk[:1], 8, 1).shape == (0, 2)
    assert _dev_lsh(sk, 0, 1).shape == (0, 2)


def test_gpu_lsh_reference_signature():
    from clustkit_amd.lsh import lsh_candidates
    sk = np.array([[1, 2, 3], [1, 2, 3], [7, 8, 9], [1, 2, 3]], np.uint64)
    pairs = lsh_candidates(sk, num_tables=8, num_bands=2)
    assert pairs.dtype == np.int32
    assert [tuple(p) for p in pairs] == [(0, 1), (0, 3), (1, 3)]
