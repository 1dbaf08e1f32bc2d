"""CPU checks of the sketch / LSH oracle (oracle/sketch_lsh.py): its
vectorised form against its step-by-step restatement of the reference's
online bottom-s buffer (clustkit/sketch.py:29-80) and pair extraction
(clustkit/lsh.py:52-90), and against closed forms."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import sketch_lsh as O


def _pack(seqs):
    ln = np.array([len(s) for s in seqs], np.int64)
    off = np.zeros(len(seqs), np.int64)
    if len(seqs):
        off[1:] = np.cumsum(ln)[:-1]
    flat = np.concatenate(seqs) if len(seqs) else np.zeros(0, np.uint8)
    return flat, off, ln


@pytest.mark.parametrize("k,s,mode", [(5, 16, "protein"), (3, 8, "nucleotide"),
                                      (14, 4, "protein"), (1, 32, "protein"),
                                      (20, 8, "protein")])
def test_vectorised_sketch_equals_online_loop(k, s, mode):
    """Random sequences 0-70 long with codes up to 21 (unknown residues enter
    the code), repeated k-mers (duplicates stay in the bottom-s), sequences
    shorter than k (zeros) and with fewer than s windows (padding); k=20
    makes the base-20 code wrap mod 2^64."""
    rng = np.random.default_rng(k * 100 + s)
    seqs = [rng.integers(0, 22, int(rng.integers(0, 70))).astype(np.uint8) for _ in range(60)]
    seqs.append(np.tile(np.array([1, 2, 3], np.uint8), 30))
    flat, off, ln = _pack(seqs)
    sk = O.compute_sketches(flat, off, ln, k, s, mode, seed=7)
    a = 20 if mode == "protein" else 4
    for i, q in enumerate(seqs):
        assert [int(x) for x in sk[i]] == O.sketch_one_loop(q, k, s, a, 7), (k, s, i)


def test_sketch_closed_forms():
    """fmix64 of 0 is 0 (so code 0 ^ seed 0 hashes to 0); a sequence of one
    repeated residue has one distinct k-mer, so its sketch is that hash
    repeated min(windows, s) times; fmix64 is a bijection (distinct codes,
    distinct hashes)."""
    assert int(O.fmix64(np.array([0], np.uint64), 0)[0]) == 0
    seq = np.full(10, 4, np.uint8)
    flat, off, ln = _pack([seq])
    sk = O.compute_sketches(flat, off, ln, k=3, sketch_size=16, seed=42)
    h = int(O.fmix64(np.array([4 * 400 + 4 * 20 + 4], np.uint64), 42)[0])
    assert [int(x) for x in sk[0][:8]] == [h] * 8
    assert all(int(x) == (1 << 64) - 1 for x in sk[0][8:])
    codes = np.arange(100000, dtype=np.uint64)
    assert len(np.unique(O.fmix64(codes, 42))) == len(codes)


@pytest.mark.parametrize("max_bucket", [1000, 3])
def test_vectorised_lsh_equals_loop_extraction(max_bucket):
    """Sketches with planted identical rows (buckets of 2-9 members in every
    table, above a cap of 3 when max_bucket=3) and random rows; the
    vectorised extraction equals the reference's loop applied to a stable
    argsort of each table's bucket ids."""
    rng = np.random.default_rng(5)
    n, s = 40, 16
    sk = rng.integers(0, 2 ** 63, size=(n, s), dtype=np.int64).astype(np.uint64)
    for g, size in enumerate((2, 4, 9)):
        rows = rng.choice(n, size=size, replace=False)
        sk[rows] = sk[rows[0]]
    sk[rng.choice(n, 6, replace=False), 3] = 12345           # one shared position
    T, b = 12, 2
    got = O.lsh_candidates(sk, T, b, seed=42, max_bucket=max_bucket)
    bands, seeds = O.draw_tables(s, T, b, 42)
    packed = set()
    for t in range(T):
        ids = O.band_hash(sk, bands[t], int(seeds[t]))
        order = np.argsort(ids, kind="stable")
        packed.update(O.extract_pairs_loop(order, ids[order], n, max_bucket))
    want = sorted(packed)
    assert [int(i) * n + int(j) for i, j in got] == want
    assert all(i < j for i, j in got)


def test_lsh_identical_rows_pair_up():
    sk = np.array([[1, 2, 3], [1, 2, 3], [7, 8, 9], [1, 2, 3]], np.uint64)
    pairs = O.lsh_candidates(sk, num_tables=8, num_bands=2)
    assert {tuple(p) for p in pairs} == {(0, 1), (0, 3), (1, 3)}


def test_band_hash_closed_form():
    """One band position: h = fmix-tail(seed * M1 + v) (lsh.py:20-29)."""
    sk = np.array([[5, 6], [7, 8]], np.uint64)
    got = O.band_hash(sk, np.array([1], np.int32), 9)
    M = (1 << 64) - 1
    for r, v in enumerate((6, 8)):
        h = (9 * 0x517CC1B727220A95 + v) & M
        h = ((h ^ (h >> 33)) * 0xC4CEB9FE1A85EC53) & M
        h ^= h >> 33
        assert int(got[r]) == h
