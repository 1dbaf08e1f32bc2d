"""Pins the CPU oracle before it is trusted as the checker.

* golden vectors written by the reference kernel compiled from its own
  sources (tests/golden/make_golden.py) -- bit-exact;
* the reference kernel itself where oracle/_ref is built (same inputs, new
  random draws);
* brute force over every local alignment of tiny sequences (score);
* closed forms: self-alignment of any sequence scores the diagonal sum with
  identity 1.0; disjoint residues with negative scores give 0;
* invariants: widening the band never lowers the score; matches <= shorter;
* Jaccard: the known answers of the reference tests
  (tests/test_pairwise.py:14-68) and the bottom-s set formula.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from clustkit_amd.scoring import BLOSUM62
from conftest import golden_files, load_golden


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.stem)
def test_oracle_matches_reference_golden(path, oracle_mod):
    g = load_golden(path)
    o = oracle_mod.sw_batch(g["pairs"], g["flat"], g["offsets"], g["lengths"], int(g["band"]),
                            BLOSUM62, hints=g["hints"])
    assert np.array_equal(o["scores"], g["scores"])
    assert np.array_equal(o["sims"].view(np.uint32), g["sims"].view(np.uint32))
    assert np.array_equal(o["mask"], g["mask"])


def test_oracle_matches_reference_build_random(oracle_mod, small_family):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built here (reference sources absent)")
    fam = small_family
    rng = np.random.default_rng(5)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(1500, 2)).astype(np.int32)
    pairs[:1000, 1] = (pairs[:1000, 0] // 10) * 10 + rng.integers(0, 10, 1000)
    hints = rng.integers(-120, 120, len(pairs)).astype(np.int32)
    for band, h in ((100, None), (12, hints), (0, None)):
        o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, band, BLOSUM62, h)
        s, sc, mk = oracle_mod.ref_sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, band,
                                            BLOSUM62, h)
        assert np.array_equal(o["scores"], sc)
        assert np.array_equal(o["sims"].view(np.uint32), s.view(np.uint32))
        assert np.array_equal(o["mask"], mk)


def _brute_local_score(a, b):
    """Best local alignment score by enumerating every alignment: a path of
    match (both advance), gap-in-b (row advances) and gap-in-a (column
    advances) steps between any start and end, each maximal gap run of
    length L costing 10 + L.  Alignments are enumerated from every start
    cell with explicit recursion over step sequences."""
    best = 0
    la, lb = len(a), len(b)

    def sub(x, y):
        return int(BLOSUM62[x, y]) if x < 20 and y < 20 else -4

    def walk(i, j, score, last):
        nonlocal best
        if last == "M" and score > best:
            best = score
        if i < la and j < lb:
            walk(i + 1, j + 1, score + sub(a[i], b[j]), "M")
        if i < la and last != "":
            walk(i + 1, j, score - (1 if last == "D" else 11), "D")
        if j < lb and last != "":
            walk(i, j + 1, score - (1 if last == "I" else 11), "I")

    for i0 in range(la):
        for j0 in range(lb):
            walk(i0, j0, 0, "")
    return best


def test_oracle_brute_force_tiny(oracle_mod):
    rng = np.random.default_rng(3)
    alphabet = np.array([0, 1, 8, 9, 17, 18, 20], np.uint8)   # includes unknown (20)
    for _ in range(80):
        la, lb = rng.integers(1, 5, 2)
        a = rng.choice(alphabet, la)
        b = rng.choice(alphabet, lb)
        if rng.random() < 0.5:                                  # plant shared words
            k = min(la, lb)
            b[:k] = a[:k] if rng.random() < 0.5 else b[:k]
        r = oracle_mod.sw_pair(a, b, 0, BLOSUM62)
        assert r["score"] == _brute_local_score(a, b), (a, b)
        assert 0 <= r["matches"] <= min(la, lb)


def test_oracle_self_alignment_closed_form(oracle_mod):
    # BLOSUM62 is diagonally dominant in the sense B[x,y] <= (B[x,x]+B[y,y])/2,
    # so no alignment of a sequence with itself beats the identity path.
    d = np.diag(BLOSUM62).astype(np.int64)
    assert np.all(2 * BLOSUM62.astype(np.int64) <= d[:, None] + d[None, :])
    rng = np.random.default_rng(9)
    for n in (1, 7, 51, 333):
        a = rng.integers(0, 20, n).astype(np.uint8)
        for band in (0, 5, 100):
            r = oracle_mod.sw_pair(a, a, band, BLOSUM62)
            assert r["score"] == int(d[a].sum())
            assert r["matches"] == n and r["identity"] == np.float32(1.0)
            assert r["end"] == (n, n)


def test_oracle_disjoint_negative_residues(oracle_mod):
    # W (18) vs D (2) scores -4 everywhere; no positive cell exists.
    r = oracle_mod.sw_pair(np.full(30, 18, np.uint8), np.full(45, 2, np.uint8), 10, BLOSUM62)
    assert r["score"] == 0 and r["matches"] == 0 and r["identity"] == 0


def test_oracle_band_monotone(oracle_mod, small_family):
    fam = small_family
    rng = np.random.default_rng(2)
    n = len(fam.lengths)
    pairs = np.stack([rng.integers(0, n, 300), rng.integers(0, n, 300)], 1).astype(np.int32)
    pairs[:200, 1] = (pairs[:200, 0] // 10) * 10 + rng.integers(0, 10, 200)
    prev = None
    for band in (1, 2, 5, 13, 40, 120, 0):
        o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, band, BLOSUM62)
        lmin = np.minimum(fam.lengths[pairs[:, 0]], fam.lengths[pairs[:, 1]])
        assert np.all(o["matches"] <= lmin) and np.all(o["scores"] >= 0)
        if prev is not None:
            assert np.all(o["scores"] >= prev)
        prev = o["scores"]


def test_identity_division_is_float32_correctly_rounded():
    # identity = fl32(m) / fl32(s) in C; NumPy float32 division is the same
    # IEEE operation; double-then-round agrees for every m <= s <= 8000.
    s = np.arange(1, 8001, dtype=np.int64)
    for m0 in range(0, 8001, 500):
        m = np.arange(m0, min(m0 + 500, 8001), dtype=np.int64)
        mm, ss = np.meshgrid(m, s, indexing="ij")
        ok = mm <= ss
        f32 = mm[ok].astype(np.float32) / ss[ok].astype(np.float32)
        f64 = (mm[ok].astype(np.float64) / ss[ok].astype(np.float64)).astype(np.float32)
        assert np.array_equal(f32, f64)


# --- Jaccard -------------------------------------------------------------

MAXV = np.iinfo(np.uint64).max


def _jac(oracle_mod, a, b):
    sk = np.stack([np.asarray(a, np.uint64), np.asarray(b, np.uint64)])
    return float(oracle_mod.jaccard_batch(np.array([[0, 1]], np.int32), sk)[0])


def test_jaccard_reference_known_answers(oracle_mod):
    # reference tests/test_pairwise.py:14-36
    assert _jac(oracle_mod, [1, 2, 3, 4, 5], [1, 2, 3, 4, 5]) == 1.0
    assert _jac(oracle_mod, [1, 2, 3, 4, 5], [6, 7, 8, 9, 10]) == 0.0
    assert 0.0 < _jac(oracle_mod, [1, 2, 3, 4, 5], [1, 2, 6, 7, 8]) < 1.0
    assert 0.0 < _jac(oracle_mod, [1, 2, 3, MAXV, MAXV], [1, 2, 4, MAXV, MAXV]) < 1.0
    # tests/test_pairwise.py:38-68 (filter at 0.5)
    sk = np.array([[1, 2, 3, 4, 5], [1, 2, 3, 4, 5], [6, 7, 8, 9, 10]], np.uint64)
    s = oracle_mod.jaccard_batch(np.array([[0, 1], [0, 2]], np.int32), sk)
    assert list(s >= np.float32(0.5)) == [True, False] and s[0] == 1.0


def test_jaccard_bottom_s_formula(oracle_mod):
    # Distinct-valued sketches: the walk visits the s smallest values of the
    # union (or stops when both reach padding); shared/visited is the classic
    # bottom-s estimator.
    rng = np.random.default_rng(4)
    for _ in range(300):
        s = int(rng.integers(1, 40))
        pool = rng.choice(1000, size=3 * s, replace=False).astype(np.uint64)
        a = np.sort(rng.choice(pool, size=int(rng.integers(0, s + 1)), replace=False))
        b = np.sort(rng.choice(pool, size=int(rng.integers(0, s + 1)), replace=False))
        A = np.concatenate([a, np.full(s - len(a), MAXV, np.uint64)])
        B = np.concatenate([b, np.full(s - len(b), MAXV, np.uint64)])
        union = np.union1d(a, b)
        # exhaustion: the walk stops when either array's s slots are consumed
        steps, shared, ia, ib = 0, 0, 0, 0
        u = list(union)
        for v in u:
            if steps == s or ia == s or ib == s:
                break
            in_a, in_b = v in set(a.tolist()), v in set(b.tolist())
            shared += in_a and in_b
            ia += in_a
            ib += in_b
            steps += 1
        want = np.float32(shared / steps) if steps else np.float32(0)
        got = oracle_mod.jaccard_batch(np.array([[0, 1]], np.int32), np.stack([A, B]))[0]
        assert got == want, (A, B)


def test_jaccard_division_double_vs_float():
    # The Numba reference divides in double and rounds; the CUDA twin divides
    # in float.  For every shared <= union <= 1024 both round identically.
    u = np.arange(1, 1025, dtype=np.int64)
    for sh in range(0, 1025):
        uu = u[u >= sh]
        f32 = np.float32(sh) / uu.astype(np.float32)
        f64 = (sh / uu.astype(np.float64)).astype(np.float32)
        assert np.array_equal(f32, f64)


def test_length_ratio_closed_form(oracle_mod):
    lengths = np.array([0, 10, 20, 29, 30, 100], np.int32)
    pairs = np.array(list(itertools.product(range(6), range(6))), np.int32)
    got = oracle_mod.length_ratio_pass(pairs, lengths, 0.3)
    for (i, j), g in zip(pairs, got):
        lo, hi = sorted((int(lengths[i]), int(lengths[j])))
        fail = hi > 0 and np.float32(lo) / np.float32(hi) < np.float32(0.3)
        assert g == (not fail)
