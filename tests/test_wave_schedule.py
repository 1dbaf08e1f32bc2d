"""CPU checks of the batched SW kernel's schedule (sw_wave.hip) through its
lane-exact emulation (tests/wave_emulator.py), against the oracle.

The regression pairs are the two pairs of the sw_family_b100 golden set on
which an earlier schedule (steady steps starting at step 30, so lane 15's
first cell skipped its state reset) scored too high on the GPU."""

from __future__ import annotations

import numpy as np
import pytest

from clustkit_amd.scoring import BLOSUM62
from conftest import GOLDEN, load_golden
from wave_emulator import run_wave


def _pack(items):
    flat = np.concatenate([x for it in items for x in it[:2]])
    lens = np.array([len(x) for it in items for x in it[:2]], np.int32)
    offs = np.zeros(len(lens), np.int64)
    offs[1:] = np.cumsum(lens)[:-1]
    pairs = np.arange(len(lens), dtype=np.int32).reshape(-1, 2)
    hints = np.array([it[2] for it in items], np.int32)
    return pairs, flat, offs, lens, hints


def _check(items, band, oracle_mod, S=16):
    res = run_wave(items, band, BLOSUM62, S=S)
    pairs, flat, offs, lens, hints = _pack(items)
    o = oracle_mod.sw_batch(pairs, flat, offs, lens, band, BLOSUM62, hints=hints)
    for q, (score, matches) in enumerate(res):
        assert score == o["scores"][q], (S, q, band, score, o["scores"][q])
        if score > 0:
            assert matches == o["matches"][q], (q, band)


@pytest.mark.parametrize("S", [16, 8])
def test_schedule_regression_pairs(S, oracle_mod):
    g = load_golden(GOLDEN / "sw_family_b100.npz")
    F, O, L = g["flat"], g["offsets"], g["lengths"]
    for p in (2148, 2993):
        a, b = g["pairs"][p]
        items = [(F[O[a]:O[a] + L[a]], F[O[b]:O[b] + L[b]], 0)]
        res = run_wave(items, int(g["band"]), BLOSUM62, S=S)
        assert res[0][0] == g["scores"][p]


@pytest.mark.parametrize("S,seed", [(16, 11), (16, 12), (8, 13), (8, 14)])
def test_schedule_random_waves_match_oracle(S, seed, oracle_mod):
    """Waves of four pairs with mixed lengths (1-300 x 1-400), related and
    unrelated sequences, diagonal hints, and bands from 0 (full DP) to 255:
    short windows (every-rule path), long windows (ramp/steady/ramp-down),
    band edges inside and outside the matrix, empty windows; S lanes per
    pair (16: four pairs per wavefront, 8: eight)."""
    rng = np.random.default_rng(seed)
    for _ in range(12):
        band = int(rng.choice([0, 1, 3, 16, 33, 40, 100, 255]))
        items = []
        for _q in range(64 // S):
            la, lb = int(rng.integers(1, 300)), int(rng.integers(1, 400))
            A = rng.integers(0, 20, la).astype(np.uint8)
            if rng.random() < 0.5:
                B = A[:lb].copy()
                m = rng.random(len(B)) < rng.uniform(0.05, 0.7)
                B[m] = rng.integers(0, 20, int(m.sum()))
                if rng.random() < 0.5:
                    B = np.concatenate([rng.integers(0, 20, int(rng.integers(0, 40))).astype(np.uint8), B])
            else:
                B = rng.integers(0, 20, lb).astype(np.uint8)
            h = int(rng.integers(-60, 60)) if rng.random() < 0.5 else 0
            items.append((A, B, h))
        _check(items, band, oracle_mod, S)


@pytest.mark.parametrize("seed", [21, 22])
def test_wide_layout_random_waves_match_oracle(seed, oracle_mod):
    """The wide-score key layout (17-bit scores, 2-bit preferences, zero by a
    compare) on the same kind of waves as the narrow one."""
    rng = np.random.default_rng(seed)
    for _ in range(10):
        band = int(rng.choice([0, 3, 16, 40, 100]))
        items = []
        for _q in range(4):
            la, lb = int(rng.integers(1, 250)), int(rng.integers(1, 300))
            A = rng.integers(0, 20, la).astype(np.uint8)
            B = A[:lb].copy() if rng.random() < 0.5 else rng.integers(0, 20, lb).astype(np.uint8)
            if len(B) and rng.random() < 0.5:
                m = rng.random(len(B)) < 0.3
                B[m] = rng.integers(0, 20, int(m.sum()))
            items.append((A, B, int(rng.integers(-30, 30)) if rng.random() < 0.5 else 0))
        res = run_wave(items, band, BLOSUM62, wide=True)
        pairs, flat, offs, lens, hints = _pack(items)
        o = oracle_mod.sw_batch(pairs, flat, offs, lens, band, BLOSUM62, hints=hints)
        for q, r in enumerate(res):
            assert r[0] == o["scores"][q], (band, q)
            if r[0] > 0:
                assert r[1] == o["matches"][q]


def test_wide_layout_takes_scores_the_narrow_one_hands_on(oracle_mod):
    """Scores between 32700 and 65500 are finished by the wide layout (the
    narrow one hands them on); 66000 is handed on by both."""
    rng = np.random.default_rng(3)
    r = rng.integers(0, 20, 7000).astype(np.uint8)
    r2 = r.copy()
    r2[rng.random(7000) < 0.02] = 5
    w = np.full(3300, 18, np.uint8)                 # W x 3300 = 36300
    big = np.full(6000, 18, np.uint8)               # W x 6000 = 66000
    items = [(r, r.copy(), 0), (r, r2, 0), (w, w.copy(), 0), (big, big.copy(), 0)]
    pairs, flat, offs, lens, hints = _pack(items)
    o = oracle_mod.sw_batch(pairs, flat, offs, lens, 5, BLOSUM62, hints=hints)
    assert o["scores"][2] == 36300 and o["scores"][3] == 66000
    assert all(o["scores"][:3] > 32700)
    narrow = run_wave(items, 5, BLOSUM62)
    assert narrow == [None, None, None, None]
    wide = run_wave(items, 5, BLOSUM62, wide=True)
    for q in range(3):
        assert wide[q] == (o["scores"][q], o["matches"][q])
    assert wide[3] is None


@pytest.mark.parametrize("seed", [41, 42])
def test_pipelined_schedule_matches_oracle(seed, oracle_mod):
    """The stripe-pipelined schedule (sw_pipe_kernel): lanes walk their
    stripes back to back, switching groups two steps apart; mixed lengths,
    related and random pairs, hints, bands 0-255 (windows shorter and longer
    than the 32-step transition)."""
    from wave_emulator import run_wave_pipe
    rng = np.random.default_rng(seed)
    for _ in range(4):
        band = int(rng.choice([0, 1, 3, 16, 33, 40, 100, 255]))
        items = []
        for _q in range(4):
            la, lb = int(rng.integers(1, 200)), int(rng.integers(1, 260))
            A = rng.integers(0, 20, la).astype(np.uint8)
            if rng.random() < 0.5:
                B = A[:lb].copy()
                m = rng.random(len(B)) < rng.uniform(0.05, 0.7)
                B[m] = rng.integers(0, 20, int(m.sum()))
            else:
                B = rng.integers(0, 20, lb).astype(np.uint8)
            items.append((A, B, int(rng.integers(-40, 40)) if rng.random() < 0.5 else 0))
        res = run_wave_pipe(items, band, BLOSUM62)
        pairs, flat, offs, lens, hints = _pack(items)
        o = oracle_mod.sw_batch(pairs, flat, offs, lens, band, BLOSUM62, hints=hints)
        for q, r in enumerate(res):
            assert r[0] == o["scores"][q], (band, q)
            if r[0] > 0:
                assert r[1] == o["matches"][q]


@pytest.mark.parametrize("S,band", [(16, 16), (16, 20), (16, 100), (8, 8), (8, 12)])
def test_pipelined_reduced_transitions_match_oracle(S, band, oracle_mod):
    """Groups whose previous windows all span the group and whose new
    windows exceed 2S positions take the kernel's reduced transition rules
    (checked against the full rules at every step inside the emulator):
    equal-length related pairs, so the interior stripes end at the band edge
    and the last ones right of the matrix in every segment at once; then
    mixed lengths, which leave the full rules to the last groups.  Bands 16
    (S = 16) and 8 (S = 8) give groups shorter than 2S + 7 steps, whose
    reduced transitions read and store the row buffer step by step."""
    from wave_emulator import run_wave_pipe
    rng = np.random.default_rng(band)
    NS = 64 // S
    for equal in (True, False):
        items = []
        for _q in range(NS):
            la = 180 if equal else int(rng.integers(120, 200))
            A = rng.integers(0, 20, la).astype(np.uint8)
            B = A.copy()
            m = rng.random(la) < rng.uniform(0.1, 0.5)
            B[m] = rng.integers(0, 20, int(m.sum()))
            items.append((A, B, 0))
        stats = {}
        res = run_wave_pipe(items, band, BLOSUM62, S=S, stats=stats)
        assert stats["fast_groups"] >= (stats["groups"] // 2 if equal else 1), stats
        pairs, flat, offs, lens, hints = _pack(items)
        o = oracle_mod.sw_batch(pairs, flat, offs, lens, band, BLOSUM62, hints=hints)
        for q, r in enumerate(res):
            assert r == (o["scores"][q], o["matches"][q]), (S, band, equal, q)


@pytest.mark.parametrize("S", [8, 16, 64])
def test_pipelined_empty_leading_stripes(S, oracle_mod):
    """A hint that puts the first stripes entirely left of column 1: those
    stripes write nothing into the row buffer, so the first stripe with
    cells reads H = 0 of the row above it everywhere (with that row's
    offset).  Planted similarity in the corner the band reaches."""
    from wave_emulator import run_wave_pipe
    rng = np.random.default_rng(183)
    i0 = 177 if S < 64 else 129      # the first row of a stripe; column 1 enters there
    items = []
    for _k in range(64 // S):
        A = rng.integers(0, 20, 251).astype(np.uint8)
        B = rng.integers(0, 20, 253).astype(np.uint8)
        A[i0 - 1:i0 + 59] = B[0:60]   # the alignment starts at (i0, 1)
        items.append((A, B, -6 - i0))
    res = run_wave_pipe(items, 7, BLOSUM62, S=S)
    pairs, flat, offs, lens, hints = _pack(items)
    o = oracle_mod.sw_batch(pairs, flat, offs, lens, 7, BLOSUM62, hints=hints)
    assert all(o["scores"] > 100)
    for q, r in enumerate(res):
        assert r == (o["scores"][q], o["matches"][q]), (S, q)


def test_pipelined_row_offset_limit(oracle_mod):
    """Narrow keys hold score + row: a pair whose score plus row reaches the
    narrow limit is handed on even though its score alone is below it (it
    goes to the wide pass), and one below the limit is finished."""
    from wave_emulator import run_wave_pipe
    w_hi = np.full(2800, 18, np.uint8)        # score 30,800 at row 2,800
    w_lo = np.full(2000, 18, np.uint8)        # score 22,000 at row 2,000
    items = [(w_hi, w_hi.copy(), 0), (w_lo, w_lo.copy(), 0)]
    pairs, flat, offs, lens, hints = _pack(items)
    o = oracle_mod.sw_batch(pairs, flat, offs, lens, 3, BLOSUM62, hints=hints)
    assert o["scores"][0] == 30800 and o["scores"][0] < 32700
    res = run_wave_pipe(items, 3, BLOSUM62)
    assert res[0] is None
    assert res[1] == (o["scores"][1], o["matches"][1])


@pytest.mark.parametrize("seed", [71, 72])
def test_pipelined_8_lane_schedule_matches_oracle(seed, oracle_mod):
    """The stripe-pipelined schedule with 8-lane segments (eight pairs per
    wavefront, 16-step transitions): narrow and wide bands, windows shorter
    and longer than the transition, hints."""
    from wave_emulator import run_wave_pipe
    rng = np.random.default_rng(seed)
    for _ in range(4):
        band = int(rng.choice([0, 1, 3, 8, 16, 20, 32, 40]))
        items = []
        for _q in range(8):
            la, lb = int(rng.integers(1, 150)), int(rng.integers(1, 200))
            A = rng.integers(0, 20, la).astype(np.uint8)
            if rng.random() < 0.5:
                B = A[:lb].copy()
                m = rng.random(len(B)) < rng.uniform(0.05, 0.7)
                B[m] = rng.integers(0, 20, int(m.sum()))
            else:
                B = rng.integers(0, 20, lb).astype(np.uint8)
            items.append((A, B, int(rng.integers(-40, 40)) if rng.random() < 0.5 else 0))
        res = run_wave_pipe(items, band, BLOSUM62, S=8)
        pairs, flat, offs, lens, hints = _pack(items)
        o = oracle_mod.sw_batch(pairs, flat, offs, lens, band, BLOSUM62, hints=hints)
        for q, r in enumerate(res):
            assert r[0] == o["scores"][q], (band, q)
            if r[0] > 0:
                assert r[1] == o["matches"][q]


@pytest.mark.parametrize("seed", [61, 62])
def test_pipelined_64_lane_chain_matches_oracle(seed, oracle_mod):
    """The one-pair-per-wavefront form (sw_pipe_kernel<64>): the 64 lanes one
    chain, 128-step transitions; narrow keys, windows shorter and longer
    than 128, hints, bands 0-255."""
    from wave_emulator import run_wave_pipe
    rng = np.random.default_rng(seed)
    for _ in range(5):
        band = int(rng.choice([0, 1, 5, 16, 40, 100, 255]))
        la, lb = int(rng.integers(1, 300)), int(rng.integers(1, 360))
        A = rng.integers(0, 20, la).astype(np.uint8)
        if rng.random() < 0.6:
            B = A[:lb].copy()
            m = rng.random(len(B)) < rng.uniform(0.05, 0.6)
            B[m] = rng.integers(0, 20, int(m.sum()))
        else:
            B = rng.integers(0, 20, lb).astype(np.uint8)
        h = int(rng.integers(-50, 50)) if rng.random() < 0.5 else 0
        got = run_wave_pipe([(A, B, h)], band, BLOSUM62, S=64)[0]
        pairs, flat, offs, lens, hints = _pack([(A, B, h)])
        o = oracle_mod.sw_batch(pairs, flat, offs, lens, band, BLOSUM62, hints=hints)
        assert got[0] == o["scores"][0], (band, la, lb, h)
        if got[0] > 0:
            assert got[1] == o["matches"][0]


def test_pipelined_64_lane_chain_wide_keys(oracle_mod):
    """The wide-score re-run pass (sw_pipe_kernel<64, wide keys>): random
    pairs, and scores between 32700 and 65500 finished, 66000 handed on."""
    from wave_emulator import run_wave_pipe
    rng = np.random.default_rng(51)
    items = []
    for _ in range(5):
        band = int(rng.choice([0, 5, 16, 100]))
        la, lb = int(rng.integers(1, 220)), int(rng.integers(1, 260))
        A = rng.integers(0, 20, la).astype(np.uint8)
        B = A[:lb].copy() if rng.random() < 0.5 else rng.integers(0, 20, lb).astype(np.uint8)
        items.append((A, B, int(rng.integers(-30, 30)) if rng.random() < 0.5 else 0, band))
    r = rng.integers(0, 20, 7000).astype(np.uint8)
    r2 = r.copy()
    r2[rng.random(7000) < 0.02] = 5
    w = np.full(3300, 18, np.uint8)
    big = np.full(6000, 18, np.uint8)
    items += [(r, r2, 0, 5), (w, w.copy(), 0, 5), (big, big.copy(), 0, 5)]
    for A, B, h, band in items:
        got = run_wave_pipe([(A, B, h)], band, BLOSUM62, S=64, wide=True)[0]
        pairs, flat, offs, lens, hints = _pack([(A, B, h)])
        o = oracle_mod.sw_batch(pairs, flat, offs, lens, band, BLOSUM62, hints=hints)
        if o["scores"][0] >= 65500:
            assert got is None
            continue
        assert got is not None and got[0] == o["scores"][0], (band, len(A), len(B))
        if got[0] > 0:
            assert got[1] == o["matches"][0]


# ---------------------------------------------------------------------------
# Global alignment: the 16-lane stripe schedule of nw_wave_kernel
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("S", [16, 8])
@pytest.mark.parametrize("scored", [False, True])
def test_nw_wave_schedule_matches_oracle(scored, S, oracle_mod):
    """Lane-by-lane emulation of nw_wave_kernel and of the packed-key
    nw_pack_kernel (DPP hand-over, row buffer, staged codes, early stop,
    final cell; int32 wrap-around for the keys), with every LDS index checked
    against the buffers the kernel allocates, against the oracle."""
    from wave_emulator import nw_pack_pair, nw_wave_pair
    from clustkit_amd.scoring import BLOSUM62
    from test_nw_oracle import _family
    for case in range(24):
        rng = np.random.default_rng(100 + case + 50 * scored)
        alpha = 20 if scored else 4
        a, b = _family(rng, int(rng.integers(5, 120)), alpha, rng.uniform(0.05, 0.5),
                       rng.uniform(0.0, 0.2))
        if rng.random() < 0.3:
            a, b = b, a
        bwp = int(rng.choice([0, 5, 10, 20, 40, 100]))
        thr = float(rng.choice([0.0, 0.3, 0.6]))
        flat = np.concatenate([a, b]).astype(np.uint8)
        off = np.array([0, len(a)], np.int64)
        lens = np.array([len(a), len(b)], np.int32)
        sub = BLOSUM62 if scored else None
        r = oracle_mod.nw_batch(np.array([[0, 1]]), flat, off, lens, thr, bwp, sub)
        la, lb = sorted((len(a), len(b)))
        ab = min(bwp, max(10, lb - la + 10))
        W = 2 * (lb if (ab <= 0 or lb <= 50) else ab) + 1
        cap = max(W, 63) if S == 8 else 479
        for emu in (nw_wave_pair, nw_pack_pair):
            got = emu(a, b, bwp, thr, sub, S=S, w_cap=cap)
            assert got[0] == r[0][0], (case, emu.__name__)
            assert bool(got[2]) == bool(r[-1][0]), (case, emu.__name__)
            if scored:
                assert got[1] == r[1][0], (case, emu.__name__)


@pytest.mark.parametrize("S", [16, 8])
@pytest.mark.parametrize("scored", [False, True])
def test_nw_pipe_schedule_matches_oracle(scored, S, oracle_mod):
    """Lane-by-lane emulation of the stripe-pipelined packed-key NW kernel
    (groups of L >= max(2S + 1, W) steps, per-lane switches in the
    transition, early stop as an integer bound checked at the switch, full
    and reduced transitions, rule-free steady range) against the oracle,
    including the wavefront-level variants a segment meets beside others:
    longer groups, more groups, a narrower steady range, full transitions
    forced, no early exit."""
    from wave_emulator import nw_pipe_pair
    from clustkit_amd.scoring import BLOSUM62
    from test_nw_oracle import _family
    for case in range(40):
        rng = np.random.default_rng(300 + case + 70 * scored + 140 * (S == 8))
        alpha = 20 if scored else 4
        a, b = _family(rng, int(rng.integers(1, 260)), alpha, rng.uniform(0.05, 0.5),
                       rng.uniform(0.0, 0.2))
        if rng.random() < 0.3:
            a, b = b, a
        bwp = int(rng.choice([0, 3, 10, 20, 40, 100]))
        thr = float(rng.choice([0.0, 0.3, 0.6, 0.9]))
        flat = np.concatenate([a, b]).astype(np.uint8)
        off = np.array([0, len(a)], np.int64)
        lens = np.array([len(a), len(b)], np.int32)
        sub = BLOSUM62 if scored else None
        r = oracle_mod.nw_batch(np.array([[0, 1]]), flat, off, lens, thr, bwp, sub)
        la, lb = sorted((len(a), len(b)))
        ab = min(bwp, max(10, lb - la + 10))
        W = 2 * (lb if (ab <= 0 or lb <= 50) else ab) + 1
        cap = max(W, 63) if S == 8 else max(W, 479)
        kw = dict(L_extra=min(int(rng.integers(0, 5)), cap - max(2 * S + 1, W)),
                  g_extra=int(rng.integers(0, 3)), wave_break=bool(rng.random() < 0.5),
                  t_hi_other=int(rng.integers(2 * S, 2 * S + 40)) if rng.random() < 0.3 else None,
                  t_lo_other=int(rng.integers(2 * S, 2 * S + 40)) if rng.random() < 0.3 else None,
                  general_other=bool(rng.random() < 0.2))
        got = nw_pipe_pair(a, b, bwp, thr, sub, S=S, w_cap=cap, **kw)
        assert got[0] == r[0][0], case
        assert bool(got[2]) == bool(r[-1][0]), case
        if scored:
            assert got[1] == r[1][0], case
