"""GPU parity of the banded SW kernels against the reference golden vectors
and the CPU oracle.  Integer outputs must be bit-exact (scores, match
counts, identity bits, masks)."""

from __future__ import annotations

import numpy as np
import pytest

from clustkit_amd.scoring import BLOSUM62
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu


def _gpu(g_pairs, flat, offsets, lengths, band, hints=None, plain=False, **kw):
    from clustkit_amd.pairwise import batch_sw_align
    return batch_sw_align(g_pairs, flat, offsets, lengths, band, BLOSUM62, diag_hints=hints,
                          plain=plain, **kw)


def _assert_same(got, sims, scores, mask, matches=None, label=""):
    g_sims, g_scores, g_matches, g_mask = got
    bad = np.flatnonzero(g_scores != scores)
    assert bad.size == 0, f"{label}: {bad.size} score mismatches, first {bad[:5]}"
    assert np.array_equal(g_sims.view(np.uint32), np.asarray(sims, np.float32).view(np.uint32)), label
    assert np.array_equal(g_mask, np.asarray(mask, np.uint8)), label
    if matches is not None:
        assert np.array_equal(g_matches, matches), label


@pytest.mark.parametrize("plain", [False, True], ids=["batched", "plain"])
@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.stem)
def test_gpu_matches_reference_golden(path, plain):
    g = load_golden(path)
    print(f"case {g['name']} plain={plain} pairs={len(g['pairs'])}", flush=True)
    got = _gpu(g["pairs"], g["flat"], g["offsets"], g["lengths"], int(g["band"]), g["hints"],
               plain=plain)
    _assert_same(got, g["sims"], g["scores"], g["mask"], label=g["name"])


@pytest.mark.parametrize("band", [100, 16, 40, 0])
def test_gpu_matches_oracle_families(band, oracle_mod, small_family):
    fam = small_family
    rng = np.random.default_rng(band + 1)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(6000, 2)).astype(np.int32)
    pairs[:4000, 1] = (pairs[:4000, 0] // 10) * 10 + rng.integers(0, 10, 4000)
    print(f"families band={band}", flush=True)
    o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, band, BLOSUM62)
    for plain in (False, True):
        got = _gpu(pairs, fam.flat, fam.offsets, fam.lengths, band, plain=plain)
        _assert_same(got, o["sims"], o["scores"], o["mask"], o["matches"],
                     label=f"band {band} plain {plain}")


def test_gpu_hints_match_oracle(oracle_mod, small_family):
    fam = small_family
    rng = np.random.default_rng(77)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(4000, 2)).astype(np.int32)
    pairs[:3000, 1] = (pairs[:3000, 0] // 10) * 10 + rng.integers(0, 10, 3000)
    hints = rng.integers(-400, 400, len(pairs)).astype(np.int32)
    hints[::3] = rng.integers(-20, 20, len(hints[::3]))
    for band in (50, 7):
        print(f"hints band={band}", flush=True)
        o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, band, BLOSUM62, hints)
        got = _gpu(pairs, fam.flat, fam.offsets, fam.lengths, band, hints)
        _assert_same(got, o["sims"], o["scores"], o["mask"], o["matches"], label=f"hints {band}")


def test_gpu_saturation_and_long_rows_rerun(oracle_mod):
    """Pairs whose 16-bit score lane saturates, whose shorter sequence
    exceeds 8191 rows or which carry raw codes > 30 are re-run in int32."""
    from clustkit_amd.device import last_stats
    from clustkit_amd.scoring import pack_sequences
    rng = np.random.default_rng(8)
    w = np.full(3300, 18, np.uint8)                           # W x 3300: 36300
    r1 = rng.integers(0, 20, 8400).astype(np.uint8)
    r2 = r1.copy()
    r2[rng.random(8400) < 0.1] = rng.integers(0, 20, 1)
    odd = rng.integers(0, 256, 300).astype(np.uint8)
    seqs = [w, w.copy(), r1, r2, odd, odd.copy(), rng.integers(0, 20, 500).astype(np.uint8)]
    flat, off, ln = pack_sequences(seqs)
    pairs = np.array([[0, 1], [2, 3], [4, 5], [6, 6], [0, 6], [4, 6]], np.int32)
    o = oracle_mod.sw_batch(pairs, flat, off, ln, 100, BLOSUM62)
    assert o["scores"][0] == 36300 and o["scores"][1] > 32767
    got = _gpu(pairs, flat, off, ln, 100)
    _assert_same(got, o["sims"], o["scores"], o["mask"], o["matches"], label="rerun")
    st = last_stats()
    assert st["rerun"] >= 3


def test_gpu_length_ratio_and_identity_mask(oracle_mod, small_family):
    fam = small_family
    rng = np.random.default_rng(12)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(3000, 2)).astype(np.int32)
    t = np.float32(0.35)
    o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, 100, BLOSUM62)
    lr = oracle_mod.length_ratio_pass(pairs, fam.lengths, float(t))
    sims = np.where(lr, o["sims"], np.float32(0))
    scores = np.where(lr, o["scores"], 0)
    mask = (lr & (o["scores"] > 0) & (o["sims"] >= t)).astype(np.uint8)
    for plain in (False, True):
        got = _gpu(pairs, fam.flat, fam.offsets, fam.lengths, 100, plain=plain,
                   threshold=t, length_ratio=True, mask_identity=True)
        _assert_same(got, sims, scores, mask, label=f"length ratio plain={plain}")


def test_gpu_empty_and_degenerate_inputs(oracle_mod):
    from clustkit_amd.scoring import pack_sequences
    seqs = [np.zeros(0, np.uint8), np.array([5], np.uint8), np.array([5], np.uint8),
            np.array([20, 20], np.uint8), np.array([20, 20, 20, 20], np.uint8)]
    flat, off, ln = pack_sequences(seqs)
    pairs = np.array([(i, j) for i in range(5) for j in range(5)], np.int32)
    o = oracle_mod.sw_batch(pairs, flat, off, ln, 10, BLOSUM62)
    got = _gpu(pairs, flat, off, ln, 10)
    _assert_same(got, o["sims"], o["scores"], o["mask"], o["matches"], label="degenerate")
    got0 = _gpu(np.zeros((0, 2), np.int32), flat, off, ln, 10)
    assert all(len(x) == 0 for x in got0)


def test_gpu_graph_replays_match_direct_launches(oracle_mod, small_family):
    """The batched call is captured once as a hipGraph and replayed: every
    replay (with eager torch work and host synchronisations in between, and
    the outputs poisoned before it) equals the directly launched call and
    the oracle; a call with other arguments re-captures."""
    import torch

    from clustkit_amd import _lib
    from clustkit_amd.device import DeviceSequences, alloc_outputs, sw_align_device
    fam = small_family
    rng = np.random.default_rng(77)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(5000, 2)).astype(np.int32)
    pairs[:3000, 1] = (pairs[:3000, 0] // 10) * 10 + rng.integers(0, 10, 3000)
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths)
    d_pairs = torch.from_numpy(pairs).cuda()
    o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, 60, BLOSUM62)
    print("graph replay: direct", flush=True)
    ref = sw_align_device(seqs, d_pairs, 60, flags=_lib.CK_SW_NO_GRAPH)
    torch.cuda.synchronize()
    ref = {k: v.cpu().numpy() for k, v in ref.items()}
    assert np.array_equal(ref["scores"], o["scores"])
    assert np.array_equal(ref["matches"][o["scores"] > 0], o["matches"][o["scores"] > 0])
    out = alloc_outputs(len(pairs))
    for rep in range(4):
        print(f"graph replay {rep}", flush=True)
        for v in out.values():
            v.fill_(0x7F)
        x = torch.randn(1 << 20, device="cuda").sum()      # eager work between replays
        sw_align_device(seqs, d_pairs, 60, out=out)
        torch.cuda.synchronize()
        assert float(x) == float(x)
        for k in ref:
            assert np.array_equal(out[k].cpu().numpy(), ref[k]), (rep, k)
    # other arguments: fewer pairs, another band -> new capture, still exact
    sub = d_pairs[:1234].contiguous()
    o2 = oracle_mod.sw_batch(pairs[:1234], fam.flat, fam.offsets, fam.lengths, 16, BLOSUM62)
    for rep in range(2):
        got = sw_align_device(seqs, sub, 16)
        torch.cuda.synchronize()
        assert np.array_equal(got["scores"].cpu().numpy(), o2["scores"]), rep


@pytest.mark.parametrize("lanes,shift,pipe", [(8, 5, 0), (16, 5, 0), (16, 3, 0), (8, 7, 0),
                                             (16, 12, 0), (16, 5, 1), (16, 3, 1), (16, 12, 1),
                                             (8, 5, 1), (8, 3, 1), (8, 12, 1)])
def test_gpu_launch_configs_match_oracle(lanes, shift, pipe, oracle_mod, small_family):
    """Every autotuner configuration (lanes per pair, cost-bin width,
    stripe-pipelined or not) gives the oracle's results, for wide, narrow,
    tiny and full-DP bands."""
    from clustkit_amd import autotune
    fam = small_family
    rng = np.random.default_rng(lanes * 100 + shift + 7 * pipe)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(3000, 2)).astype(np.int32)
    pairs[:2000, 1] = (pairs[:2000, 0] // 10) * 10 + rng.integers(0, 10, 2000)
    hints = rng.integers(-30, 30, 3000).astype(np.int32)
    autotune.apply(lanes, shift, pipe)
    try:
        for band in (100, 16, 3, 0):
            print(f"config lanes={lanes} shift={shift} pipe={pipe} band={band}", flush=True)
            h = hints if band == 16 else None
            o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, band, BLOSUM62,
                                    hints=h)
            got = _gpu(pairs, fam.flat, fam.offsets, fam.lengths, band, h)
            _assert_same(got, o["sims"], o["scores"], o["mask"], o["matches"],
                         label=f"lanes {lanes} shift {shift} pipe {pipe} band {band}")
    finally:
        autotune.apply(0, 5, 1)


def test_gpu_autotune_picks_a_config_and_stays_exact(oracle_mod, small_family, tmp_path):
    import torch

    from clustkit_amd import autotune
    from clustkit_amd.device import DeviceSequences, sw_align_device
    fam = small_family
    rng = np.random.default_rng(5)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(4000, 2)).astype(np.int32)
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths)
    d_pairs = torch.from_numpy(pairs).cuda()
    cache = tmp_path / "tune.json"
    try:
        print("autotune", flush=True)
        res = autotune.tune(seqs, d_pairs, 24, sample=2500, repeats=1, cache_path=cache)
        assert (res["lanes"], res["bin_shift"], res["pipelined"]) in autotune.CANDIDATES
        assert not res["cached"]
        assert autotune.current() == (res["lanes"], res["bin_shift"], res["pipelined"])
        again = autotune.tune(seqs, d_pairs, 24, sample=2500, repeats=1, cache_path=cache)
        assert again["cached"] and again["lanes"] == res["lanes"]
        o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, 24, BLOSUM62)
        got = sw_align_device(seqs, d_pairs, 24)
        torch.cuda.synchronize()
        assert np.array_equal(got["scores"].cpu().numpy(), o["scores"])
    finally:
        autotune.apply(0, 5, 1)


def test_gpu_wide_score_pass_takes_saturated_pairs(oracle_mod):
    """Pairs whose scores pass 32700 are finished by the wide-score
    wavefront pass (not the one-pair-per-thread kernel) unless they reach
    65500; results stay bit-exact."""
    from clustkit_amd.device import last_paths, last_stats
    from clustkit_amd.scoring import pack_sequences
    rng = np.random.default_rng(12)
    seqs = []
    for i in range(40):
        r = rng.integers(0, 20, int(rng.integers(6500, 8000))).astype(np.uint8)
        r2 = r.copy()
        r2[rng.random(len(r)) < rng.uniform(0.0, 0.05)] = rng.integers(0, 20)
        seqs += [r, r2]
    seqs += [np.full(6000, 18, np.uint8), np.full(6000, 18, np.uint8)]   # 66000
    flat, off, ln = pack_sequences(seqs)
    pairs = np.array([[2 * i, 2 * i + 1] for i in range(41)], np.int32)
    for band in (16, 64):
        print(f"wide pass band={band}", flush=True)
        o = oracle_mod.sw_batch(pairs, flat, off, ln, band, BLOSUM62)
        assert (o["scores"][:40] > 32700).sum() >= 20 and o["scores"][40] == 66000
        got = _gpu(pairs, flat, off, ln, band)
        _assert_same(got, o["sims"], o["scores"], o["mask"], o["matches"], label=f"wide {band}")
        st, paths = last_stats(), last_paths()
        sat = int((o["scores"] >= 32700).sum())
        assert st["rerun"] == sat and paths["wide"] == sat and paths["int32"] == 1, (st, paths)


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.stem)
def test_gpu_latency_form_matches_reference_golden(path):
    """One pair per wavefront (CK_SW_LATENCY) on the reference's golden sets."""
    g = load_golden(path)
    print(f"latency case {g['name']} pairs={len(g['pairs'])}", flush=True)
    got = _gpu(g["pairs"], g["flat"], g["offsets"], g["lengths"], int(g["band"]), g["hints"],
               latency=True)
    _assert_same(got, g["sims"], g["scores"], g["mask"], label=g["name"])


def test_gpu_latency_form_matches_oracle(oracle_mod, small_family):
    """Latency form on family pairs (bands 0-100, with and without hints),
    a small batch of a few pairs, saturating and long pairs that take the
    re-run passes, and degenerate lengths."""
    from clustkit_amd.scoring import pack_sequences
    fam = small_family
    rng = np.random.default_rng(90)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(3000, 2)).astype(np.int32)
    pairs[:2000, 1] = (pairs[:2000, 0] // 10) * 10 + rng.integers(0, 10, 2000)
    hints = rng.integers(-300, 300, len(pairs)).astype(np.int32)
    for band, h in ((100, None), (16, None), (0, None), (40, hints), (7, hints)):
        print(f"latency families band={band} hints={h is not None}", flush=True)
        o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, band, BLOSUM62, h)
        got = _gpu(pairs, fam.flat, fam.offsets, fam.lengths, band, h, latency=True)
        _assert_same(got, o["sims"], o["scores"], o["mask"], o["matches"],
                     label=f"latency band {band}")
        few = pairs[:5]
        got = _gpu(few, fam.flat, fam.offsets, fam.lengths, band,
                   None if h is None else h[:5], latency=True)
        _assert_same(got, o["sims"][:5], o["scores"][:5], o["mask"][:5], o["matches"][:5],
                     label=f"latency few band {band}")
    w = np.full(3300, 18, np.uint8)
    r1 = rng.integers(0, 20, 8400).astype(np.uint8)
    r2 = r1.copy()
    r2[rng.random(8400) < 0.1] = rng.integers(0, 20, 1)
    seqs = [w, w.copy(), r1, r2, np.zeros(0, np.uint8), np.array([5], np.uint8),
            rng.integers(0, 20, 500).astype(np.uint8)]
    flat, off, ln = pack_sequences(seqs)
    pairs = np.array([[0, 1], [2, 3], [4, 5], [5, 5], [6, 6], [0, 6], [5, 6]], np.int32)
    print("latency rerun and degenerate", flush=True)
    o = oracle_mod.sw_batch(pairs, flat, off, ln, 100, BLOSUM62)
    got = _gpu(pairs, flat, off, ln, 100, latency=True)
    _assert_same(got, o["sims"], o["scores"], o["mask"], o["matches"], label="latency rerun")


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.stem)
def test_gpu_pipelined_matches_reference_golden(path):
    """The stripe-pipelined kernel (16 lanes) on the reference's golden sets."""
    from clustkit_amd import autotune
    g = load_golden(path)
    print(f"pipelined case {g['name']} pairs={len(g['pairs'])}", flush=True)
    autotune.apply(16, 5, 1)
    try:
        got = _gpu(g["pairs"], g["flat"], g["offsets"], g["lengths"], int(g["band"]),
                   g["hints"])
    finally:
        autotune.apply(0, 5, 1)
    _assert_same(got, g["sims"], g["scores"], g["mask"], label=g["name"])
