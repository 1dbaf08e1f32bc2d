"""GPU parity of the banded global-alignment (NW identity) kernels against
the CPU oracle (oracle/nw_oracle.c, pinned in tests/test_nw_oracle.py).
Identity bits, scores and masks must be bit-exact through every kernel: the
packed-key kernels (8 and 16 lanes per pair; stripe-pipelined, the default,
and one stripe at a time), the int32 16-lane kernel for pairs outside the
16-bit range, and the one-pair-per-thread kernel."""


from __future__ import annotations

import numpy as np
import pytest

from clustkit_amd.scoring import BLOSUM62

pytestmark = pytest.mark.gpu

# kernel selections: default (stripe-pipelined packed keys), packed keys one
# stripe at a time, one pair per thread
KERNELS = (dict(), dict(stripe_kernel=True), dict(plain_kernel=True))


def _flat(seqs):
    lens = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(len(seqs), np.int64)
    off[1:] = np.cumsum(lens)[:-1]
    flat = np.concatenate(seqs).astype(np.uint8) if len(seqs) else np.zeros(0, np.uint8)
    return flat, off, lens


def _same(got, want, label):
    g_sims, g_scores, g_mask = got
    w_sims, w_scores, w_mask = want
    assert np.array_equal(g_sims.view(np.uint32), np.asarray(w_sims, np.float32).view(np.uint32)), \
        f"{label}: identity mismatches at {np.flatnonzero(g_sims != w_sims)[:5]}"
    if w_scores is not None:
        bad = np.flatnonzero(g_scores != w_scores)
        assert bad.size == 0, f"{label}: {bad.size} score mismatches, first {bad[:5]}"
    assert np.array_equal(g_mask, np.asarray(w_mask, bool)), label


def _oracle(oracle_mod, pairs, flat, off, lens, thr, band, scored):
    if scored:
        s, sc, m = oracle_mod.nw_batch(pairs, flat, off, lens, thr, band, BLOSUM62)
        return s, sc, m
    s, m = oracle_mod.nw_batch(pairs, flat, off, lens, thr, band)
    return s, None, m


@pytest.mark.parametrize("scored", [True, False], ids=["scored", "plain"])
@pytest.mark.parametrize("band", [100, 10, 30, 0, 239, 300])
def test_gpu_nw_families(band, scored, oracle_mod, small_family):
    from clustkit_amd.pairwise import batch_nw_align
    fam = small_family
    rng = np.random.default_rng(band + 7 * scored)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(5000, 2)).astype(np.int32)
    pairs[:3500, 1] = (pairs[:3500, 0] // 10) * 10 + rng.integers(0, 10, 3500)
    for thr in (0.3, 0.0, 0.8):
        print(f"nw families band={band} scored={scored} t={thr}", flush=True)
        want = _oracle(oracle_mod, pairs, fam.flat, fam.offsets, fam.lengths, thr, band, scored)
        for kk in KERNELS:
            got = batch_nw_align(pairs, fam.flat, fam.offsets, fam.lengths, thr, band,
                                 BLOSUM62 if scored else None, **kk)
            if not scored:
                got = (got[0], None, got[2])
            _same(got, want, f"band {band} t {thr} {kk}")


def test_gpu_nw_edge_cases(oracle_mod):
    """Empty sequences, lengths around the 50-residue full-matrix cut-off,
    unknown and high byte codes, equal lengths in both orders, length ratio
    and length difference at their limits."""
    from clustkit_amd.pairwise import batch_nw_align
    rng = np.random.default_rng(5)
    seqs = []
    for n in (0, 1, 2, 5, 49, 50, 51, 52, 60, 100, 101, 130, 400, 700):
        seqs.append(rng.integers(0, 20, n).astype(np.uint8))
    base = rng.integers(0, 20, 120).astype(np.uint8)
    seqs.append(base)
    seqs.append(np.where(rng.random(120) < 0.1, 20, base).astype(np.uint8))   # unknowns
    seqs.append(np.where(rng.random(120) < 0.1, 255, base).astype(np.uint8))  # raw 255
    seqs.append(base[::-1].copy())
    flat, off, lens = _flat(seqs)
    n = len(seqs)
    pairs = np.array([(i, j) for i in range(n) for j in range(n)], np.int32)
    for scored in (True, False):
        for band in (0, 1, 10, 45, 100, 239):
            for thr in (0.0, 0.4, 0.95):
                want = _oracle(oracle_mod, pairs, flat, off, lens, thr, band, scored)
                for kk in KERNELS:
                    got = batch_nw_align(pairs, flat, off, lens, thr, band,
                                         BLOSUM62 if scored else None, **kk)
                    if not scored:
                        got = (got[0], None, got[2])
                    _same(got, want, f"edge band {band} t {thr} scored {scored} {kk}")


def test_gpu_nw_nucleotide_reference_tests():
    """tests/test_pairwise.py:71-142 of the reference through the drop-in
    dispatcher (use_sw=False, band_width=0)."""
    from clustkit_amd.pairwise import compute_pairwise_alignment

    def enc(s):
        return np.array(["ACGT".index(c) if c in "ACGT" else 4 for c in s], np.uint8)

    def run(a, b):
        ea, eb = enc(a), enc(b)
        n = max(len(ea), len(eb))
        seqs = np.full((2, n), 4, np.uint8)
        seqs[0, :len(ea)] = ea
        seqs[1, :len(eb)] = eb
        lens = np.array([len(ea), len(eb)], np.int32)
        return compute_pairwise_alignment(np.array([[0, 1]], np.int32), seqs, lens,
                                          threshold=0.5, band_width=0)

    seq = "ATGCTAGCTAGCATGCTAGCTAGCATGCTAGCTAGCATGCTAGCTAGC"
    f, ids = run(seq, seq)
    assert len(f) == 1 and ids[0] == np.float32(1.0)
    f, ids = run(seq, "ATGCTCGCTAGCATGCTTGCTAGCATGCTAGCTAGCATGCTAGCTTGC")
    assert len(f) == 1 and 0.8 < ids[0] < 1.0
    f, ids = run("A" * 40, "C" * 40)
    assert len(f) == 0


def test_gpu_nw_inner_function(oracle_mod):
    """_nw_identity / _nw_identity_and_score on single oriented pairs (the
    band as given, no length tests) against the oracle's inner function."""
    from clustkit_amd.pairwise import _nw_identity, _nw_identity_and_score
    rng = np.random.default_rng(9)
    for case in range(40):
        la = int(rng.integers(0, 160))
        lb = int(rng.integers(la, la + 80))
        a = rng.integers(0, 20, la).astype(np.uint8)
        b = rng.integers(0, 20, lb).astype(np.uint8)
        if case % 3 == 0 and la:
            b[:la] = np.where(rng.random(la) < 0.8, a, b[:la])
        band = int(rng.choice([0, 3, 10, 25, 60]))
        thr = float(rng.choice([0.0, 0.2, 0.5]))
        want = oracle_mod.nw_pair(a, b, band, thr)
        assert _nw_identity(a, la, b, lb, band, thr) == want[0], case
        want_s = oracle_mod.nw_pair(a, b, band, thr, BLOSUM62)
        got_s = _nw_identity_and_score(a, la, b, lb, band, thr, BLOSUM62)
        assert (got_s[0], int(got_s[1])) == (want_s[0], want_s[1]), case


def test_gpu_nw_dispatcher_and_refinement(oracle_mod, small_family):
    """compute_pairwise_alignment: use_sw=False (global identity) and
    refine_global_margin > 0 (borderline SW hits re-scored, pairwise.py:
    1004-1044), against compositions of the oracles."""
    from clustkit_amd.pairwise import compute_pairwise_alignment
    fam = small_family
    rng = np.random.default_rng(21)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(3000, 2)).astype(np.int32)
    pairs[:2000, 1] = (pairs[:2000, 0] // 10) * 10 + rng.integers(0, 10, 2000)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    t, bw = 0.4, 60
    order = np.argsort(np.minimum(pairs[:, 0], pairs[:, 1]), kind="mergesort")
    sp = pairs[order]

    got_p, got_s = compute_pairwise_alignment(pairs, None, fam.lengths, t, band_width=bw,
                                              flat_sequences=fam.flat, offsets=fam.offsets)
    s, m = oracle_mod.nw_batch(sp, fam.flat, fam.offsets, fam.lengths, t, bw)
    assert np.array_equal(got_p, sp[m]) and np.array_equal(got_s, s[m])

    margin = 0.15
    for use_c in (True, False):
        got_p, got_s = compute_pairwise_alignment(
            pairs, None, fam.lengths, t, band_width=bw, flat_sequences=fam.flat,
            offsets=fam.offsets, use_sw=True, use_c_sw=use_c, refine_global_margin=margin)
        o = oracle_mod.sw_batch(sp, fam.flat, fam.offsets, fam.lengths, bw, BLOSUM62)
        keep = o["sims"] >= np.float32(t)
        if not use_c:
            keep &= o["scores"] > 0
        kp, ks = sp[keep], o["sims"][keep]
        ref = ks < (t + margin)
        assert ref.any()
        ns, nm = oracle_mod.nw_batch(kp[ref], fam.flat, fam.offsets, fam.lengths,
                                     np.float32(t), bw)
        ks = ks.copy()
        ks[ref] = ns
        fm = np.ones(len(kp), bool)
        fm[ref] = nm
        assert np.array_equal(got_p, kp[fm]), use_c
        assert np.array_equal(got_s.view(np.uint32), ks[fm].view(np.uint32)), use_c


def test_gpu_nw_long_sequences(oracle_mod):
    """Sequences of 2,800-4,200 residues: scored pairs beyond 2,909 rows leave
    the 16-bit key range (11 x rows > 32,000) and run the int32 kernel, the
    others and all plain pairs the packed-key kernels."""
    from clustkit_amd.pairwise import batch_nw_align
    from clustkit_amd.workloads import mutate
    rng = np.random.default_rng(31)
    seqs = []
    for n in (2800, 2950, 3300, 4200):
        root = rng.integers(0, 20, n).astype(np.uint8)
        seqs.append(root)
        for ident in (0.9, 0.6):
            seqs.append(mutate(rng, root, ident).astype(np.uint8))
    flat, off, lens = _flat(seqs)
    n = len(seqs)
    pairs = np.array([(i, j) for i in range(n) for j in range(n) if i != j], np.int32)
    for scored in (True, False):
        for band in (100, 30):
            print(f"nw long scored={scored} band={band}", flush=True)
            want = _oracle(oracle_mod, pairs, flat, off, lens, 0.3, band, scored)
            got = batch_nw_align(pairs, flat, off, lens, 0.3, band,
                                 BLOSUM62 if scored else None)
            if not scored:
                got = (got[0], None, got[2])
            _same(got, want, f"long band {band} scored {scored}")


def test_gpu_nw_device_entry_matches_host_entry(oracle_mod, small_family):
    """ck_nw_align_device on HBM tensors (the form bench.py and the search
    use), on the default and on a side stream, equals the host-pointer form
    and the oracle."""
    import torch

    from clustkit_amd.device import DeviceSequences, nw_align_device
    fam = small_family
    rng = np.random.default_rng(44)
    n = len(fam.lengths)
    pairs = rng.integers(0, n, size=(4000, 2)).astype(np.int32)
    pairs[:3000, 1] = (pairs[:3000, 0] // 10) * 10 + rng.integers(0, 10, 3000)
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths, torch.device("cuda:0"))
    d_pairs = torch.from_numpy(pairs).cuda()
    side = torch.cuda.Stream()
    for scored in (True, False):
        want = _oracle(oracle_mod, pairs, fam.flat, fam.offsets, fam.lengths, 0.3, 80, scored)
        for stream in (None, side):
            out = nw_align_device(seqs, d_pairs, 80, BLOSUM62 if scored else None,
                                  threshold=0.3, stream=stream)
            torch.cuda.synchronize()
            got = (out["sims"].cpu().numpy(), out["scores"].cpu().numpy() if scored else None,
                   out["mask"].cpu().numpy().astype(bool))
            _same(got, want, f"device entry scored={scored} side={stream is not None}")
