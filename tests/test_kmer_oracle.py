"""Pins the k-mer scoring oracle (Phase A/B, clustkit/csrc/kmer_score.c) and the
host index build against golden vectors from the reference object and
against closed forms."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN

KMER_FILES = sorted(GOLDEN.glob("kmer_*.npz"))


def _load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", KMER_FILES, ids=lambda p: p.stem)
def test_kmer_oracle_matches_reference_golden(path, oracle_mod):
    from clustkit_amd.kmer_index import build_kmer_index
    g = _load(path)
    k = int(g["k"])
    ko, ke, kf = build_kmer_index(g["db_flat"], g["db_offsets"], g["db_lengths"], k)
    mt, md, w, mc, ta = (int(x) for x in g["params"])
    tg, cn, sc, dg = oracle_mod.kmer_score_batch(
        g["q_flat"], g["q_offsets"], g["q_lengths"], k, 20, ko, ke, kf, int(g["freq_thresh"]),
        len(g["db_lengths"]), min_total_hits=mt, min_diag_hits=md, diag_bin_width=w,
        max_cands=mc, phase_a_topk=ta)
    assert np.array_equal(cn, g["counts"])
    for q in range(len(cn)):
        n = cn[q]
        assert np.array_equal(tg[q, :n], g["targets"][q, :n]), q
        assert np.array_equal(sc[q, :n], g["scores"][q, :n]), q
        if md > 1:
            assert np.array_equal(dg[q, :n], g["diags"][q, :n]), q


def test_kmer_index_layout_closed_form():
    from clustkit_amd.kmer_index import build_kmer_index, compute_freq_threshold
    seqs = [np.array([0, 1, 2, 0, 1, 2], np.uint8), np.array([1, 2, 20, 1, 2, 0], np.uint8),
            np.array([3], np.uint8)]
    from clustkit_amd.scoring import pack_sequences
    flat, off, ln = pack_sequences(seqs)
    ko, ke, kf = build_kmer_index(flat, off, ln, 2)
    assert len(ko) == 401 and ko[-1] == len(ke) == kf.sum()
    # every (seq, pos) window without a code >= 20 appears once, under its code
    want = {}
    for s, seq in enumerate(seqs):
        for p in range(len(seq) - 1):
            if seq[p] < 20 and seq[p + 1] < 20:
                want.setdefault(int(seq[p]) * 20 + int(seq[p + 1]), []).append((s << 32) | p)
    for code in range(400):
        got = list(ke[ko[code]:ko[code + 1]])
        assert got == want.get(code, []), code
    assert compute_freq_threshold(kf, 3, 95.0) == 100
    assert compute_freq_threshold(np.zeros(5, np.int32), 7) == 7


def test_kmer_index_matches_reference_golden_db():
    """The index of a golden database reproduces the candidate sets: checked
    through the oracle test above; here the CSR invariants on that database."""
    from clustkit_amd.kmer_index import build_kmer_index
    g = _load(KMER_FILES[0])
    ko, ke, kf = build_kmer_index(g["db_flat"], g["db_offsets"], g["db_lengths"], int(g["k"]))
    assert np.all(np.diff(ko) == kf)
    tid = ke >> 32
    for code in np.flatnonzero(kf > 1)[:200]:
        seg = ke[ko[code]:ko[code + 1]]
        assert np.all(np.diff(seg) > 0)               # (seq, pos) ascending
    assert tid.max() < len(g["db_lengths"])


def test_tile_offsets_numpy_and_c_agree_and_bound_every_posting():
    """The per-(k-mer, 8192-target tile) table that steers the LDS-tiled
    Phase A/B: every posting in a tile's range has its target in the tile."""
    import ctypes
    from clustkit_amd import _lib
    from clustkit_amd.kmer_index import build_kmer_index
    from clustkit_amd.kmer_search import tile_offsets
    from clustkit_amd.workloads import FamilySpec, generate_families
    db = generate_families(FamilySpec(n_roots=2000, n_desc=9, median=120, clip_lo=20,
                                      clip_hi=300, seed=9))
    nd = len(db.lengths)
    ko, ke, kf = build_kmer_index(db.flat, db.offsets, db.lengths, 2)
    to = tile_offsets(ko, ke, nd)
    nt = (nd + 8191) // 8192
    tid = (ke >> 32)
    assert to.shape == (len(ko) - 1, nt + 1)
    for c in range(0, len(ko) - 1, 7):
        assert to[c, 0] == ko[c] and to[c, nt] == ko[c + 1]
        for j in range(nt):
            seg = tid[to[c, j]:to[c, j + 1]]
            assert np.all((seg >= j * 8192) & (seg < (j + 1) * 8192))
    if _lib.LIB_PATH.exists():
        lib = _lib.load()
        out = np.zeros_like(to)
        t32 = np.ascontiguousarray(tid, np.int32)
        assert lib.ck_kmer_tile_offsets_host(ko.ctypes.data, t32.ctypes.data, len(ko) - 1,
                                             nd, out.ctypes.data) == 0
        assert np.array_equal(out, to)


def test_cell_counts_closed_form_matches_loop():
    from clustkit_amd.cells import pair_cells
    rng = np.random.default_rng(0)
    L = rng.integers(0, 120, 60)
    pairs = rng.integers(0, 60, (500, 2))
    for band in (30, 3, 0):
        for hints in (None, rng.integers(-150, 150, 500)):
            got = pair_cells(pairs, L, band, hints)
            for p, (x, y) in enumerate(pairs):
                lx, ly = L[x], L[y]
                la, lb = min(lx, ly), max(lx, ly)
                h = 0 if hints is None else int(hints[p])
                if h > la or h < -la:
                    h = 0
                if lx > ly:
                    h = -h
                bw = lb if (band <= 0 or lb <= 50) else band
                want = sum(max(0, min(lb, i + h + bw) - max(1, i + h - bw) + 1)
                           for i in range(1, la + 1))
                assert got[p] == want
