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


def _similar_setup(seed, n_db=40, n_q=5):
    from clustkit_amd.kmer_index import build_kmer_index
    from clustkit_amd.scoring import pack_sequences
    from clustkit_amd.workloads import mutate
    rng = np.random.default_rng(seed)
    roots = [rng.integers(0, 20, int(n)).astype(np.uint8) for n in rng.integers(20, 90, n_db // 4)]
    db = []
    for r in roots:
        db.append(r)
        for _ in range(3):
            x = (mutate(rng, r, float(rng.uniform(0.5, 0.9))) % 20).astype(np.uint8)
            if rng.random() < 0.3:
                x[rng.integers(0, len(x), 2)] = 20          # unknown residues
            db.append(x)
    flat, off, lens = pack_sequences(db)
    ko, ke, kf = build_kmer_index(flat, off, lens, 5)
    qs = [(mutate(rng, roots[i], 0.7) % 20).astype(np.uint8) for i in rng.choice(len(roots), n_q)]
    qs[0][3] = 22
    qf, qo, ql = pack_sequences(qs)
    return flat, off, lens, ko, ke, kf, qf, qo, ql


def _similar_brute(qf, qo, ql, flat, off, lens, ko, ke, kf, ft, T, sub):
    """Weighted counts by enumerating every one of the 20^5 codes per query
    position (no pruning): S(q, a) = sum_j sub[q_j, a_j], query residue as
    the row index as in kmer_index.py:1246-1272."""
    B = np.asarray(sub, np.int64)
    nd = len(lens)
    tid = (np.asarray(ke, np.int64) >> 32).astype(np.int64)
    out = []
    for q in range(len(ql)):
        s = qf[qo[q]:qo[q] + ql[q]].astype(np.int64)
        cnt = np.zeros(nd, np.int64)
        for p in range(len(s) - 4):
            w = s[p:p + 5]
            if (w >= 20).any():
                continue
            S = (B[w[0]][:, None, None, None, None] + B[w[1]][None, :, None, None, None]
                 + B[w[2]][None, None, :, None, None] + B[w[3]][None, None, None, :, None]
                 + B[w[4]][None, None, None, None, :]).reshape(-1)
            codes = np.nonzero(S >= T)[0]
            f = kf[codes]
            codes = codes[(f > 0) & (f <= ft)]
            for c in codes:
                np.add.at(cnt, tid[ko[c]:ko[c + 1]], S[c])
        out.append(cnt)
    return out


def _select(cnt, min_total, mc):
    ids = np.nonzero(cnt >= min_total)[0]
    if len(ids) > mc:
        ids = ids[np.lexsort((ids, -cnt[ids]))][:mc]
    return ids


@pytest.mark.parametrize("T", [9, 13, 18])
def test_similar_kmer_oracle_matches_brute_force(T, oracle_mod):
    """The oracle's pruned walk counts exactly the codes with S >= T: against
    an enumeration of all 20^5 codes per position, with the frequency filter,
    unknown residues, and truncation to max_cands (count desc, id asc)."""
    from clustkit_amd.scoring import BLOSUM62
    flat, off, lens, ko, ke, kf, qf, qo, ql = _similar_setup(7 + T)
    ft = int(np.percentile(kf[kf > 0], 90))
    for mc in (2000, 7):
        tg, cn, sc = oracle_mod.kmer_similar_batch(qf, qo, ql, ko, ke, kf, ft, len(lens), BLOSUM62,
                                                   T, min_total_hits=2, max_cands=mc)
        brute = _similar_brute(qf, qo, ql, flat, off, lens, ko, ke, kf, ft, T, BLOSUM62)
        for q, cnt in enumerate(brute):
            ids = _select(cnt, 2, mc)
            assert cn[q] == len(ids), (q, mc)
            assert np.array_equal(tg[q, :cn[q]], ids), (q, mc)
            assert np.array_equal(sc[q, :cn[q]], cnt[ids]), (q, mc)


def test_similar_kmer_oracle_closed_form_all_codes(oracle_mod):
    """With a threshold below every score every 5-mer is a neighbour, so a
    target's count is the sum over (valid query 5-mer, target 5-mer) pairs of
    their column score, restricted to target 5-mers whose frequency passes --
    computed target by target, without enumerating codes."""
    from clustkit_amd.scoring import BLOSUM62
    flat, off, lens, ko, ke, kf, qf, qo, ql = _similar_setup(3, n_db=16, n_q=3)
    ft = int(np.percentile(kf[kf > 0], 80))
    B = np.asarray(BLOSUM62, np.int64)
    tg, cn, sc = oracle_mod.kmer_similar_batch(qf, qo, ql, ko, ke, kf, ft, len(lens), BLOSUM62,
                                               -1000, min_total_hits=-10 ** 9, max_cands=2000)
    for q in range(len(ql)):
        s = qf[qo[q]:qo[q] + ql[q]].astype(np.int64)
        qw = [s[p:p + 5] for p in range(len(s) - 4) if (s[p:p + 5] < 20).all()]
        want = np.zeros(len(lens), np.int64)
        for t in range(len(lens)):
            u = flat[off[t]:off[t] + lens[t]].astype(np.int64)
            for p in range(len(u) - 4):
                w = u[p:p + 5]
                if (w >= 20).any():
                    continue
                f = kf[int(np.polyval(w, 20))]
                if not (0 < f <= ft):
                    continue
                want[t] += sum(int(B[a, w].sum()) for a in qw)
        assert cn[q] == len(lens)
        assert np.array_equal(tg[q, :cn[q]], np.arange(len(lens)))
        assert np.array_equal(sc[q, :cn[q]], want)
