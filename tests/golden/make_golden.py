"""Write the SW golden vectors from the reference kernel.

Runs in the build container only: it drives oracle/_ref/sw_align.so, the
reference's clustkit/csrc/sw_align.c compiled from its own sources by
``make -C oracle ref``, through ctypes exactly as clustkit/pairwise.py:69-108
does, and stores inputs + outputs as small .npz fixtures next to this script.
The fixture sequences of the reference's test suite (tests/conftest.py:16-92)
are read from that file as text (regex over the string constants, nothing is
executed).

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import re
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from clustkit_amd.scoring import BLOSUM62, encode_sequence, pack_sequences  # noqa: E402
from clustkit_amd.workloads import FamilySpec, generate_families  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent
REF_CONFTEST = Path("/root/reference/tests/conftest.py")


def _fixture_sequences():
    """Protein and nucleotide fixtures of the reference test suite."""
    txt = REF_CONFTEST.read_text()
    consts = dict(re.findall(r'^([A-Z_]+_(?:BASE|SEQ))\s*=\s*"([A-Z]+)"', txt, re.M))
    a, b, c = consts["CLUSTER_A_BASE"], consts["CLUSTER_B_BASE"], consts["CLUSTER_C_SEQ"]
    prot = [a, a[:10] + "XX" + a[12:], a[:40] + "XX" + a[42:], b, b[:20] + "XX" + b[22:], c]
    nts = [consts[k] for k in ("NT_CLUSTER_A_BASE", "NT_CLUSTER_B_BASE", "NT_CLUSTER_C_BASE")]
    return prot, nts


def _run(name, seqs_or_packed, pairs, band, hints=None, note=""):
    if isinstance(seqs_or_packed, tuple):
        flat, offsets, lengths = seqs_or_packed
    else:
        flat, offsets, lengths = pack_sequences(seqs_or_packed)
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    sims, scores, mask = oracle.ref_sw_batch(pairs, flat, offsets, lengths, band, BLOSUM62,
                                             hints=hints, n_threads=8)
    np.savez_compressed(
        OUT / f"sw_{name}.npz", flat=flat, offsets=offsets, lengths=lengths, pairs=pairs,
        band=np.int32(band), hints=(np.zeros(0, np.int32) if hints is None else
                                    np.asarray(hints, np.int32)),
        has_hints=np.bool_(hints is not None), sims=sims, scores=scores, mask=mask,
        note=np.array(note))
    print(f"{name}: {len(pairs)} pairs, band {band}, max score {scores.max()}")


def main():
    if not oracle.ref_available():
        oracle.build(ref=True)
    rng = np.random.default_rng(20261203)

    prot, nts = _fixture_sequences()
    enc = [encode_sequence(s, "protein") for s in prot]
    n = len(enc)
    allp = np.array([(i, j) for i in range(n) for j in range(n)], np.int32)
    for band in (0, 3, 10, 50, 100):
        _run(f"fixture_protein_b{band}", enc, allp, band,
             note="reference tests/conftest.py protein fixtures, all ordered pairs")
    nenc = [encode_sequence(s, "nucleotide") for s in nts]
    nenc += [np.where(rng.random(len(s)) < 0.1, rng.integers(0, 4, len(s)), s).astype(np.uint8)
             for s in nenc]
    m = len(nenc)
    _run("fixture_nucleotide_b100",
         nenc, np.array([(i, j) for i in range(m) for j in range(m)], np.int32), 100,
         note="reference nucleotide fixtures + 10% mutants, codes 0..4 scored by BLOSUM62 rows")

    fam = generate_families(FamilySpec(n_roots=24, n_desc=9, median=220, clip_lo=20,
                                       clip_hi=600, id_lo=0.3, id_hi=0.95, seed=7))
    N = len(fam.lengths)
    fp = np.stack([rng.integers(0, N, 3000), np.zeros(3000, np.int64)], 1)
    fp[:2000, 1] = (fp[:2000, 0] // 10) * 10 + rng.integers(0, 10, 2000)
    fp[2000:, 1] = rng.integers(0, N, 1000)
    packed = (fam.flat, fam.offsets, fam.lengths)
    for band in (100, 16, 0, 1):
        _run(f"family_b{band}", packed, fp, band, note="synthetic families, 2/3 related pairs")
    hints = rng.integers(-250, 250, len(fp)).astype(np.int32)
    hints[::7] = 0
    _run("family_b30_hints", packed, fp, 30, hints=hints, note="random diagonal hints")
    _run("family_b5_hints", packed, fp[:1500], 5, hints=hints[:1500], note="narrow band + hints")

    # Edge cases: empty sequences, length 1, unknown and out-of-alphabet codes,
    # all-mismatch, identical long sequences (16-bit saturation), a shorter
    # sequence above 8191 residues, near-threshold lengths around 50.
    w = np.full(3200, 18, np.uint8)                      # 3200 x W: score 35200
    big = rng.integers(0, 20, 8300).astype(np.uint8)     # shorter side > 8191
    big2 = big.copy()
    big2[rng.random(8300) < 0.2] = rng.integers(0, 20, 1)
    seqs = [np.zeros(0, np.uint8), np.array([0], np.uint8), np.array([20, 20, 20], np.uint8),
            np.array([200, 7, 255, 7], np.uint8), np.array([255, 7, 200, 7], np.uint8),
            np.full(40, 1, np.uint8), np.full(40, 18, np.uint8), w, w.copy(),
            rng.integers(0, 21, 50).astype(np.uint8), rng.integers(0, 21, 51).astype(np.uint8),
            big, big2, rng.integers(0, 20, 700).astype(np.uint8)]
    k = len(seqs)
    ep = np.array([(i, j) for i in range(k) for j in range(k)
                   if not (i >= 11 and j >= 11 and (i, j) not in ((11, 12), (12, 11)))], np.int32)
    _run("edge_b100", seqs, ep, 100, note="empty/short/unknown/saturating/long edge cases")
    _run("edge_b0", [s[:400] for s in seqs], ep, 0, note="edge cases truncated to 400, full DP")

    _kmer_golden(rng)


def _kmer_golden(rng):
    """Phase A/B k-mer scoring vectors from the reference kmer_score.c."""
    from clustkit_amd.kmer_index import build_kmer_index, compute_freq_threshold
    db = generate_families(FamilySpec(n_roots=60, n_desc=9, median=250, clip_lo=20,
                                      clip_hi=900, id_lo=0.3, id_hi=0.95, seed=21))
    # queries: planted descendants of database roots, unrelated proteins, a
    # query with unknown residues and a query shorter than k
    qseqs = []
    from clustkit_amd.workloads import mutate
    for r in range(0, 60, 3):
        s = db.flat[db.offsets[r * 10]:db.offsets[r * 10] + db.lengths[r * 10]]
        qseqs.append(mutate(rng, s, float(rng.uniform(0.3, 0.9))))
    qseqs += [rng.integers(0, 20, int(n)).astype(np.uint8) for n in (60, 300, 900)]
    unk = qseqs[0].copy()
    unk[::7] = 20
    qseqs += [unk, np.array([3, 4], np.uint8), np.zeros(0, np.uint8)]
    qf, qo, ql = pack_sequences(qseqs)
    for k in (3, 5):
        ko, ke, kf = build_kmer_index(db.flat, db.offsets, db.lengths, k)
        nd = len(db.lengths)
        ft = compute_freq_threshold(kf, nd, 95.0)
        for tag, kw in (("default", {}), ("topk", dict(phase_a_topk=40, max_cands=16)),
                        ("phase_a", dict(min_diag_hits=1, max_cands=64))):
            tg, cn, sc, dg = oracle.ref_kmer_score_batch(qf, qo, ql, k, 20, ko, ke, kf, ft, nd,
                                                         n_threads=8, **kw)
            mc = tg.shape[1]
            valid = np.arange(mc)[None, :] < cn[:, None]
            np.savez_compressed(
                OUT / f"kmer_k{k}_{tag}.npz", db_flat=db.flat, db_offsets=db.offsets,
                db_lengths=db.lengths, q_flat=qf, q_offsets=qo, q_lengths=ql, k=np.int32(k),
                freq_thresh=np.int32(ft), params=np.array(
                    [kw.get("min_total_hits", 2), kw.get("min_diag_hits", 2),
                     kw.get("diag_bin_width", 10), kw.get("max_cands", 2000),
                     kw.get("phase_a_topk", 10000)], np.int32),
                targets=np.where(valid, tg, -1), counts=cn, scores=np.where(valid, sc, 0),
                diags=np.where(valid, dg, 0) if kw.get("min_diag_hits", 2) > 1 else
                np.zeros_like(dg))
            print(f"kmer k={k} {tag}: {int(cn.sum())} candidates over {len(ql)} queries")


if __name__ == "__main__":
    main()
