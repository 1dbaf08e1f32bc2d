"""Synthetic protein families for the alignment-stage workloads.

The reference ships no benchmark data, so workloads are generated:

* root residues i.i.d. from the BLOSUM62 background frequencies;
* root lengths lognormal (median m, sigma) clipped to [lo, hi], optionally
  mixed with a Pareto tail;
* each root has ``n_desc`` descendants at a target identity drawn
  uniformly from [id_lo, id_hi]: every site is substituted with probability
  1 - identity, the new residue drawn from P(b | a) proportional to
  p_b * 2**(S_ab / 2), b != a (the conditional target frequencies implied by
  BLOSUM62's half-bit scores); indels occur at 0.02 per site, half
  insertions (background residues) and half deletions, each with a geometric
  length of mean 3;
* sequences are integer-encoded (0..19) and packed flat.

Named workloads (sizes as used by the benchmarks and tests):

``small``   2,000 proteins  (200 roots x 10), median 300, clip 50-1000, id 0.3-0.9
``mixed2m`` 50,000 proteins (5,000 roots x 10), median 350, clip 30-2000, id 0.3-0.95
``heavy``   100,000 proteins, 90% lognormal median 250 + 10% Pareto(1.5, 1000..8000)
``fast200k`` 200,000 proteins (20,000 roots x 10), median 300, clip 30-3000, id 0.25-0.95
``serve1m`` 1,000,000 proteins (100,000 roots x 10), median 350, clip 30-5000, id 0.3-0.9
            (search-serving database; generated with one random stream per root,
            in parallel worker processes)
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .scoring import BLOSUM62, BLOSUM62_BACKGROUND, pack_sequences

_BG = BLOSUM62_BACKGROUND / BLOSUM62_BACKGROUND.sum()
_BG_CDF = np.cumsum(_BG)


def _conditional_cdf() -> np.ndarray:
    w = _BG[None, :] * np.power(2.0, BLOSUM62.astype(np.float64) / 2.0)
    np.fill_diagonal(w, 0.0)
    w /= w.sum(axis=1, keepdims=True)
    return np.cumsum(w, axis=1)


_SUB_CDF = _conditional_cdf()


@dataclass
class FamilySpec:
    n_roots: int
    n_desc: int = 9
    median: float = 300.0
    sigma: float = 0.5
    clip_lo: int = 50
    clip_hi: int = 1000
    id_lo: float = 0.3
    id_hi: float = 0.9
    indel_rate: float = 0.02
    tail_frac: float = 0.0
    tail_alpha: float = 1.5
    tail_min: int = 1000
    tail_max: int = 8000
    seed: int = 0


@dataclass
class Family:
    flat: np.ndarray
    offsets: np.ndarray
    lengths: np.ndarray
    root_of: np.ndarray
    identity: np.ndarray
    spec: FamilySpec = field(repr=False, default=None)


def _draw(rng, cdf, n):
    return np.searchsorted(cdf, rng.random(n) * cdf[-1], side="right").clip(0, 19).astype(np.uint8)


def _root_lengths(rng, spec: FamilySpec, n: int) -> np.ndarray:
    lens = rng.lognormal(np.log(spec.median), spec.sigma, size=n)
    lens = np.clip(np.rint(lens), spec.clip_lo, spec.clip_hi)
    if spec.tail_frac > 0:
        tail = rng.random(n) < spec.tail_frac
        pareto = spec.tail_min * (1.0 + rng.pareto(spec.tail_alpha, size=n))
        lens = np.where(tail, np.clip(np.rint(pareto), spec.tail_min, spec.tail_max), lens)
    return lens.astype(np.int64)


def mutate(rng, root: np.ndarray, identity: float, indel_rate: float = 0.02) -> np.ndarray:
    """One descendant of ``root`` at the target identity."""
    n = len(root)
    seq = root.copy()
    sub = rng.random(n) < (1.0 - identity)
    if sub.any():
        src = seq[sub]
        u = rng.random(src.size)
        seq[sub] = (u[:, None] > _SUB_CDF[src]).sum(axis=1).clip(0, 19).astype(np.uint8)
    ev = np.flatnonzero(rng.random(n) < indel_rate)
    if ev.size == 0:
        return seq
    kinds = rng.random(ev.size) < 0.5
    glen = rng.geometric(1.0 / 3.0, size=ev.size)
    parts = []
    cur = 0
    for pos, ins, ln in zip(ev.tolist(), kinds.tolist(), glen.tolist()):
        if pos < cur:
            continue
        parts.append(seq[cur:pos])
        if ins:
            parts.append(_draw(rng, _BG_CDF, ln))
            cur = pos
        else:
            cur = min(n, pos + ln)
    parts.append(seq[cur:])
    out = np.concatenate(parts)
    return out if out.size else seq[:1].copy()


def generate_families(spec: FamilySpec) -> Family:
    rng = np.random.default_rng(spec.seed)
    rl = _root_lengths(rng, spec, spec.n_roots)
    seqs, root_of, ident = [], [], []
    for r in range(spec.n_roots):
        root = _draw(rng, _BG_CDF, int(rl[r]))
        seqs.append(root)
        root_of.append(r)
        ident.append(1.0)
        for _ in range(spec.n_desc):
            t = rng.uniform(spec.id_lo, spec.id_hi)
            seqs.append(mutate(rng, root, t, spec.indel_rate))
            root_of.append(r)
            ident.append(t)
    flat, offsets, lengths = pack_sequences(seqs)
    return Family(flat, offsets, lengths, np.asarray(root_of, np.int32),
                  np.asarray(ident, np.float32), spec)


def _roots_chunk(args):
    spec, r0, r1 = args
    seqs, root_of, ident = [], [], []
    for r in range(r0, r1):
        rng = np.random.default_rng([spec.seed, r])
        root = _draw(rng, _BG_CDF, int(_root_lengths(rng, spec, 1)[0]))
        seqs.append(root)
        root_of.append(r)
        ident.append(1.0)
        for _ in range(spec.n_desc):
            t = rng.uniform(spec.id_lo, spec.id_hi)
            seqs.append(mutate(rng, root, t, spec.indel_rate))
            root_of.append(r)
            ident.append(t)
    flat, _, lengths = pack_sequences(seqs)
    return flat, lengths, np.asarray(root_of, np.int32), np.asarray(ident, np.float32)


def generate_families_streams(spec: FamilySpec, workers: int = 1, chunk: int = 2000,
                              progress=None) -> Family:
    """Like ``generate_families`` but every root family draws from its own
    random stream (seeded by (seed, root)), so the output does not depend on
    how roots are split over ``workers`` processes (spawned, not forked, so
    that a parent holding a GPU context is safe).  ``progress(msg)`` is called
    now and then when generating in-process."""
    jobs = [(spec, r, min(r + chunk, spec.n_roots)) for r in range(0, spec.n_roots, chunk)]
    if workers > 1 and len(jobs) > 1:
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(min(workers, len(jobs))) as pool:
            parts = pool.map(_roots_chunk, jobs)
    else:
        parts = []
        for i, j in enumerate(jobs):
            parts.append(_roots_chunk(j))
            if progress is not None and (i + 1) % 10 == 0:
                progress(f"generated {j[2]} of {spec.n_roots} root families")
    flat = np.concatenate([p[0] for p in parts]) if parts else np.zeros(0, np.uint8)
    lengths = np.concatenate([p[1] for p in parts]) if parts else np.zeros(0, np.int32)
    offsets = np.zeros(len(lengths), np.int64)
    if len(lengths):
        np.cumsum(lengths[:-1].astype(np.int64), out=offsets[1:])
    return Family(flat, offsets, lengths, np.concatenate([p[2] for p in parts]),
                  np.concatenate([p[3] for p in parts]), spec)


def make_queries(db: Family, n: int, seed: int, planted_frac: float = 0.5,
                 len_lo: int = 50, len_hi: int = 2000, id_lo: float = 0.3, id_hi: float = 0.9,
                 median: float = 350.0, sigma: float = 0.5):
    """Search queries: a ``planted_frac`` share are descendants of random
    database sequences at identity U(id_lo, id_hi); the rest are random
    background proteins with lognormal lengths.  Lengths are kept within
    [len_lo, len_hi].  Returns (list of uint8 arrays, source db index or -1)."""
    rng = np.random.default_rng(seed)
    out, src = [], []
    ok = np.flatnonzero((db.lengths >= len_lo) & (db.lengths <= len_hi))
    while len(out) < n:
        if rng.random() < planted_frac and len(ok):
            s = int(ok[rng.integers(len(ok))])
            q = mutate(rng, db.flat[db.offsets[s]:db.offsets[s] + db.lengths[s]],
                       float(rng.uniform(id_lo, id_hi)))
        else:
            s = -1
            ln = int(np.clip(np.rint(rng.lognormal(np.log(median), sigma)), len_lo, len_hi))
            q = _draw(rng, _BG_CDF, ln)
        if len(q) < len_lo:
            continue
        out.append(q[:len_hi])
        src.append(s)
    return out, np.asarray(src, np.int64)


SPECS = {
    "small": FamilySpec(n_roots=200, median=300, clip_lo=50, clip_hi=1000, id_lo=0.3,
                        id_hi=0.9, seed=1),
    "mixed2m": FamilySpec(n_roots=5000, median=350, clip_lo=30, clip_hi=2000, id_lo=0.3,
                          id_hi=0.95, seed=2),
    "heavy": FamilySpec(n_roots=10000, median=250, clip_lo=30, clip_hi=2000, id_lo=0.3,
                        id_hi=0.95, tail_frac=0.1, seed=3),
    "fast200k": FamilySpec(n_roots=20000, median=300, clip_lo=30, clip_hi=3000, id_lo=0.25,
                           id_hi=0.95, seed=4),
    "serve1m": FamilySpec(n_roots=100000, median=350, clip_lo=30, clip_hi=5000, id_lo=0.3,
                          id_hi=0.9, seed=5),
}
