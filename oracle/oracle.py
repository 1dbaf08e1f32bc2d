"""ctypes wrappers over the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker or as the timed CPU baseline.
The shipped package never imports it.

* ``sw_batch`` / ``sw_pair``      -> oracle/liboracle.so (own restatement of
  clustkit/csrc/sw_align.c:25-266).
* ``ref_sw_batch``                -> oracle/_ref/sw_align.so, the reference
  kernel compiled from its own sources (``make -C oracle ref``), called the
  way clustkit/pairwise.py:69-108 calls it.
* ``jaccard_batch``               -> restatement of pairwise.py:1564-1612.
* ``length_ratio_pass``           -> closed form of pairwise.py:1243-1252.
* ``nw_batch`` / ``nw_pair``      -> restatement of the banded global
  alignment of pairwise.py:258-505 and its batch forms 684-736, 804-850.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None
_REF = None
_KREF = None

_P = ctypes.c_void_p


def build(ref: bool = True) -> None:
    """Compile the oracle (and the reference objects where the reference
    sources exist).  Always recompiles."""
    subprocess.run(["make", "-s", "-B", "-C", str(_HERE), "liboracle.so"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-B", "-C", str(_HERE), "ref"], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build(ref=False)
        lib = ctypes.CDLL(str(path))
        lib.sw_oracle_batch.restype = None
        lib.sw_oracle_batch.argtypes = [_P, _P, _P, _P, ctypes.c_int64, ctypes.c_int32,
                                        _P, _P, ctypes.c_int32, _P, _P, _P, _P, _P]
        lib.sw_oracle_pair.restype = None
        lib.sw_oracle_pair.argtypes = [_P, ctypes.c_int32, _P, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, _P, _P, _P]
        lib.kmer_oracle_batch.restype = None
        lib.kmer_oracle_batch.argtypes = [_P, _P, _P, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, _P, _P, _P] + [ctypes.c_int32] * 8 + \
            [_P, _P, _P, _P]
        lib.kmer_similar_oracle_batch.restype = None
        lib.kmer_similar_oracle_batch.argtypes = [_P, _P, _P, ctypes.c_int32, _P, _P, _P] + \
            [ctypes.c_int32] * 4 + [_P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P]
        lib.nw_oracle_batch.restype = None
        lib.nw_oracle_batch.argtypes = [_P, ctypes.c_int64, _P, _P, _P, ctypes.c_float,
                                        ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_int32,
                                        _P, _P, _P]
        lib.nw_oracle_pair.restype = ctypes.c_float
        lib.nw_oracle_pair.argtypes = [_P, ctypes.c_int32, _P, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_float, ctypes.c_int32, _P, _P, _P]
        lib.jaccard_oracle_batch.restype = None
        lib.jaccard_oracle_batch.argtypes = [_P, _P, ctypes.c_int64, ctypes.c_int32,
                                             ctypes.c_int32, _P]
        _LIB = lib
    return _LIB


def ref_available() -> bool:
    return (_HERE / "_ref" / "sw_align.so").exists()


def _ref():
    global _REF
    if _REF is None:
        lib = ctypes.CDLL(str(_HERE / "_ref" / "sw_align.so"))
        lib.batch_sw_align_c.restype = None
        lib.batch_sw_align_c.argtypes = [_P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32,
                                         _P, ctypes.c_float, _P, _P, _P, _P]
        lib.sw_set_num_threads.restype = None
        lib.sw_set_num_threads.argtypes = [ctypes.c_int32]
        _REF = lib
    return _REF


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def sw_batch(pairs, flat, offsets, lengths, band_width, sub_matrix, hints=None,
             n_threads: int = 0):
    """Oracle batch SW.  Returns dict of sims f32, scores i32, mask u8,
    matches i32, end (M,2) i32 (1-based row/col of the first best cell in the
    oriented problem, shorter sequence on rows)."""
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    m = len(pairs)
    flat = _c(flat, np.uint8)
    if flat.size == 0:
        flat = np.zeros(1, np.uint8)
    offsets = _c(offsets, np.int64)
    lengths = _c(lengths, np.int32)
    sub = _c(sub_matrix, np.int8).reshape(-1)
    h = None if hints is None else _c(hints, np.int32)
    sims = np.zeros(m, np.float32)
    scores = np.zeros(m, np.int32)
    mask = np.zeros(m, np.uint8)
    matches = np.zeros(m, np.int32)
    end = np.zeros((m, 2), np.int32)
    if m:
        _lib().sw_oracle_batch(pairs.ctypes.data, flat.ctypes.data, offsets.ctypes.data,
                               lengths.ctypes.data, m, int(band_width), sub.ctypes.data,
                               None if h is None else h.ctypes.data, int(n_threads),
                               sims.ctypes.data, scores.ctypes.data, mask.ctypes.data,
                               matches.ctypes.data, end.ctypes.data)
    return {"sims": sims, "scores": scores, "mask": mask, "matches": matches, "end": end}


def nw_batch(pairs, flat, offsets, lengths, threshold, band_width, sub_matrix=None,
             n_threads: int = 0):
    """Oracle batch global alignment.  sub_matrix None -> the plain form
    (pairwise.py:804-850: +2/-1, gaps -10/-1, returns sims, mask); else the
    scored form (pairwise.py:684-736: returns sims, scores, mask)."""
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    m = len(pairs)
    flat = _c(flat, np.uint8)
    if flat.size == 0:
        flat = np.zeros(1, np.uint8)
    offsets = _c(offsets, np.int64)
    lengths = _c(lengths, np.int32)
    scored = sub_matrix is not None
    sub = _c(sub_matrix if scored else np.zeros(400, np.int8), np.int8).reshape(-1)
    sims = np.zeros(m, np.float32)
    scores = np.zeros(m, np.int32)
    mask = np.zeros(m, np.uint8)
    if m:
        _lib().nw_oracle_batch(pairs.ctypes.data, m, flat.ctypes.data, offsets.ctypes.data,
                               lengths.ctypes.data, float(np.float32(threshold)),
                               int(band_width), int(scored), sub.ctypes.data, int(n_threads),
                               sims.ctypes.data, scores.ctypes.data, mask.ctypes.data)
    if scored:
        return sims, scores, mask.astype(bool)
    return sims, mask.astype(bool)


def nw_pair(seq_a, seq_b, band_width, threshold=0.0, sub_matrix=None):
    """One oriented problem (rows = seq_a; the caller puts the shorter first):
    (identity f32, score i32), the inner function without the batch tests."""
    a = _c(seq_a, np.uint8)
    b = _c(seq_b, np.uint8)
    scored = sub_matrix is not None
    sub = _c(sub_matrix if scored else np.zeros(400, np.int8), np.int8).reshape(-1)
    ws = np.zeros(8 * (max(len(a), len(b)) + 1), np.int32)
    sc = np.zeros(1, np.int32)
    a1 = a if a.size else np.zeros(1, np.uint8)
    b1 = b if b.size else np.zeros(1, np.uint8)
    ident = _lib().nw_oracle_pair(a1.ctypes.data, len(a), b1.ctypes.data, len(b),
                                  int(band_width), float(np.float32(threshold)), int(scored),
                                  sub.ctypes.data, ws.ctypes.data, sc.ctypes.data)
    return np.float32(ident), int(sc[0])


def sw_pair(seq_i, seq_j, band_width, sub_matrix, hint: int = 0):
    a = _c(seq_i, np.uint8)
    b = _c(seq_j, np.uint8)
    sub = _c(sub_matrix, np.int8).reshape(-1)
    out5 = np.zeros(5, np.int32)
    ident = np.zeros(1, np.float32)
    _lib().sw_oracle_pair(a.ctypes.data if a.size else None, len(a),
                          b.ctypes.data if b.size else None, len(b),
                          int(band_width), int(hint), sub.ctypes.data,
                          out5.ctypes.data, ident.ctypes.data)
    return {"score": int(out5[0]), "matches": int(out5[1]), "end": (int(out5[2]), int(out5[3])),
            "swapped": bool(out5[4]), "identity": np.float32(ident[0])}


def ref_sw_batch(pairs, flat, offsets, lengths, band_width, sub_matrix, hints=None,
                 n_threads: int = 1, threshold: float = 0.0):
    """The reference kernel (compiled from its own sources) called as
    clustkit/pairwise.py:69-108 does.  Returns (sims, scores, mask)."""
    lib = _ref()
    pairs = _c(pairs, np.int32).reshape(-1)
    m = len(pairs) // 2
    flat = _c(flat, np.uint8)
    if flat.size == 0:
        flat = np.zeros(1, np.uint8)
    offsets = _c(offsets, np.int64)
    lengths = _c(lengths, np.int32)
    sub = _c(sub_matrix, np.int8).reshape(-1)
    h = None if hints is None else _c(hints, np.int32)
    sims = np.empty(m, np.float32)
    scores = np.empty(m, np.int32)
    mask = np.empty(m, np.uint8)
    lib.sw_set_num_threads(int(n_threads))
    lib.batch_sw_align_c(pairs.ctypes.data, flat.ctypes.data, offsets.ctypes.data,
                         lengths.ctypes.data, m, int(band_width), sub.ctypes.data,
                         np.float32(threshold), None if h is None else h.ctypes.data,
                         sims.ctypes.data, scores.ctypes.data, mask.ctypes.data)
    return sims, scores, mask


def jaccard_batch(pairs, sketches, n_threads: int = 0):
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    sk = _c(sketches, np.uint64)
    out = np.zeros(len(pairs), np.float32)
    if len(pairs):
        _lib().jaccard_oracle_batch(pairs.ctypes.data, sk.ctypes.data, len(pairs),
                                    sk.shape[1], int(n_threads), out.ctypes.data)
    return out


def jaccard_prefilter(pairs, sketches, threshold: float, min_filter_rate: float = 0.05):
    """Restatement of pairwise.py:949-1001 on top of ``jaccard_batch``."""
    m = len(pairs)
    if m == 0:
        return np.ones(0, dtype=bool)
    floor = max(0.001, 0.3 * threshold ** 5)
    if min(2000, m) < m:
        idx = np.random.RandomState(0).choice(m, size=2000, replace=False)
        s = jaccard_batch(np.asarray(pairs)[idx], sketches)
        if float(np.mean(s < floor)) < min_filter_rate:
            return None
    return jaccard_batch(pairs, sketches) >= np.float32(floor)


def length_ratio_pass(pairs, lengths, threshold: float):
    """pairwise.py:1243-1252: fails when longer > 0 and
    fl32(shorter)/fl32(longer) < fl32(threshold)."""
    pairs = np.asarray(pairs).reshape(-1, 2)
    li = np.asarray(lengths)[pairs[:, 0]].astype(np.int64)
    lj = np.asarray(lengths)[pairs[:, 1]].astype(np.int64)
    sh = np.minimum(li, lj).astype(np.float32)
    lo = np.maximum(li, lj)
    ratio = np.where(lo > 0, sh / np.maximum(lo, 1).astype(np.float32), np.float32(1))
    return ~((lo > 0) & (ratio < np.float32(threshold)))


# --- k-mer index scoring (clustkit/csrc/kmer_score.c:198-472) ----------------

_KMER_DEFAULTS = dict(min_total_hits=2, min_diag_hits=2, diag_bin_width=10,
                      max_cands=2000, phase_a_topk=10000)


def _kmer_args(q_flat, q_off, q_len, kmer_offsets, kmer_entries, kmer_freqs):
    return (_c(q_flat if len(q_flat) else np.zeros(1, np.uint8), np.uint8),
            _c(q_off, np.int64), _c(q_len, np.int32), _c(kmer_offsets, np.int64),
            _c(kmer_entries, np.int64), _c(kmer_freqs, np.int32))


def kmer_score_batch(q_flat, q_off, q_len, k, alpha, kmer_offsets, kmer_entries, kmer_freqs,
                     freq_thresh, num_db, n_threads: int = 0, **kw):
    """Oracle Phase A/B scoring.  Returns (targets (nq, mc), counts, scores, diags)."""
    o = dict(_KMER_DEFAULTS, **kw)
    qf, qo, ql, ko, ke, kf = _kmer_args(q_flat, q_off, q_len, kmer_offsets, kmer_entries,
                                        kmer_freqs)
    nq, mc = len(ql), int(o["max_cands"])
    tg = np.zeros((nq, mc), np.int32)
    cn = np.zeros(nq, np.int32)
    sc = np.zeros((nq, mc), np.int32)
    dg = np.zeros((nq, mc), np.int32)
    if nq:
        _lib().kmer_oracle_batch(qf.ctypes.data, qo.ctypes.data, ql.ctypes.data, nq, int(k),
                                 int(alpha), ko.ctypes.data, ke.ctypes.data, kf.ctypes.data,
                                 int(freq_thresh), int(num_db), int(o["min_total_hits"]),
                                 int(o["min_diag_hits"]), int(o["diag_bin_width"]), mc,
                                 int(o["phase_a_topk"]), int(n_threads), tg.ctypes.data,
                                 cn.ctypes.data, sc.ctypes.data, dg.ctypes.data)
    return tg, cn, sc, dg


def kmer_similar_batch(q_flat, q_off, q_len, kmer_offsets, kmer_entries, kmer_freqs,
                       freq_thresh, num_db, sub_matrix, kmer_score_thresh, min_total_hits=2,
                       max_cands=2000, n_threads: int = 0):
    """Oracle BLOSUM-similar 5-mer scoring (k = 5, alphabet 20).  Returns
    (targets (nq, max_cands), counts, scores (the weighted counts))."""
    qf, qo, ql, ko, ke, kf = _kmer_args(q_flat, q_off, q_len, kmer_offsets, kmer_entries,
                                        kmer_freqs)
    sub = np.ascontiguousarray(np.asarray(sub_matrix, np.int8).reshape(-1))
    nq, mc = len(ql), int(max_cands)
    tg = np.zeros((nq, mc), np.int32)
    cn = np.zeros(nq, np.int32)
    sc = np.zeros((nq, mc), np.int32)
    if nq:
        _lib().kmer_similar_oracle_batch(qf.ctypes.data, qo.ctypes.data, ql.ctypes.data, nq,
                                         ko.ctypes.data, ke.ctypes.data, kf.ctypes.data,
                                         int(freq_thresh), int(num_db), int(min_total_hits),
                                         mc, sub.ctypes.data, int(kmer_score_thresh),
                                         int(n_threads), tg.ctypes.data, cn.ctypes.data,
                                         sc.ctypes.data)
    return tg, cn, sc


def ref_kmer_available() -> bool:
    return (_HERE / "_ref" / "kmer_score.so").exists()


def ref_kmer_score_batch(q_flat, q_off, q_len, k, alpha, kmer_offsets, kmer_entries,
                         kmer_freqs, freq_thresh, num_db, compact=True, n_threads: int = 1,
                         **kw):
    """The reference batch_score_queries_c (compiled from its sources), called as
    clustkit/kmer_index.py:1640-1665 does."""
    global _KREF
    if _KREF is None:
        lib = ctypes.CDLL(str(_HERE / "_ref" / "kmer_score.so"))
        lib.batch_score_queries_c.restype = None
        lib.batch_score_queries_c.argtypes = [_P, _P, _P, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_int32, _P, _P, _P, _P] + \
            [ctypes.c_int32] * 7 + [_P, _P, _P, _P]
        lib.kmer_set_num_threads.restype = None
        lib.kmer_set_num_threads.argtypes = [ctypes.c_int32]
        _KREF = lib
    o = dict(_KMER_DEFAULTS, **kw)
    qf, qo, ql, ko, ke, kf = _kmer_args(q_flat, q_off, q_len, kmer_offsets, kmer_entries,
                                        kmer_freqs)
    comp = _c(ke >> 32, np.int32) if compact else None
    nq, mc = len(ql), int(o["max_cands"])
    tg = np.zeros((nq, mc), np.int32)
    cn = np.zeros(nq, np.int32)
    sc = np.zeros((nq, mc), np.int32)
    dg = np.zeros((nq, mc), np.int32)
    _KREF.kmer_set_num_threads(int(n_threads))
    if nq:
        _KREF.batch_score_queries_c(qf.ctypes.data, qo.ctypes.data, ql.ctypes.data, nq, int(k),
                                    int(alpha), ko.ctypes.data, ke.ctypes.data,
                                    None if comp is None else comp.ctypes.data, kf.ctypes.data,
                                    int(freq_thresh), int(num_db), int(o["min_total_hits"]),
                                    int(o["min_diag_hits"]), int(o["diag_bin_width"]), mc,
                                    int(o["phase_a_topk"]), tg.ctypes.data, cn.ctypes.data,
                                    sc.ctypes.data, dg.ctypes.data)
    return tg, cn, sc, dg
