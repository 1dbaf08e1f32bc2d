"""Drop-in for the alignment stage of clustkit/pairwise.py, on the MI355X.

Keeps the reference's entry points, argument meaning and return types:

* ``compute_pairwise_alignment``  (pairwise.py:1047-1208)
* ``_batch_sw_c``                 (pairwise.py:69-108)
* ``_batch_sw_compact_scored``    (pairwise.py:636-681)
* ``batch_sw_align``              the C kernel's contract (csrc/sw_align.c:182-266)
* ``_jaccard_prefilter``          (pairwise.py:949-1001)
* ``_batch_jaccard``, ``jaccard_from_sketches``, ``compute_pairwise_jaccard``
                                  (pairwise.py:1564-1661)
* ``BLOSUM62`` and the gap constants (pairwise.py:129-154)
* ``_batch_align_compact``, ``_batch_align_compact_scored``, ``_batch_align``,
  ``_nw_identity``, ``_nw_identity_and_score`` -- banded global alignment
                                  (pairwise.py:258-505, 684-850), and the
                                  refinement of borderline SW hits
                                  (pairwise.py:1004-1044)

Every alignment and Jaccard evaluation runs in the HIP library
(libclustkit_hip.so) through its C-ABI; nothing here computes a score on the
host.  Without the library or a GPU the calls raise.

Semantics per reference branch of ``compute_pairwise_alignment`` (the
reference's own branches differ in edge cases; each is reproduced):

* ``device != "cpu"`` -- the reference's GPU branch (pairwise.py:1458-1557):
  SW with the length-ratio pre-filter, pass = score > 0 and identity >= t.
* ``device == "cpu"``, ``use_sw`` and ``use_c_sw`` -- the C branch
  (pairwise.py:1151-1171): no length-ratio filter, pass = identity >= t.
* ``device == "cpu"``, ``use_sw`` without ``use_c_sw`` -- the Numba branch
  (pairwise.py:1175-1194): pass = score > 0 and identity >= t.
* ``device == "cpu"`` without ``use_sw`` -- global NW identity
  (pairwise.py:1197-1206), pass = identity >= t.
* ``refine_global_margin > 0`` on the CPU SW branches -- SW hits below
  t + margin are re-scored by global NW identity (pairwise.py:1004-1044).

The "cpu" device value selects the reference branch whose results are
reproduced; the arithmetic still runs on the GPU.  Output order follows the
selected branch unless ``pair_order`` says otherwise: the CPU branches return
passing pairs in stable min(i, j) order (pairwise.py:1109-1111); the GPU
branch re-sorts that list stably by max(len_i, len_j) and returns the passing
pairs in that order (pairwise.py:1480-1487, 1553-1555).  ``pair_order`` =
"min_index" or "max_length" forces one of the two for any branch.
"""

from __future__ import annotations

import logging

import numpy as np

from . import _lib
from .scoring import BLOSUM62, GAP_EXTEND, GAP_OPEN

_logger = logging.getLogger("clustkit")

BLOSUM62_GAP_OPEN = np.int32(GAP_OPEN)
BLOSUM62_GAP_EXTEND = np.int32(GAP_EXTEND)

_EMPTY_PAIRS = np.empty((0, 2), dtype=np.int32)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def batch_sw_align(pairs, flat_sequences, offsets, lengths, band_width, sub_matrix=BLOSUM62,
                   diag_hints=None, threshold: float = 0.0, length_ratio: bool = False,
                   mask_identity: bool = False, plain: bool = False, latency: bool = False):
    """Banded SW over candidate pairs with the contract of batch_sw_align_c
    (sw_align.c:182-266).  Returns (sims f32, scores i32, matches i32,
    mask u8).  ``mask`` is score > 0 unless ``mask_identity`` (then
    score > 0 and identity >= threshold).  ``latency`` runs the batched
    kernel one pair per wavefront (small batches; same results)."""
    lib = _lib.require_gpu()
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    m = len(pairs)
    sims = np.zeros(m, np.float32)
    scores = np.zeros(m, np.int32)
    matches = np.zeros(m, np.int32)
    mask = np.zeros(m, np.uint8)
    if m == 0:
        return sims, scores, matches, mask
    flat = _c(flat_sequences, np.uint8)
    if flat.size == 0:
        flat = np.zeros(1, np.uint8)
    offsets = _c(offsets, np.int64)
    lengths = _c(lengths, np.int32)
    sub = _c(sub_matrix, np.int8).reshape(-1)
    if sub.size != 400:
        raise ValueError("sub_matrix must be 20x20")
    hints = None if diag_hints is None else _c(diag_hints, np.int32)
    if hints is not None and len(hints) != m:
        raise ValueError("diag_hints must have one entry per pair")
    flags = ((_lib.CK_SW_LENGTH_RATIO if length_ratio else 0)
             | (_lib.CK_SW_MASK_IDENTITY if mask_identity else 0)
             | (_lib.CK_SW_PLAIN if plain else 0)
             | (_lib.CK_SW_LATENCY if latency else 0))
    _lib.check(lib.ck_sw_align_host(
        pairs.ctypes.data, m, flat.ctypes.data, flat.size, offsets.ctypes.data,
        lengths.ctypes.data, len(lengths), int(band_width), sub.ctypes.data,
        None if hints is None else hints.ctypes.data, float(np.float32(threshold)), flags,
        sims.ctypes.data, scores.ctypes.data, matches.ctypes.data, mask.ctypes.data))
    return sims, scores, matches, mask


def _batch_sw_c(pairs, flat_sequences, offsets, lengths, threshold, band_width, sub_matrix,
                n_threads=4):
    """pairwise.py:69-108: returns (sims, sims >= threshold).  ``n_threads``
    configured OpenMP in the reference and is accepted for signature
    compatibility."""
    sims, _, _, _ = batch_sw_align(pairs, flat_sequences, offsets, lengths, band_width,
                                   sub_matrix)
    return sims, sims >= np.float32(threshold)


def _batch_sw_compact_scored(pairs, flat_sequences, offsets, lengths, threshold, band_width,
                             sub_matrix):
    """pairwise.py:636-681: returns (sims, scores, mask = score > 0)."""
    sims, scores, _, mask = batch_sw_align(pairs, flat_sequences, offsets, lengths,
                                           band_width, sub_matrix)
    return sims, scores, mask.astype(bool)


# ---------------------------------------------------------------------------
# Jaccard pre-filter
# ---------------------------------------------------------------------------

def _batch_jaccard(pairs, sketches, threshold):
    """pairwise.py:1598-1612: (sims f32, sims >= threshold)."""
    lib = _lib.require_gpu()
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    sk = _c(sketches, np.uint64)
    if sk.ndim != 2:
        raise ValueError("sketches must be (N, sketch_size)")
    sims = np.zeros(len(pairs), np.float32)
    if len(pairs):
        _lib.check(lib.ck_jaccard_host(pairs.ctypes.data, len(pairs), sk.ctypes.data,
                                       sk.shape[0], sk.shape[1], sims.ctypes.data))
    return sims, sims >= np.float32(threshold)


def jaccard_from_sketches(sketch_a: np.ndarray, sketch_b: np.ndarray) -> float:
    """Jaccard estimate of two sorted sketches (pairwise.py:1615-1620)."""
    sk = np.stack([_c(sketch_a, np.uint64), _c(sketch_b, np.uint64)])
    sims, _ = _batch_jaccard(np.array([[0, 1]], np.int32), sk, 0.0)
    return float(sims[0])


def compute_pairwise_jaccard(candidate_pairs, sketches, threshold, device="cpu"):
    """pairwise.py:1623-1661: (filtered_pairs, similarities)."""
    if len(candidate_pairs) == 0:
        return _EMPTY_PAIRS.copy(), np.empty(0, dtype=np.float32)
    pairs = _c(candidate_pairs, np.int32).reshape(-1, 2)
    sims, mask = _batch_jaccard(pairs, sketches, np.float32(threshold))
    return pairs[mask], sims[mask]


def _jaccard_prefilter(candidate_pairs, sketches, threshold, min_filter_rate: float = 0.05):
    """pairwise.py:949-1001.  Floor max(0.001, 0.3 * t**5); a 2000-pair
    sample drawn with RandomState(0) decides whether the filter is worth
    applying (None = skip)."""
    m = len(candidate_pairs)
    if m == 0:
        return np.ones(0, dtype=np.bool_)
    floor = max(0.001, 0.3 * threshold ** 5)
    sample_size = min(2000, m)
    if sample_size < m:
        idx = np.random.RandomState(0).choice(m, size=sample_size, replace=False)
        s, _ = _batch_jaccard(np.asarray(candidate_pairs)[idx], sketches, np.float32(floor))
        rate = float(np.mean(s < floor))
        if rate < min_filter_rate:
            _logger.info(f"  Jaccard pre-filter skipped: estimated {rate:.1%} removal rate "
                         f"below {min_filter_rate:.0%} threshold")
            return None
    sims, _ = _batch_jaccard(candidate_pairs, sketches, np.float32(floor))
    return sims >= floor


# ---------------------------------------------------------------------------
# Dispatcher
# ---------------------------------------------------------------------------

# ---------------------------------------------------------------------------
# Banded global alignment (NW identity)
# ---------------------------------------------------------------------------

def batch_nw_align(pairs, flat_sequences, offsets, lengths, threshold, band_width,
                   sub_matrix=None, *, plain_kernel: bool = False, raw: bool = False,
                   stripe_kernel: bool = False):
    """Global-alignment identity of every pair on the GPU (ck_nw_align_host).

    sub_matrix None: the plain scoring of pairwise.py:258-362 (+2 / -1, gaps
    -10 / -1); a 20x20 table: pairwise.py:365-505 (gaps -11 / -1).  Batch
    rules of pairwise.py:684-736 / 804-850 unless ``raw`` (the inner function
    alone, band as given).  ``plain_kernel`` / ``stripe_kernel`` select the
    one-pair-per-thread kernel / the packed-key kernel one stripe at a time
    instead of the stripe-pipelined one (same results).  Returns (sims f32,
    scores i32, mask bool)."""
    lib = _lib.require_gpu()
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    m = len(pairs)
    flat = _c(flat_sequences, np.uint8)
    offsets = _c(offsets, np.int64)
    lengths = _c(lengths, np.int32)
    sims = np.zeros(m, np.float32)
    scores = np.zeros(m, np.int32)
    mask = np.zeros(m, np.uint8)
    if m == 0:
        return sims, scores, mask.astype(bool)
    sub = None if sub_matrix is None else _c(sub_matrix, np.int8).reshape(-1)
    if sub is not None and sub.size != 400:
        raise ValueError("sub_matrix must be 20x20")
    flags = ((_lib.CK_NW_PLAIN if plain_kernel else 0) | (_lib.CK_NW_RAW if raw else 0)
             | (_lib.CK_NW_STRIPE if stripe_kernel else 0))
    fl = flat if flat.size else np.zeros(1, np.uint8)
    _lib.check(lib.ck_nw_align_host(pairs.ctypes.data, m, fl.ctypes.data, flat.size,
                                    offsets.ctypes.data, lengths.ctypes.data, len(lengths),
                                    int(band_width), None if sub is None else sub.ctypes.data,
                                    float(np.float32(threshold)), flags, sims.ctypes.data,
                                    scores.ctypes.data, mask.ctypes.data))
    return sims, scores, mask.astype(bool)


def _batch_align_compact(pairs, flat_sequences, offsets, lengths, threshold, band_width):
    """pairwise.py:804-850: (sims f32, mask = identity >= threshold)."""
    sims, _, mask = batch_nw_align(pairs, flat_sequences, offsets, lengths, threshold,
                                   band_width)
    return sims, mask


def _batch_align_compact_scored(pairs, flat_sequences, offsets, lengths, threshold,
                                band_width, sub_matrix):
    """pairwise.py:684-736: (sims f32, scores i32, mask = identity >= threshold);
    score -1,000,000 for pairs removed by the length tests or stopped early."""
    return batch_nw_align(pairs, flat_sequences, offsets, lengths, threshold, band_width,
                          sub_matrix)


def _padded_to_flat(sequences, lengths):
    enc = np.asarray(sequences)
    lengths = _c(lengths, np.int32)
    offsets = np.zeros(len(lengths), np.int64)
    if len(lengths):
        np.cumsum(lengths[:-1].astype(np.int64), out=offsets[1:])
    flat = np.concatenate([enc[k, :lengths[k]] for k in range(len(lengths))]) \
        if len(lengths) else np.zeros(0, np.uint8)
    return _c(flat, np.uint8), offsets, lengths


def _batch_align(pairs, sequences, lengths, threshold, band_width):
    """pairwise.py:739-801 over a padded (N, max_len) matrix."""
    flat, offsets, lengths = _padded_to_flat(sequences, lengths)
    return _batch_align_compact(pairs, flat, offsets, lengths, threshold, band_width)


def _nw_one(seq_a, len_a, seq_b, len_b, band_width, threshold, sub_matrix):
    len_a, len_b = int(len_a), int(len_b)
    if len_a > len_b:
        raise ValueError("the shorter sequence goes first (as every caller in the reference "
                         "passes it)")
    a = _c(np.asarray(seq_a)[:len_a], np.uint8)
    b = _c(np.asarray(seq_b)[:len_b], np.uint8)
    flat = np.concatenate([a, b])
    offsets = np.array([0, len_a], np.int64)
    lengths = np.array([len_a, len_b], np.int32)
    sims, scores, _ = batch_nw_align(np.array([[0, 1]], np.int32), flat, offsets, lengths,
                                     threshold, band_width, sub_matrix, raw=True)
    return np.float32(sims[0]), np.int32(scores[0])


def _nw_identity(seq_a, len_a, seq_b, len_b, band_width, threshold):
    """pairwise.py:258-362 on one pair (len_a <= len_b): identity f32."""
    return _nw_one(seq_a, len_a, seq_b, len_b, band_width, threshold, None)[0]


def _nw_identity_and_score(seq_a, len_a, seq_b, len_b, band_width, threshold, sub_matrix):
    """pairwise.py:365-505 on one pair (len_a <= len_b): (identity f32, score i32)."""
    return _nw_one(seq_a, len_a, seq_b, len_b, band_width, threshold, sub_matrix)


def _refine_borderline_sw_hits(kept_pairs, kept_sims, *, threshold, refine_global_margin,
                               flat_sequences, offsets, lengths, band_width):
    """pairwise.py:1004-1044: SW hits below threshold + margin are re-scored
    with global NW identity and kept only if that passes."""
    if refine_global_margin <= 0 or len(kept_pairs) == 0:
        return kept_pairs, kept_sims
    refine_mask = kept_sims < (threshold + refine_global_margin)
    if not np.any(refine_mask):
        return kept_pairs, kept_sims
    _logger.info(f"  Refining {int(np.sum(refine_mask))} borderline SW hits with global NW")
    nw_sims, nw_keep = _batch_align_compact(kept_pairs[refine_mask], flat_sequences, offsets,
                                            lengths, np.float32(threshold), np.int32(band_width))
    kept_sims = kept_sims.copy()
    kept_sims[refine_mask] = nw_sims
    final_mask = np.ones(len(kept_pairs), dtype=np.bool_)
    final_mask[refine_mask] = nw_keep
    return kept_pairs[final_mask], kept_sims[final_mask]


def compute_pairwise_alignment(candidate_pairs, encoded_sequences, lengths, threshold,
                               band_width: int = 50, device: str = "cpu",
                               mode: str = "protein", sketches=None, flat_sequences=None,
                               offsets=None, use_sw: bool = False, use_c_sw: bool = False,
                               n_threads: int = 4, refine_global_margin: float = 0.0,
                               pair_order: str | None = None):
    """pairwise.py:1047-1208: (filtered_pairs (K,2) int32, identities (K,) f32).

    ``pair_order`` (not in the reference signature; default None = the order
    of the reference branch that ``device`` / ``use_sw`` select): "min_index"
    = stable by min(i, j); "max_length" = that list re-sorted stably by
    max(len_i, len_j), the reference GPU branch's order.  The set of pairs
    and each pair's identity do not depend on it."""
    if pair_order not in (None, "min_index", "max_length"):
        raise ValueError("pair_order must be None, 'min_index' or 'max_length'")
    if len(candidate_pairs) == 0:
        return _EMPTY_PAIRS.copy(), np.empty(0, dtype=np.float32)
    pairs = _c(candidate_pairs, np.int32).reshape(-1, 2)
    lengths = _c(lengths, np.int32)

    if sketches is not None:
        m_before = len(pairs)
        jmask = _jaccard_prefilter(pairs, sketches, threshold)
        if jmask is not None:
            pairs = pairs[jmask]
            if m_before > len(pairs):
                _logger.info(f"  Jaccard pre-filter: {m_before} -> {len(pairs)} pairs")
            if len(pairs) == 0:
                return _EMPTY_PAIRS.copy(), np.empty(0, dtype=np.float32)

    order = np.argsort(np.minimum(pairs[:, 0], pairs[:, 1]), kind="mergesort")
    pairs = pairs[order]

    if flat_sequences is None or offsets is None:
        if encoded_sequences is None:
            raise ValueError("need flat_sequences/offsets or encoded_sequences")
        enc = np.asarray(encoded_sequences)
        offsets = np.zeros(len(lengths), np.int64)
        if len(lengths):
            np.cumsum(lengths[:-1].astype(np.int64), out=offsets[1:])
        flat_sequences = np.concatenate([enc[k, :lengths[k]] for k in range(len(lengths))]) \
            if len(lengths) else np.zeros(0, np.uint8)

    gpu_branch = device != "cpu"
    if pair_order is None:
        pair_order = "max_length" if gpu_branch else "min_index"
    if pair_order == "max_length":
        # pairwise.py:1480-1487: stable argsort of max(len_i, len_j) over the
        # min-index-sorted list; outputs stay in this order (1553-1555).
        pairs = pairs[np.argsort(np.maximum(lengths[pairs[:, 0]], lengths[pairs[:, 1]]),
                                 kind="mergesort")]
    t32 = np.float32(threshold)
    if not gpu_branch and not use_sw:
        sims, mask = _batch_align_compact(pairs, flat_sequences, offsets, lengths, t32,
                                          np.int32(band_width))
        return pairs[mask], sims[mask]

    if gpu_branch:
        sims, scores, _, mask = batch_sw_align(pairs, flat_sequences, offsets, lengths,
                                               band_width, BLOSUM62, threshold=t32,
                                               length_ratio=True, mask_identity=True)
        keep = mask.astype(bool)
    elif use_c_sw:
        sims, scores, _, _ = batch_sw_align(pairs, flat_sequences, offsets, lengths,
                                            band_width, BLOSUM62)
        keep = sims >= t32
    else:
        sims, scores, _, _ = batch_sw_align(pairs, flat_sequences, offsets, lengths,
                                            band_width, BLOSUM62)
        keep = (scores > 0) & (sims >= t32)
    if gpu_branch:
        return pairs[keep], sims[keep]
    return _refine_borderline_sw_hits(pairs[keep], sims[keep], threshold=threshold,
                                      refine_global_margin=refine_global_margin,
                                      flat_sequences=flat_sequences, offsets=offsets,
                                      lengths=lengths, band_width=band_width)
