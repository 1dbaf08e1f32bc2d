"""CSR k-mer inverted index of a sequence database (host build, NumPy).

Layout of clustkit/kmer_index.py:117-162 (and the database format of
clustkit/database.py:127-136):

* ``kmer_offsets`` int64[alpha**k + 1]  CSR row pointers by k-mer code
* ``kmer_entries`` int64[total]         ``seq_id << 32 | position``, ordered by
                                         code, then sequence, then position
* ``kmer_freqs``   int32[alpha**k]       occurrences per code
* compact Phase-A entries = ``kmer_entries >> 32`` as int32

A k-mer is indexed only when all its residue codes are below the alphabet
size (codes >= alpha, e.g. unknown residues, break the window).  The
frequency cap follows kmer_index.py:451-463.
"""

from __future__ import annotations

import numpy as np


def window_codes(flat, offsets, lengths, k: int, alpha: int):
    """Codes of every valid k-mer window, in (sequence, position) order.

    Returns (codes int64, seq int64, pos int64)."""
    flat = np.asarray(flat, dtype=np.uint8)
    offsets = np.asarray(offsets, dtype=np.int64)
    lengths = np.asarray(lengths, dtype=np.int64)
    nk = np.maximum(lengths - k + 1, 0)
    total = int(nk.sum())
    seq = np.repeat(np.arange(len(lengths), dtype=np.int64), nk)
    first = np.zeros(len(lengths), dtype=np.int64)
    if len(lengths):
        np.cumsum(nk[:-1], out=first[1:])
    pos = np.arange(total, dtype=np.int64) - np.repeat(first, nk)
    start = np.repeat(offsets, nk) + pos
    codes = np.zeros(total, dtype=np.int64)
    valid = np.ones(total, dtype=bool)
    for j in range(k):
        r = flat[start + j].astype(np.int64)
        valid &= r < alpha
        codes = codes * alpha + np.minimum(r, alpha - 1)
    return codes[valid], seq[valid], pos[valid]


def build_kmer_index(flat_sequences, offsets, lengths, k: int, mode: str = "protein",
                     alpha_size: int | None = None):
    """(kmer_offsets, kmer_entries, kmer_freqs) as in kmer_index.py:117-162."""
    alpha = alpha_size or (20 if mode == "protein" else 4)
    num = alpha ** k
    codes, seq, pos = window_codes(flat_sequences, offsets, lengths, k, alpha)
    order = np.argsort(codes, kind="stable")
    counts = np.bincount(codes, minlength=num).astype(np.int64)
    kmer_offsets = np.zeros(num + 1, dtype=np.int64)
    np.cumsum(counts, out=kmer_offsets[1:])
    kmer_entries = (seq[order] << 32) | pos[order]
    return kmer_offsets, kmer_entries, counts.astype(np.int32)


def compute_freq_threshold(kmer_freqs, num_sequences: int, percentile: float = 99.5) -> int:
    """max(int(percentile of nonzero freqs), max(100, N // 200))."""
    nonzero = np.asarray(kmer_freqs)[np.asarray(kmer_freqs) > 0]
    if len(nonzero) == 0:
        return int(num_sequences)
    return int(max(int(np.percentile(nonzero, percentile)), max(100, num_sequences // 200)))
