"""Bottom-s MinHash sketches of rolling k-mer codes (host, NumPy).

Produces the sketches the Jaccard pre-filter reads.  Semantics follow
clustkit/sketch.py:17-80:

* k-mer code = base-``alphabet_size`` number of the k residue codes (codes
  >= alphabet_size are folded in unchanged, as the reference does);
* hash = MurmurHash3 fmix64 of ``code ^ seed``;
* sketch = the ``sketch_size`` smallest hashes of the multiset, ascending,
  padded with 2**64-1 when the sequence has fewer k-mers;
* a sequence shorter than k gets an all-zero sketch.

This is input preparation for the alignment stage, vectorised over the flat
sequence buffer; it is not on the timed path.
"""

from __future__ import annotations

import numpy as np

U64_MAX = np.uint64(0xFFFFFFFFFFFFFFFF)
_C1 = np.uint64(0xFF51AFD7ED558CCD)
_C2 = np.uint64(0xC4CEB9FE1A85EC53)
_S33 = np.uint64(33)


def fmix64(x: np.ndarray, seed: int) -> np.ndarray:
    """MurmurHash3 64-bit finaliser of ``x ^ seed`` (sketch.py:17-26)."""
    h = np.asarray(x, dtype=np.uint64) ^ np.uint64(seed)
    h = h ^ (h >> _S33)
    h = h * _C1
    h = h ^ (h >> _S33)
    h = h * _C2
    h = h ^ (h >> _S33)
    return h


def kmer_codes(flat: np.ndarray, offsets: np.ndarray, lengths: np.ndarray, k: int,
               alphabet_size: int):
    """Rolling k-mer codes of every window that fits inside its sequence.

    Returns (codes uint64[T], seq_id int64[T]) in sequence order."""
    flat = np.asarray(flat, dtype=np.uint8)
    lengths = np.asarray(lengths, dtype=np.int64)
    offsets = np.asarray(offsets, dtype=np.int64)
    nk = np.maximum(lengths - k + 1, 0)
    total = int(nk.sum())
    seq_id = np.repeat(np.arange(len(lengths), dtype=np.int64), nk)
    start = np.repeat(offsets, nk)
    first = np.zeros(len(lengths), dtype=np.int64)
    if len(lengths):
        np.cumsum(nk[:-1], out=first[1:])
    pos = start + (np.arange(total, dtype=np.int64) - np.repeat(first, nk))
    base = np.uint64(alphabet_size)
    codes = np.zeros(total, dtype=np.uint64)
    for j in range(k):
        codes = codes * base + flat[pos + j].astype(np.uint64)
    return codes, seq_id


def compute_sketches(flat: np.ndarray, offsets: np.ndarray, lengths: np.ndarray, k: int,
                     sketch_size: int, mode: str = "protein", seed: int = 42,
                     chunk: int = 8192) -> np.ndarray:
    """(N, sketch_size) uint64 sketches (clustkit/sketch.py:137-183)."""
    alphabet_size = 20 if mode == "protein" else 4
    lengths = np.asarray(lengths, dtype=np.int64)
    n = len(lengths)
    out = np.full((n, sketch_size), U64_MAX, dtype=np.uint64)
    short = lengths < k
    out[short] = 0
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        codes, sid = kmer_codes(flat, offsets[lo:hi], lengths[lo:hi], k, alphabet_size)
        if codes.size == 0:
            continue
        h = fmix64(codes, seed)
        order = np.lexsort((h, sid))
        h = h[order]
        sid = sid[order]
        counts = np.bincount(sid, minlength=hi - lo)
        starts = np.zeros(hi - lo, dtype=np.int64)
        np.cumsum(counts[:-1], out=starts[1:])
        rank = np.arange(len(h), dtype=np.int64) - starts[sid]
        keep = rank < sketch_size
        out[lo + sid[keep], rank[keep]] = h[keep]
    return out
