"""MinHash sketches and LSH candidate pairs on the CPU -- TEST INFRASTRUCTURE
ONLY (checker of ck_sketch_device / ck_lsh_candidates_device).

Only tests/ and bench.py's CPU comparison import this module.  Two forms:

* vectorised NumPy (``compute_sketches``, ``lsh_candidates``): the whole
  table at once, for sizes of tens of thousands of sequences;
* literal loops (``sketch_one_loop``, ``extract_pairs_loop``) that follow
  the reference's per-sequence online algorithm and its pair extraction
  step by step, for small inputs -- a second, independent statement the
  vectorised form is checked against.

Semantics (clustkit/sketch.py:17-80, 137-183; clustkit/lsh.py:20-178):

* k-mer code = base-``alphabet_size`` number of the k residue codes, first
  residue most significant, mod 2^64 (codes >= alphabet_size enter
  unchanged, as in the reference's rolling code, sketch.py:46-53, 75-78);
* hash = MurmurHash3 fmix64 of ``code ^ seed`` (sketch.py:17-26);
* sketch = the ``sketch_size`` smallest hashes of the multiset, ascending,
  padded with 2^64-1; a sequence shorter than k gets an all-zero sketch;
* LSH: per table a band of ``num_bands`` positions and a seed drawn from
  ``RandomState(seed)`` (lsh.py:131-137); bucket id = multiply-add chain and
  fmix tail (lsh.py:20-29); the first min(size, 1000) members of each bucket,
  in ascending sequence order, form all pairs, packed as i*N + j (i < j),
  de-duplicated and unpacked in ascending order (lsh.py:52-90, 161-178).
  Ascending member order is the order of the reference's GPU path (a stable
  device argsort, lsh.py:258-261); its CPU path uses NumPy's default
  (unstable) argsort, which only matters for buckets above the cap.
"""

from __future__ import annotations

import numpy as np

U64_MAX = np.uint64(0xFFFFFFFFFFFFFFFF)
_C1 = np.uint64(0xFF51AFD7ED558CCD)
_C2 = np.uint64(0xC4CEB9FE1A85EC53)
_M1 = np.uint64(0x517CC1B727220A95)
_S33 = np.uint64(33)
_MASK = (1 << 64) - 1


def fmix64(x: np.ndarray, seed: int) -> np.ndarray:
    """MurmurHash3 64-bit finaliser of ``x ^ seed`` (sketch.py:17-26)."""
    h = np.asarray(x, dtype=np.uint64) ^ np.uint64(seed)
    h = h ^ (h >> _S33)
    h = h * _C1
    h = h ^ (h >> _S33)
    h = h * _C2
    h = h ^ (h >> _S33)
    return h


def kmer_codes(flat, offsets, lengths, k: int, alphabet_size: int):
    """Codes of every window that fits inside its sequence: (codes uint64[T],
    seq_id int64[T]) in sequence order."""
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
    with np.errstate(over="ignore"):
        for j in range(k):
            codes = codes * base + flat[pos + j].astype(np.uint64)
    return codes, seq_id


def compute_sketches(flat, offsets, lengths, k: int, sketch_size: int, mode: str = "protein",
                     seed: int = 42, chunk: int = 8192) -> np.ndarray:
    """(N, sketch_size) uint64 sketches (sketch.py:137-183), vectorised."""
    alphabet_size = 20 if mode == "protein" else 4
    lengths = np.asarray(lengths, dtype=np.int64)
    offsets = np.asarray(offsets, dtype=np.int64)
    n = len(lengths)
    out = np.full((n, sketch_size), U64_MAX, dtype=np.uint64)
    out[lengths < k] = 0
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        codes, sid = kmer_codes(flat, offsets[lo:hi], lengths[lo:hi], k, alphabet_size)
        if codes.size == 0:
            continue
        with np.errstate(over="ignore"):
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


def _fmix_int(h: int) -> int:
    h ^= h >> 33
    h = (h * 0xFF51AFD7ED558CCD) & _MASK
    h ^= h >> 33
    h = (h * 0xC4CEB9FE1A85EC53) & _MASK
    h ^= h >> 33
    return h


def sketch_one_loop(seq, k: int, sketch_size: int, alphabet_size: int, seed: int) -> list:
    """sketch.py:29-80 step by step: rolling code, online bottom-s buffer
    with replacement of the current maximum, final sort."""
    seq = [int(x) for x in seq]
    L = len(seq)
    if L < k:
        return [0] * sketch_size
    num = L - k + 1
    out = [_MASK] * sketch_size
    buf_count, buf_max, buf_max_pos = 0, 0, 0
    place = 1
    for _ in range(k - 1):
        place = (place * alphabet_size) & _MASK
    val = 0
    for j in range(k):
        val = (val * alphabet_size + seq[j]) & _MASK
    for i in range(num):
        h = _fmix_int(val ^ seed)
        if buf_count < sketch_size:
            out[buf_count] = h
            if buf_count == 0 or h > buf_max:
                buf_max, buf_max_pos = h, buf_count
            buf_count += 1
        elif h < buf_max:
            out[buf_max_pos] = h
            buf_max, buf_max_pos = out[0], 0
            for s in range(1, sketch_size):
                if out[s] > buf_max:
                    buf_max, buf_max_pos = out[s], s
        if i + 1 < num:
            val = (((val - seq[i] * place) & _MASK) * alphabet_size + seq[i + k]) & _MASK
    return sorted(out)


def draw_tables(sketch_size: int, num_tables: int, num_bands: int, seed: int = 42):
    """Band positions and seeds per table, drawn as lsh.py:131-137 does."""
    rng = np.random.RandomState(seed)
    bands = np.empty((num_tables, num_bands), dtype=np.int32)
    seeds = np.empty(num_tables, dtype=np.int64)
    for t in range(num_tables):
        bands[t] = rng.choice(sketch_size, size=num_bands, replace=False).astype(np.int32)
        seeds[t] = int(rng.randint(0, 2 ** 31))
    return bands, seeds


def band_hash(sketches: np.ndarray, band: np.ndarray, seed: int) -> np.ndarray:
    """Bucket ids of every row for one table (lsh.py:20-29)."""
    h = np.full(sketches.shape[0], np.uint64(seed), dtype=np.uint64)
    with np.errstate(over="ignore"):
        for idx in band:
            h = h * _M1 + sketches[:, idx]
        h = (h ^ (h >> _S33)) * _C2
    return h ^ (h >> _S33)


def _bucket_pairs(order, ids_sorted, n: int, max_bucket: int):
    if len(ids_sorted) < 2:
        return np.zeros(0, np.int64)
    brk = np.flatnonzero(ids_sorted[1:] != ids_sorted[:-1]) + 1
    starts = np.concatenate(([0], brk))
    sizes = np.diff(np.concatenate((starts, [len(ids_sorted)])))
    sizes = np.minimum(sizes, max_bucket)
    out = []
    for s in np.unique(sizes[sizes >= 2]):
        st = starts[sizes == s]
        members = order[st[:, None] + np.arange(s)[None, :]].astype(np.int64)
        iu, ju = np.triu_indices(int(s), 1)
        x = members[:, iu].ravel()
        y = members[:, ju].ravel()
        out.append(np.minimum(x, y) * n + np.maximum(x, y))
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def lsh_candidates(sketches: np.ndarray, num_tables: int, num_bands: int, seed: int = 42,
                   max_bucket: int = 1000) -> np.ndarray:
    """(M, 2) int32 candidate pairs (i < j), ascending (lsh.py:93-178)."""
    n, s = sketches.shape
    if n < 2:
        return np.empty((0, 2), dtype=np.int32)
    bands, seeds = draw_tables(s, num_tables, num_bands, seed)
    packed = []
    for t in range(num_tables):
        ids = band_hash(sketches, bands[t], int(seeds[t]))
        order = np.argsort(ids, kind="stable")
        p = _bucket_pairs(order, ids[order], n, max_bucket)
        if len(p):
            packed.append(np.unique(p))
    if not packed:
        return np.empty((0, 2), dtype=np.int32)
    u = np.unique(np.concatenate(packed))
    return np.stack([(u // n).astype(np.int32), (u % n).astype(np.int32)], axis=1)


def extract_pairs_loop(order, sorted_ids, n: int, max_bucket: int) -> list:
    """lsh.py:52-90 step by step: runs of equal ids, the first max_bucket
    members of each, every pair packed as lo * n + hi."""
    out = []
    i = 0
    while i < n:
        j = i + 1
        while j < n and sorted_ids[j] == sorted_ids[i]:
            j += 1
        end = min(j, i + max_bucket)
        for a in range(i, end):
            for b in range(a + 1, end):
                x, y = int(order[a]), int(order[b])
                if x > y:
                    x, y = y, x
                out.append(x * n + y)
        i = j
    return out
