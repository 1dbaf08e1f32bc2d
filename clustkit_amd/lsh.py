"""LSH candidate pairs from sketches (host, NumPy).

Follows clustkit/lsh.py:20-178 and clustkit/utils.py:99-141: per table, a
band of ``num_bands`` sketch positions and a seed are drawn from
``RandomState(seed)``; the band is hashed with the multiply-add chain and
fmix tail of lsh.py:20-29; sequences sharing a bucket id become pairs
(buckets capped at 1000 members); pairs are packed as ``i*N + j`` (i < j),
de-duplicated and unpacked in ascending order.

Within-bucket order uses a stable sort (the reference's default argsort is
not specified for ties; it only matters for buckets above the cap).  This is
input preparation for the alignment stage, not on the timed path.
"""

from __future__ import annotations

import math

import numpy as np

_M1 = np.uint64(0x517CC1B727220A95)
_C2 = np.uint64(0xC4CEB9FE1A85EC53)
_S33 = np.uint64(33)


def auto_lsh_params(threshold: float, sensitivity: str = "medium", k: int = 5) -> dict:
    """Tables/bands from the collision-probability rule of utils.py:99-141."""
    target_miss = {"low": 0.05, "medium": 0.01, "high": 0.001}.get(sensitivity, 0.01)
    jac = max(threshold ** k, 1e-9)
    b = 1 if jac < 0.15 else (2 if jac < 0.5 else 3)
    p = min(jac ** b, 1.0 - 1e-12)
    tables = math.ceil(math.log(target_miss) / math.log(1.0 - p))
    return {"num_tables": max(8, min(tables, 512)), "num_bands": b}


def band_hash(sketches: np.ndarray, band: np.ndarray, seed: int) -> np.ndarray:
    h = np.full(sketches.shape[0], np.uint64(seed), dtype=np.uint64)
    for idx in band:
        h = h * _M1 + sketches[:, idx]
    h = (h ^ (h >> _S33)) * _C2
    return h ^ (h >> _S33)


def _bucket_pairs(order: np.ndarray, ids_sorted: np.ndarray, n: int, max_bucket: int):
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
        lo = np.minimum(x, y)
        hi = np.maximum(x, y)
        out.append(lo * n + hi)
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def lsh_candidates(sketches: np.ndarray, num_tables: int, num_bands: int, seed: int = 42,
                   max_bucket: int = 1000) -> np.ndarray:
    """(M, 2) int32 candidate pairs (i < j), ascending (lsh.py:93-178)."""
    n, s = sketches.shape
    if n < 2:
        return np.empty((0, 2), dtype=np.int32)
    rng = np.random.RandomState(seed)
    bands = np.empty((num_tables, num_bands), dtype=np.int32)
    seeds = np.empty(num_tables, dtype=np.int64)
    for t in range(num_tables):
        bands[t] = rng.choice(s, size=num_bands, replace=False).astype(np.int32)
        seeds[t] = int(rng.randint(0, 2 ** 31))
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
