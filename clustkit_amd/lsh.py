"""LSH candidate pairs on the device (ck_lsh_candidates_device).

Signature and results of clustkit/lsh.py:93-178 ``lsh_candidates``: per
table a band of ``num_bands`` sketch positions and a seed drawn from
``RandomState(seed)`` (lsh.py:131-137, drawn here on the host exactly so);
sequences sharing a bucket id (lsh.py:20-29) become pairs, the first 1000
members of a bucket in ascending sequence order (the reference GPU path's
stable sort, lsh.py:258-261); pairs (i < j) de-duplicated, ascending.  The
work runs on the GPU whatever ``device`` names.

``auto_lsh_params`` is the table/band rule of clustkit/utils.py:99-141.
"""

from __future__ import annotations

import math

import numpy as np

from . import _lib
from .device import _stream_ptr, _torch

MAX_BUCKET = 1000


def auto_lsh_params(threshold: float, sensitivity: str = "medium", k: int = 5) -> dict:
    """Tables/bands from the collision-probability rule of utils.py:99-141."""
    target_miss = {"low": 0.05, "medium": 0.01, "high": 0.001}.get(sensitivity, 0.01)
    jac = max(threshold ** k, 1e-9)
    b = 1 if jac < 0.15 else (2 if jac < 0.5 else 3)
    p = min(jac ** b, 1.0 - 1e-12)
    tables = math.ceil(math.log(target_miss) / math.log(1.0 - p))
    return {"num_tables": max(8, min(tables, 512)), "num_bands": b}


def draw_tables(sketch_size: int, num_tables: int, num_bands: int, seed: int = 42):
    """(bands int32 (T, b), seeds int64 (T,)) as lsh.py:131-137 draws them."""
    rng = np.random.RandomState(seed)
    bands = np.empty((num_tables, num_bands), dtype=np.int32)
    seeds = np.empty(num_tables, dtype=np.int64)
    for t in range(num_tables):
        bands[t] = rng.choice(sketch_size, size=num_bands, replace=False).astype(np.int32)
        seeds[t] = int(rng.randint(0, 2 ** 31))
    return bands, seeds


def lsh_candidates_device(sketches, num_tables: int, num_bands: int, seed: int = 42,
                          max_bucket: int = MAX_BUCKET, stream=None):
    """Candidate pairs of device-resident sketches (torch int64 (N, s), the
    uint64 bit patterns): torch int32 (M, 2) on the same device."""
    torch = _torch()
    lib = _lib.require_gpu()
    n, s = int(sketches.shape[0]), int(sketches.shape[1])
    assert sketches.is_contiguous() and sketches.dtype.itemsize == 8
    bands, seeds = draw_tables(s, num_tables, num_bands, seed)
    cnt = np.zeros(1, np.int64)
    sp = _stream_ptr(stream)
    _lib.check(lib.ck_lsh_candidates_device(sketches.data_ptr(), n, s, bands.ctypes.data,
                                            seeds.ctypes.data, int(num_tables),
                                            int(num_bands), int(max_bucket), cnt.ctypes.data,
                                            sp))
    m = int(cnt[0])
    out = torch.empty((m, 2), dtype=torch.int32, device=sketches.device)
    _lib.check(lib.ck_lsh_fetch_pairs_device(out.data_ptr() if m else None, m, sp))
    return out


def lsh_candidates(sketches: np.ndarray, num_tables: int, num_bands: int, seed: int = 42,
                   device: str = "cpu") -> np.ndarray:
    """lsh.py:93-178: (M, 2) int32 candidate pairs (i < j), ascending."""
    torch = _torch()
    sk = np.ascontiguousarray(sketches, dtype=np.uint64)
    n = sk.shape[0]
    if n < 2:
        return np.empty((0, 2), dtype=np.int32)
    d = torch.from_numpy(sk.view(np.int64)).to("cuda:0")
    out = lsh_candidates_device(d, num_tables, num_bands, seed)
    torch.cuda.current_stream().synchronize()
    return out.cpu().numpy()
