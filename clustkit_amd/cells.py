"""Algorithmic DP cell counts of banded SW (the GCUPS numerator).

Cells of one pair = sum over rows i = 1..la of the in-matrix band width
min(lb, i + h + bw) - max(1, i + h - bw) + 1 (clipped at 0), with la <= lb,
the effective half band bw of clustkit/csrc/sw_align.c:55-58 (bw <= 0 or
max(len) <= 50 -> bw = max(len)) and the clamped, orientation-adjusted
diagonal hint h of sw_align.c:237-256.  These are the cells the reference
loop evaluates (sw_align.c:110-160).
"""

from __future__ import annotations

import numpy as np


def _sum_min_lin(n, c, cap):
    """sum_{i=1..n} min(cap, i + c) for int64 arrays (cap may be < 1 + c)."""
    t = np.clip(cap - c, 0, n)                 # rows where i + c <= cap
    return t * (t + 1) // 2 + t * c + (n - t) * cap


def _sum_max_lin(n, c, floor):
    """sum_{i=1..n} max(floor, i + c)."""
    u = np.clip(floor - c, 0, n)               # rows where i + c <= floor
    s_lin = (n * (n + 1) // 2 - u * (u + 1) // 2) + (n - u) * c
    return u * floor + s_lin


def _band_cells(la, lb, bw, h):
    """Cells of rows 1..la inside |j - i - h| <= bw and 1 <= j <= lb."""
    c1, c2 = h + bw, h - bw
    # rows with a non-empty band: i + h + bw >= 1 and i + h - bw <= lb
    i0 = np.maximum(1, 1 - c1)
    i1 = np.minimum(la, lb - c2)

    def prefix(n):
        n = np.maximum(n, 0)
        return _sum_min_lin(n, c1, lb) - _sum_max_lin(n, c2, 1) + n

    cells = np.where(i1 >= i0, prefix(i1) - prefix(i0 - 1), 0)
    cells[(la == 0)] = 0
    return cells


def pair_cells(pairs, lengths, band_width: int, hints=None) -> np.ndarray:
    pairs = np.asarray(pairs).reshape(-1, 2)
    lengths = np.asarray(lengths, dtype=np.int64)
    lx = lengths[pairs[:, 0]]
    ly = lengths[pairs[:, 1]]
    la = np.minimum(lx, ly)
    lb = np.maximum(lx, ly)
    h = np.zeros(len(pairs), np.int64) if hints is None else np.asarray(hints, np.int64).copy()
    h[(h > la) | (h < -la)] = 0
    h = np.where(lx <= ly, h, -h)
    bw = np.where((band_width <= 0) | (lb <= 50), lb, np.int64(band_width))
    return _band_cells(la, lb, bw, h)


def nw_pair_cells(pairs, lengths, band_width: int, threshold: float = 0.0) -> np.ndarray:
    """Cells of the banded global alignment of clustkit/pairwise.py:684-850:
    0 for pairs removed by the length-ratio or length-difference test, else
    rows 1..la of the adaptive half band min(bw, max(10, diff + 10)) (the
    full matrix when that is <= 0 or the longer sequence has <= 50
    residues).  Rows skipped by the early stop are counted (the figure is
    the work of a call with threshold 0, where no pair stops early)."""
    pairs = np.asarray(pairs).reshape(-1, 2)
    lengths = np.asarray(lengths, dtype=np.int64)
    lx = lengths[pairs[:, 0]]
    ly = lengths[pairs[:, 1]]
    la = np.minimum(lx, ly)
    lb = np.maximum(lx, ly)
    diff = lb - la
    t = np.float32(threshold)
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio_fail = (lb > 0) & (la.astype(np.float32) / lb.astype(np.float32) < t)
    removed = ratio_fail | (diff > band_width)
    ab = np.minimum(np.int64(band_width), np.maximum(10, diff + 10))
    bw = np.where((ab <= 0) | (lb <= 50), lb, ab)
    cells = _band_cells(la, lb, bw, np.zeros(len(pairs), np.int64))
    cells[removed] = 0
    return cells
