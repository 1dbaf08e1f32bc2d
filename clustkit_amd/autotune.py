"""Autotuner of the batched SW kernel's launch configuration.

The knobs change how the work is laid out on the GPU, never the results:

* band tile / pairs per wavefront: a pair's band is swept by a segment of
  ``lanes`` lanes -- stripes of ``lanes`` rows by the band's positions --
  so a wavefront holds 64 / lanes pairs (16 lanes: 4 pairs, short-ramp
  8 lanes: 8 pairs; narrow bands favour 8);
* bin boundaries: the device work queue orders pairs by estimated kernel
  steps in bins of ``2**bin_shift`` steps (coarse bins: less sorting work,
  fine bins: closer cost matching inside a wavefront and a tighter tail);
* schedule: one stripe at a time, or stripe-pipelined (each lane walks its
  stripes back to back, hiding the per-stripe ramp).

``tune`` times every candidate on a random sample of the caller's
device-resident pairs and keeps the fastest; the choice is cached per
workload signature (band and pair-length quantiles) in memory and, when a
path is given, in a JSON file.  ``apply`` sets a configuration
process-wide (``ck_sw_set_config``).

The sample measures throughput only if no pair in it sets its duration:
each pair slot of the GPU (CUs x 64 pairs in flight) should carry at least
twice the cost of the sample's longest pair, otherwise the kernels are
ranked by how they run a few long pairs.  On the heavy-tailed workload at
band 16 (2.13 M pairs, 10% of them 1,000-8,000 residues) a 200,000-pair
sample made one stripe at a time look faster (5.05 against 5.78 ms; on the
whole set 41.5 against 20.1 ms), and even 520,000 pairs preferred 8-lane
segments one stripe at a time (8.76 against 8.88 ms; 28.3 against 20.1 ms
on the whole set).  So up to 4 M pairs the whole set is timed (a few
seconds once per workload signature); beyond that a random sample,
``sample_size`` growing it until the slot rule holds.
"""

from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np

from . import _lib
from .scoring import BLOSUM62

# (lanes per pair, bin shift, pipelined)
CANDIDATES = [(16, 5, 1), (16, 5, 0), (8, 5, 1), (8, 5, 0), (16, 3, 1), (16, 7, 1), (8, 3, 1),
              (8, 7, 1)]

_memo: dict[str, dict] = {}


def current() -> tuple[int, int, int]:
    import ctypes
    lib = _lib.load()
    a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _lib.check(lib.ck_sw_get_config(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return int(a.value), int(b.value), int(c.value)


def apply(lanes: int, bin_shift: int, pipelined: int = 1) -> None:
    _lib.check(_lib.load().ck_sw_set_config(int(lanes), int(bin_shift), int(pipelined)))


def signature(pair_lengths: np.ndarray, band: int) -> str:
    """Workload key: band and the 10/50/90/99% quantiles of the shorter and
    longer length of the pairs, each rounded to a power of two."""
    pl = np.asarray(pair_lengths)
    if pl.size == 0:
        return f"b{band}:empty"
    la, lb = pl.min(axis=1), pl.max(axis=1)
    q = lambda x: [int(2 ** np.round(np.log2(max(1.0, v))))  # noqa: E731
                   for v in np.quantile(x, [0.1, 0.5, 0.9, 0.99])]
    return f"b{band}:a{q(la)}:b{q(lb)}"


def pair_costs(pair_lengths: np.ndarray, band: int) -> np.ndarray:
    """Estimated kernel steps of each pair: stripes of 16 rows of the shorter
    sequence times the band's positions (the whole longer sequence for
    band <= 0 or a longer side of at most 50, as the kernel does)."""
    pl = np.asarray(pair_lengths, np.int64).reshape(-1, 2)
    la, lb = pl.min(axis=1), pl.max(axis=1)
    bw = np.where((band <= 0) | (lb <= 50), lb, band)
    return ((la + 15) // 16) * (2 * bw + 1)


def sample_size(pair_lengths: np.ndarray, band: int, slots: int, m: int) -> int:
    """Pairs to time, from a random sample's lengths: the sample's size grown
    until its cost per pair slot (total / ``slots``) is at least twice its
    longest pair's cost (capped at the ``m`` pairs there are)."""
    cost = pair_costs(pair_lengths, band)
    n = int(cost.size)
    if n == 0 or n >= m:
        return n
    per_slot = cost.sum() / max(1, slots)
    need = 2.0 * float(cost.max())
    if per_slot >= need:
        return n
    return int(min(m, np.ceil(n * need / per_slot)))


def tune(seqs, pairs, band: int, sub_matrix=BLOSUM62, hints=None, sample: int = 4_000_000,
         repeats: int = 2, candidates=CANDIDATES, cache_path: str | Path | None = None,
         seed: int = 0, stream=None) -> dict:
    """Pick the fastest configuration for ``pairs`` (torch int32 [M, 2] on
    the device) against ``seqs`` (a DeviceSequences).  Applies it and returns
    ``{"lanes", "bin_shift", "pipelined", "signature", "ms": {cand: ms},
    "cached"}``."""
    import torch

    from .device import alloc_outputs, sw_align_device
    m = int(pairs.shape[0])
    rng = np.random.default_rng(seed)
    slots = _lib.device_props(int(pairs.device.index or 0))["num_cu"] * 64

    def draw(n):
        if m > n:
            idx = torch.from_numpy(np.sort(rng.choice(m, size=n, replace=False))).to(
                pairs.device)
            return pairs[idx].contiguous(), (hints[idx].contiguous()
                                             if hints is not None else None)
        return pairs, hints

    sp, sh = draw(sample)
    lens = seqs.lengths[sp.long()].cpu().numpy()
    grown = sample_size(lens, band, slots, m)
    if grown > sp.shape[0]:
        sp, sh = draw(grown)
        lens = seqs.lengths[sp.long()].cpu().numpy()
    sig = signature(lens, band)
    cache = {}
    if cache_path is not None and Path(cache_path).exists():
        cache = json.loads(Path(cache_path).read_text())
    hit = _memo.get(sig) or cache.get(sig)
    if hit is not None:
        apply(hit["lanes"], hit["bin_shift"], hit.get("pipelined", 0))
        return dict(hit, signature=sig, cached=True)
    out = alloc_outputs(int(sp.shape[0]), sp.device)
    s = stream if stream is not None else torch.cuda.current_stream()
    times = {}
    saved = current()
    try:
        for lanes, shift, pipe in candidates:
            apply(lanes, shift, pipe)
            sw_align_device(seqs, sp, band, sub_matrix, hints=sh, out=out, stream=s)  # warm
            best = float("inf")
            for _ in range(repeats):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(s)
                sw_align_device(seqs, sp, band, sub_matrix, hints=sh, out=out, stream=s)
                e1.record(s)
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            times[f"{lanes}{'p' if pipe else ''}x{shift}"] = round(best, 4)
    except Exception:
        apply(*saved)
        raise
    lanes, shift, pipe = min(candidates,
                             key=lambda c: times[f"{c[0]}{'p' if c[2] else ''}x{c[1]}"])
    apply(lanes, shift, pipe)
    res = {"lanes": lanes, "bin_shift": shift, "pipelined": pipe, "ms": times,
           "sample_pairs": int(sp.shape[0]),
           "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
    _memo[sig] = res
    if cache_path is not None:
        cache[sig] = res
        Path(cache_path).write_text(json.dumps(cache, indent=1, sort_keys=True))
    return dict(res, signature=sig, cached=False)
