"""The no-allocation SW entry point (ck_sw_align_device_noalloc) against
ck_sw_align_device, through the C-ABI: identical outputs on calls in a row,
each later one smaller, covering the batched kernel, the wide-score re-run
(scores of 32,700 or more) and the int32 kernel (shorter side above 8,191
rows), so every region of the fixed scratch array is used; sizes beyond its
fixed maximum are rejected before any device work."""

from __future__ import annotations

import numpy as np
import pytest

from clustkit_amd.scoring import BLOSUM62

pytestmark = pytest.mark.gpu

def _noalloc():
    from clustkit_amd import _lib
    lib = _lib.require_gpu()
    return lib, lib.ck_sw_align_device_noalloc


def _table(small_family):
    """The family table plus two 8,300-residue copies (int32 kernel) and a
    7,000-residue near-identical pair (wide-score pass)."""
    from clustkit_amd.scoring import pack_sequences
    fam = small_family
    rng = np.random.default_rng(99)
    seqs = [fam.flat[o:o + n] for o, n in zip(fam.offsets, fam.lengths)]
    big = rng.integers(0, 20, 8300).astype(np.uint8)
    mid = rng.integers(0, 20, 7000).astype(np.uint8)
    mid2 = mid.copy()
    mid2[rng.random(7000) < 0.02] = rng.integers(0, 20, 1)[0]
    seqs += [big, big.copy(), mid, mid2]
    return pack_sequences(seqs)


def test_noalloc_matches_device_form(small_family, oracle_mod):
    import torch

    from clustkit_amd import _lib
    from clustkit_amd.device import DeviceSequences, alloc_outputs, last_paths, sw_align_device
    lib, fn = _noalloc()
    flat, offsets, lengths = _table(small_family)
    n = len(lengths)
    seqs = DeviceSequences.upload(flat, offsets, lengths)
    rng = np.random.default_rng(5)
    sub = np.ascontiguousarray(BLOSUM62, np.int8).reshape(-1)
    specials = np.array([[n - 4, n - 3], [n - 2, n - 1]], np.int32)
    calls = [(60000, 100, True, 0), (5000, 30, False, _lib.CK_SW_LATENCY),
             (700, 50, False, _lib.CK_SW_PLAIN | _lib.CK_SW_LENGTH_RATIO |
              _lib.CK_SW_MASK_IDENTITY)]
    for m, band, with_hints, flags in calls:
        print(f"noalloc call: {m} pairs, band {band}, flags {flags:#x}", flush=True)
        p = rng.integers(0, n - 4, (m, 2)).astype(np.int32)
        p[: m // 2, 1] = np.minimum((p[: m // 2, 0] // 10) * 10 + rng.integers(0, 10, m // 2),
                                    n - 5)
        p[:2] = specials
        hints = rng.integers(-40, 41, m).astype(np.int32) if with_hints else None
        d_p = torch.from_numpy(p).cuda()
        d_h = None if hints is None else torch.from_numpy(hints).cuda()
        want = sw_align_device(seqs, d_p, band, BLOSUM62, hints=d_h, threshold=0.3,
                               flags=flags)
        torch.cuda.synchronize()
        got = alloc_outputs(m)
        for v in got.values():
            v.fill_(7)
        rc = fn(d_p.data_ptr(), m, seqs.flat.data_ptr(), seqs.offsets.data_ptr(),
                seqs.lengths.data_ptr(), seqs.max_len, band, sub.ctypes.data,
                None if d_h is None else d_h.data_ptr(), 0.3, flags, got["sims"].data_ptr(),
                got["scores"].data_ptr(), got["matches"].data_ptr(), got["mask"].data_ptr(),
                torch.cuda.current_stream().cuda_stream)
        _lib.check(rc)
        torch.cuda.synchronize()
        paths = last_paths()
        if band == 100:
            assert paths["wide"] >= 1 and paths["int32"] >= 1
        for k in ("scores", "matches", "mask"):
            assert torch.equal(got[k], want[k]), k
        assert torch.equal(got["sims"].view(torch.int32), want["sims"].view(torch.int32))
        o = oracle_mod.sw_batch(p[:2000], flat, offsets, lengths, band, BLOSUM62,
                                None if hints is None else hints[:2000])
        want_scores = o["scores"]
        if flags & _lib.CK_SW_LENGTH_RATIO:
            want_scores = np.where(oracle_mod.length_ratio_pass(p[:2000], lengths, 0.3),
                                   want_scores, 0)
        assert np.array_equal(got["scores"][:2000].cpu().numpy(), want_scores)


def test_noalloc_rejects_beyond_maximum():
    import torch
    lib, fn = _noalloc()
    buf = torch.zeros(64, dtype=torch.int32, device="cuda:0")
    ptr = buf.data_ptr()
    sub = np.ascontiguousarray(BLOSUM62, np.int8).reshape(-1)

    def call(m, max_len, band, flags=0):
        return fn(ptr, m, ptr, ptr, ptr, max_len, band, sub.ctypes.data, None, 0.0, flags,
                  ptr, ptr, ptr, ptr, torch.cuda.current_stream().cuda_stream)

    assert call((1 << 20) + 1, 100, 100) != 0
    assert b"at most" in lib.ck_last_error()
    assert call(10, 5000, 0) != 0                      # full-matrix rows of 10,002 entries
    assert b"no-allocation workspace" in lib.ck_last_error()
    assert call(10, 100, 5000) != 0                    # band 5000
    assert call(10, 100, 100, 0x10) != 0               # CK_SW_TIMING
    assert call(0, 100, 100) == 0
