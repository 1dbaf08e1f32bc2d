"""The C-ABI as integrators bind it, on the GPU.

* ``ck_batch_sw_align`` bound with ctypes exactly as INTEGRATION.md shows
  for clustkit/pairwise.py:_load_c_sw (pairwise.py:42-66), called the way
  _batch_sw_c calls batch_sw_align_c (pairwise.py:69-108), against the
  golden vectors written by the reference kernel; its null-pointer checks.
* Host-pointer entry points (ck_sw_align_host, ck_jaccard_host,
  ck_nw_align_host) called from three host threads at once with changing
  sizes: every result equals the oracle's.
* A device-form call whose ``max_len`` is below a pair's length (band 0,
  full matrix) reports the overflow through ck_sw_last_status instead of
  returning scores silently.
"""

from __future__ import annotations

import ctypes
import threading

import numpy as np
import pytest

from clustkit_amd.scoring import BLOSUM62
from conftest import golden_files, load_golden

pytestmark = pytest.mark.gpu


def _reference_side_binding():
    """The stub of INTEGRATION.md section 2, verbatim in effect."""
    from clustkit_amd import _lib
    _lib.require_gpu()            # loads torch's HIP runtime first (one runtime per process)
    lib = ctypes.cdll.LoadLibrary(str(_lib.LIB_PATH))
    lib.ck_last_error.restype = ctypes.c_char_p
    lib.ck_batch_sw_align.restype = ctypes.c_int
    lib.ck_batch_sw_align.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,   # pairs, flat, offsets
        ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,     # lengths, M, band_width
        ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,    # sub_matrix, threshold, diag_hints
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,   # out_sims, out_scores, out_mask
    ]

    def _check(rc):
        if rc != 0:
            raise RuntimeError(lib.ck_last_error())

    lib.batch_sw_align_c = lambda *a: _check(lib.ck_batch_sw_align(*a))
    lib.sw_set_num_threads = lambda n: None
    return lib


def _batch_sw_c_via_binding(lib, pairs, flat_sequences, offsets, lengths, threshold,
                            band_width, sub_matrix, hints=None):
    """The body of the reference _batch_sw_c (pairwise.py:69-108) on the
    rebound symbol, with the diagonal hints the search path passes."""
    lib.sw_set_num_threads(4)
    M = len(pairs)
    pf = np.ascontiguousarray(pairs.flatten().astype(np.int32))
    fl = np.ascontiguousarray(flat_sequences)
    off = offsets.astype(np.int64)
    lens = lengths.astype(np.int32)
    sm = sub_matrix.astype(np.int8)
    h = None if hints is None else np.ascontiguousarray(hints, np.int32)
    c_sims = np.empty(M, dtype=np.float32)
    c_scores = np.empty(M, dtype=np.int32)
    c_mask = np.empty(M, dtype=np.uint8)
    lib.batch_sw_align_c(pf.ctypes.data, fl.ctypes.data, off.ctypes.data, lens.ctypes.data, M,
                         band_width, sm.ctypes.data, np.float32(threshold),
                         None if h is None else h.ctypes.data,
                         c_sims.ctypes.data, c_scores.ctypes.data, c_mask.ctypes.data)
    return c_sims, c_scores, c_mask, c_sims >= threshold


@pytest.mark.parametrize("path", golden_files(), ids=lambda p: p.stem)
def test_gpu_ck_batch_sw_align_reference_binding(path):
    g = load_golden(path)
    print(f"ck_batch_sw_align {g['name']}", flush=True)
    lib = _reference_side_binding()
    flat = g["flat"] if g["flat"].size else np.zeros(1, np.uint8)
    sims, scores, mask, keep = _batch_sw_c_via_binding(
        lib, g["pairs"].reshape(-1, 2), flat, g["offsets"], g["lengths"], 0.4,
        int(g["band"]), BLOSUM62, g["hints"])
    assert np.array_equal(scores, g["scores"])
    assert np.array_equal(sims.view(np.uint32), g["sims"].view(np.uint32))
    assert np.array_equal(mask, g["mask"])
    assert np.array_equal(keep, g["sims"] >= np.float32(0.4))


def test_gpu_ck_batch_sw_align_null_pointers():
    lib = _reference_side_binding()
    pairs = np.array([0, 1], np.int32)
    flat = np.arange(20, dtype=np.uint8)
    off = np.array([0, 10], np.int64)
    lens = np.array([10, 10], np.int32)
    sub = np.ascontiguousarray(BLOSUM62, np.int8)
    outs = [np.zeros(1, np.float32), np.zeros(1, np.int32), np.zeros(1, np.uint8)]
    args = [pairs.ctypes.data, flat.ctypes.data, off.ctypes.data, lens.ctypes.data, 1, 10,
            sub.ctypes.data, 0.0, None] + [o.ctypes.data for o in outs]
    assert lib.ck_batch_sw_align(*args) == 0
    for i in (1, 6, 9, 10, 11):                 # flat, sub_matrix, each output
        bad = list(args)
        bad[i] = None
        assert lib.ck_batch_sw_align(*bad) != 0, i
        assert b"null" in lib.ck_last_error()


def test_gpu_host_forms_from_concurrent_threads(small_family, oracle_mod):
    """Three host threads call the SW, Jaccard and NW host forms at once,
    eight times each with growing and shrinking sizes (so the shared staging
    buffers are reallocated between calls); each result is compared with the
    oracle's for the same inputs."""
    from clustkit_amd.pairwise import _batch_jaccard, batch_nw_align, batch_sw_align
    fam = small_family
    n = len(fam.lengths)
    rng = np.random.default_rng(17)
    sizes = [5000, 300, 12000, 40, 7000, 1, 20000, 900]
    sw_in, nw_in, jc_in = [], [], []
    sk = np.sort(rng.integers(0, 2000, (n, 64)).astype(np.uint64), axis=1)
    for m in sizes:
        p = rng.integers(0, n, (m, 2)).astype(np.int32)
        p[: m // 2, 1] = (p[: m // 2, 0] // 10) * 10 + rng.integers(0, 10, m // 2)
        o = oracle_mod.sw_batch(p, fam.flat, fam.offsets, fam.lengths, 100, BLOSUM62)
        sw_in.append((p, o))
        q = p[: max(1, m // 4)]
        nw_in.append((q, oracle_mod.nw_batch(q, fam.flat, fam.offsets, fam.lengths, 0.3, 50,
                                             BLOSUM62)))
        jc_in.append((p, oracle_mod.jaccard_batch(p, sk)))
    errors = []

    def run(name, fn):
        try:
            fn()
        except BaseException as e:  # noqa: BLE001 - reported below
            errors.append(f"{name}: {e!r}")

    def sw():
        for p, o in sw_in:
            sims, scores, matches, mask = batch_sw_align(p, fam.flat, fam.offsets, fam.lengths,
                                                         100)
            assert np.array_equal(scores, o["scores"]) and np.array_equal(matches, o["matches"])
            assert np.array_equal(sims.view(np.uint32), o["sims"].view(np.uint32))

    def jc():
        for p, want in jc_in:
            got = _batch_jaccard(p, sk, 0.0)[0]
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))

    def nw():
        for q, (ws, wsc, wm) in nw_in:
            sims, scores, mask = batch_nw_align(q, fam.flat, fam.offsets, fam.lengths, 0.3, 50,
                                                BLOSUM62)
            assert np.array_equal(scores, wsc) and np.array_equal(mask, wm)
            assert np.array_equal(sims.view(np.uint32), ws.view(np.uint32))

    ts = [threading.Thread(target=run, args=(nm, f)) for nm, f in
          (("sw", sw), ("jaccard", jc), ("nw", nw))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts)
    assert not errors, errors


def test_gpu_sw_device_max_len_too_small_is_reported(small_family, oracle_mod):
    """band 0 (full matrix) sizes the one-pair-per-thread workspace from
    max_len; a max_len below a pair's longer sequence must surface as an
    error (ck_sw_last_status), and the next correct call is clean."""
    import torch

    from clustkit_amd import _lib
    from clustkit_amd.device import DeviceSequences, last_status, sw_align_device
    fam = small_family
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths)
    pairs = np.array([[0, 1], [2, 3], [4, 14]], np.int32)
    d_pairs = torch.from_numpy(pairs).cuda()
    short = DeviceSequences(seqs.flat, seqs.offsets, seqs.lengths, max_len=10, n=seqs.n)
    for flags in (0, _lib.CK_SW_PLAIN):
        sw_align_device(short, d_pairs, 0, BLOSUM62, flags=flags)
        torch.cuda.synchronize()
        with pytest.raises(_lib.HipLibraryError, match="max_len"):
            last_status()
        out = sw_align_device(seqs, d_pairs, 0, BLOSUM62, flags=flags)
        torch.cuda.synchronize()
        last_status()
        o = oracle_mod.sw_batch(pairs, fam.flat, fam.offsets, fam.lengths, 0, BLOSUM62)
        assert np.array_equal(out["scores"].cpu().numpy(), o["scores"])
