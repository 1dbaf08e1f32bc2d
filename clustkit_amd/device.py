"""HBM-resident sequence sets and device-pointer calls into the C-ABI.

PyTorch is used only to own device memory and streams: the arrays handed to
the library are plain device pointers (``tensor.data_ptr()``) and the stream
is the raw ``hipStream_t`` of a torch stream.

Layout in HBM (the compact layout of clustkit/io.py:35-131):
  flat     uint8[sum(len)]   residue codes of all sequences, concatenated
  offsets  int64[N]          start of sequence i in ``flat``
  lengths  int32[N]          length of sequence i
  pairs    int32[M, 2]       candidate pairs (row-major)
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .scoring import BLOSUM62


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.HipLibraryError("no GPU visible to torch: HBM-resident data needs the MI355X")
    return torch


@dataclass
class DeviceSequences:
    flat: "object"
    offsets: "object"
    lengths: "object"
    max_len: int
    n: int

    @classmethod
    def upload(cls, flat, offsets, lengths, device: str = "cuda:0") -> "DeviceSequences":
        torch = _torch()
        flat = np.ascontiguousarray(flat, dtype=np.uint8)
        if flat.size == 0:
            flat = np.zeros(1, np.uint8)
        lengths = np.ascontiguousarray(lengths, dtype=np.int32)
        return cls(flat=torch.from_numpy(flat).to(device),
                   offsets=torch.from_numpy(np.ascontiguousarray(offsets, np.int64)).to(device),
                   lengths=torch.from_numpy(lengths).to(device),
                   max_len=int(lengths.max()) if lengths.size else 0, n=len(lengths))


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def alloc_outputs(m: int, device: str = "cuda:0"):
    torch = _torch()
    return {"sims": torch.empty(m, dtype=torch.float32, device=device),
            "scores": torch.empty(m, dtype=torch.int32, device=device),
            "matches": torch.empty(m, dtype=torch.int32, device=device),
            "mask": torch.empty(m, dtype=torch.uint8, device=device)}


def sw_align_device(seqs: DeviceSequences, pairs, band_width: int, sub_matrix=BLOSUM62,
                    hints=None, threshold: float = 0.0, flags: int = 0, out=None,
                    stream=None):
    """Enqueue banded SW for device-resident ``pairs`` (torch int32 [M,2]) on
    ``stream`` (default: torch's current stream).  Returns the output dict
    (torch tensors, valid once the stream has run)."""
    lib = _lib.require_gpu()
    m = int(pairs.shape[0])
    if out is None:
        out = alloc_outputs(m, pairs.device)
    if m == 0:
        return out
    assert pairs.is_contiguous() and pairs.dtype.itemsize == 4
    sub = np.ascontiguousarray(sub_matrix, np.int8).reshape(-1)
    _lib.check(lib.ck_sw_align_device(
        pairs.data_ptr(), m, seqs.flat.data_ptr(), seqs.offsets.data_ptr(),
        seqs.lengths.data_ptr(), seqs.max_len, int(band_width), sub.ctypes.data,
        None if hints is None else hints.data_ptr(), float(np.float32(threshold)), int(flags),
        out["sims"].data_ptr(), out["scores"].data_ptr(), out["matches"].data_ptr(),
        out["mask"].data_ptr(), _stream_ptr(stream)))
    return out


def nw_align_device(seqs: DeviceSequences, pairs, band_width: int, sub_matrix=None,
                    threshold: float = 0.0, flags: int = 0, out=None, stream=None):
    """Enqueue banded global alignment (NW identity, ck_nw_align_device) for
    device-resident ``pairs``; sub_matrix None = the plain +2/-1 scoring.
    Returns the output dict (sims, scores, mask)."""
    lib = _lib.require_gpu()
    m = int(pairs.shape[0])
    if out is None:
        out = alloc_outputs(m, pairs.device)
    if m == 0:
        return out
    assert pairs.is_contiguous() and pairs.dtype.itemsize == 4
    sub = None if sub_matrix is None else np.ascontiguousarray(sub_matrix, np.int8).reshape(-1)
    _lib.check(lib.ck_nw_align_device(
        pairs.data_ptr(), m, seqs.flat.data_ptr(), seqs.offsets.data_ptr(),
        seqs.lengths.data_ptr(), seqs.max_len, int(band_width),
        None if sub is None else sub.ctypes.data, float(np.float32(threshold)), int(flags),
        out["sims"].data_ptr(), out["scores"].data_ptr(), out["mask"].data_ptr(),
        _stream_ptr(stream)))
    return out


def graph_csr_device(pairs, sims, num_sequences: int, indptr=None, indices=None, data=None,
                     nnz=None, stream=None):
    """Enqueue the symmetric CSR of device-resident ``pairs`` (int32 [K,2])
    and ``sims`` (float32 [K]) (ck_graph_csr_device).  Returns (indptr,
    indices, data, nnz) torch tensors; entries past nnz are unused."""
    torch = _torch()
    lib = _lib.require_gpu()
    k = int(pairs.shape[0])
    dev = pairs.device
    n = int(num_sequences)
    if indptr is None:
        indptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
    if indices is None:
        indices = torch.empty(max(1, 2 * k), dtype=torch.int32, device=dev)
    if data is None:
        data = torch.empty(max(1, 2 * k), dtype=torch.float32, device=dev)
    if nnz is None:
        nnz = torch.zeros(1, dtype=torch.int64, device=dev)
    _lib.check(lib.ck_graph_csr_device(pairs.data_ptr(), k, sims.data_ptr(), n,
                                       indptr.data_ptr(), indices.data_ptr(), data.data_ptr(),
                                       nnz.data_ptr(), _stream_ptr(stream)))
    return indptr, indices, data, nnz


def last_stats():
    lib = _lib.load()
    a, b, c = (np.zeros(1, np.int64) for _ in range(3))
    _lib.check(lib.ck_sw_last_stats(a.ctypes.data, b.ctypes.data, c.ctypes.data))
    return {"batched": int(a[0]), "rerun": int(b[0]), "skipped": int(c[0])}


def last_paths():
    """How the last call's re-run pairs were finished: wide-score wavefront
    pass and int32 one-pair-per-thread kernel."""
    lib = _lib.load()
    a, b = (np.zeros(1, np.int64) for _ in range(2))
    _lib.check(lib.ck_sw_last_paths(a.ctypes.data, b.ctypes.data))
    return {"wide": int(a[0]), "int32": int(b[0])}


def last_status():
    """Raise when the last SW / NW device call on this thread met a pair its
    one-pair-per-thread workspace could not hold (``max_len`` below a pair's
    longer sequence with band_width <= 0); waits for that call's counters."""
    _lib.check(_lib.load().ck_sw_last_status())


def last_timing():
    """(queue_ms, main_ms, rerun_ms) of the last call made with CK_SW_TIMING."""
    lib = _lib.load()
    v = np.zeros(3, np.float32)
    _lib.check(lib.ck_sw_last_timing(v[0:].ctypes.data, v[1:].ctypes.data, v[2:].ctypes.data))
    return float(v[0]), float(v[1]), float(v[2])


def jaccard_device(pairs, sketches, out=None, stream=None):
    torch = _torch()
    lib = _lib.require_gpu()
    m = int(pairs.shape[0])
    if out is None:
        out = torch.empty(m, dtype=torch.float32, device=pairs.device)
    if m:
        _lib.check(lib.ck_jaccard_device(pairs.data_ptr(), m, sketches.data_ptr(),
                                         int(sketches.shape[1]), out.data_ptr(),
                                         _stream_ptr(stream)))
    return out
