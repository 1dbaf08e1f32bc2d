"""Bottom-s MinHash sketches on the device (ck_sketch_device).

Signature and results of clustkit/sketch.py:137-183 ``compute_sketches``:
(N, sketch_size) uint64, each row the sketch_size smallest MurmurHash3
fmix64(code ^ seed) over the sequence's k-mer windows (code in base 20 for
protein, 4 for nucleotide), ascending, padded with 2^64-1; zeros for a
sequence shorter than k.  The reference's CPU and GPU paths give the same
sketches, so ``device`` only names the reference branch; the work runs on
the GPU either way (there is no CPU path here).
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceSequences, _stream_ptr, _torch


def _alphabet(mode: str) -> int:
    return 20 if mode == "protein" else 4


def compute_sketches_device(seqs: DeviceSequences, k: int, sketch_size: int,
                            mode: str = "protein", seed: int = 42, stream=None):
    """Sketches of HBM-resident sequences: torch int64 (N, sketch_size) on the
    device holding the uint64 bit patterns (view as uint64 on the host)."""
    torch = _torch()
    lib = _lib.require_gpu()
    out = torch.empty((seqs.n, sketch_size), dtype=torch.int64, device=seqs.flat.device)
    if seqs.n == 0:
        return out
    _lib.check(lib.ck_sketch_device(seqs.flat.data_ptr(), seqs.offsets.data_ptr(),
                                    seqs.lengths.data_ptr(), seqs.n, int(k), int(sketch_size),
                                    _alphabet(mode), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                    out.data_ptr(), _stream_ptr(stream)))
    return out


def compute_sketches(encoded_sequences, lengths, k: int, sketch_size: int, mode: str,
                     seed: int = 42, device: str = "cpu", flat_sequences=None,
                     offsets=None) -> np.ndarray:
    """sketch.py:137-183: (N, sketch_size) uint64 sketches on the host."""
    torch = _torch()
    lengths = np.ascontiguousarray(lengths, dtype=np.int32)
    if flat_sequences is None or offsets is None:
        enc = np.asarray(encoded_sequences, dtype=np.uint8)
        offsets = np.zeros(len(lengths), np.int64)
        if len(lengths):
            np.cumsum(lengths[:-1].astype(np.int64), out=offsets[1:])
        flat_sequences = (np.concatenate([enc[i, :lengths[i]] for i in range(len(lengths))])
                          if len(lengths) else np.zeros(0, np.uint8))
    seqs = DeviceSequences.upload(flat_sequences, offsets, lengths)
    out = compute_sketches_device(seqs, k, sketch_size, mode, seed)
    torch.cuda.current_stream().synchronize()
    return out.cpu().numpy().view(np.uint64)
