"""k-mer index search on the MI355X: drop-in for the reference's
``batch_score_queries_c`` (clustkit/csrc/kmer_score.c:420-472) and for the
standard path of ``search_kmer_index`` (clustkit/kmer_index.py:1475-2349):
k-mer Phase A/B scoring -> banded SW on the candidates (or banded global
identity with ``local_alignment=False``) -> E-value-style rank -> top-k hits
per query.

The optional recall add-ons of search_kmer_index (BLOSUM-similar k-mers,
IDF weighting, reduced alphabets, spaced seeds, ungapped and ML pre-filters)
are not part of this build; asking for them raises ``NotImplementedError``.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .kmer_index import build_kmer_index, compute_freq_threshold
from .scoring import BLOSUM62


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------------------
# Result types (clustkit/search.py:39-60)
# ---------------------------------------------------------------------------

@dataclass
class SearchHit:
    query_id: str
    target_id: str
    identity: float
    query_length: int
    target_length: int
    score: float = 0.0


@dataclass
class SearchResults:
    hits: list
    num_queries: int
    num_targets: int
    num_candidates: int
    num_aligned: int
    runtime_seconds: float


@dataclass
class SequenceSet:
    """Compact sequence storage (the flat/offsets/lengths part of
    clustkit/io.py:SequenceDataset)."""
    flat_sequences: np.ndarray
    offsets: np.ndarray
    lengths: np.ndarray
    ids: list = field(default_factory=list)

    @property
    def num_sequences(self) -> int:
        return len(self.lengths)

    @classmethod
    def from_arrays(cls, flat, offsets, lengths, ids=None):
        n = len(lengths)
        return cls(_c(flat, np.uint8), _c(offsets, np.int64), _c(lengths, np.int32),
                   list(ids) if ids is not None else [f"s{i}" for i in range(n)])


@dataclass
class DatabaseIndex:
    """The arrays of clustkit/database.py:19-39 that the k-mer search reads,
    plus the sketch / LSH arrays a saved database carries (optional here)."""
    dataset: SequenceSet
    params: dict
    kmer_offsets: np.ndarray
    kmer_entries: np.ndarray
    kmer_freqs: np.ndarray
    kmer_entries_compact: np.ndarray | None = None
    sketches: np.ndarray | None = None
    lsh_bucket_ids: np.ndarray | None = None
    lsh_band_indices: np.ndarray | None = None
    lsh_seeds: np.ndarray | None = None
    freq_thresholds: dict | None = None


# File layout of clustkit/database.py:223-300 (save_database) and 304-442
# (load_database): one .npy per array, ids and params as JSON.
_DB_ARRAYS = ("sketches", "lsh_bucket_ids", "lsh_band_indices", "lsh_seeds")


def save_database(db_index: DatabaseIndex, output_dir) -> None:
    """database.py:223-300: flat_sequences / offsets / lengths, the k-mer
    index (offsets, entries, freqs, compact ids), sketches and LSH arrays
    when present, freq_thresholds.json, ids.json, params.json.  The
    reduced-alphabet indices of the reference's full build are not part of
    this build and are not written."""
    import json
    from pathlib import Path
    out = Path(output_dir)
    out.mkdir(parents=True, exist_ok=True)
    ds = db_index.dataset
    np.save(out / "flat_sequences.npy", ds.flat_sequences)
    np.save(out / "offsets.npy", ds.offsets)
    np.save(out / "lengths.npy", ds.lengths)
    for name in _DB_ARRAYS:
        arr = getattr(db_index, name)
        if arr is not None:
            np.save(out / f"{name}.npy", arr)
    if db_index.kmer_offsets is not None:
        np.save(out / "kmer_offsets.npy", db_index.kmer_offsets)
        np.save(out / "kmer_entries.npy", db_index.kmer_entries)
        np.save(out / "kmer_freqs.npy", db_index.kmer_freqs)
        if db_index.kmer_entries_compact is not None:
            np.save(out / "kmer_entries_compact.npy", db_index.kmer_entries_compact)
    if db_index.freq_thresholds:
        with open(out / "freq_thresholds.json", "w") as f:
            json.dump(db_index.freq_thresholds, f)
    with open(out / "ids.json", "w") as f:
        json.dump(list(ds.ids), f)
    params = dict(db_index.params)
    params.setdefault("num_sequences", ds.num_sequences)
    params.setdefault("max_length", int(ds.lengths.max()) if ds.num_sequences else 0)
    with open(out / "params.json", "w") as f:
        json.dump(params, f, indent=2)


def load_database(db_dir, mmap: bool = True) -> DatabaseIndex:
    """database.py:304-442 for the standard search path: the sequence table
    (compact, or the padded encoded_sequences.npy converted to it), the
    k-mer index, sketches / LSH arrays and thresholds when present.  Arrays
    are memory-mapped (``mmap``) so a large database is read once, on its
    way to HBM.  Loaded with allow_pickle off (plain .npy arrays only)."""
    import json
    from pathlib import Path
    d = Path(db_dir)
    mode = "r" if mmap else None

    def load(name, required=True):
        f = d / f"{name}.npy"
        if not f.exists():
            if required:
                raise FileNotFoundError(f)
            return None
        return np.load(f, mmap_mode=mode, allow_pickle=False)

    with open(d / "params.json") as f:
        params = json.load(f)
    lengths = np.ascontiguousarray(load("lengths"), np.int32)
    flat, offsets = load("flat_sequences", False), load("offsets", False)
    if flat is None or offsets is None:
        enc = load("encoded_sequences")
        offsets = np.zeros(len(lengths), np.int64)
        if len(lengths):
            np.cumsum(lengths[:-1].astype(np.int64), out=offsets[1:])
        flat = np.concatenate([np.asarray(enc[k, :lengths[k]]) for k in range(len(lengths))]) \
            if len(lengths) else np.zeros(0, np.uint8)
    with open(d / "ids.json") as f:
        ids = json.load(f)
    ft = None
    if (d / "freq_thresholds.json").exists():
        with open(d / "freq_thresholds.json") as f:
            ft = json.load(f)
    ds = SequenceSet(flat_sequences=flat, offsets=offsets, lengths=lengths, ids=ids)
    return DatabaseIndex(dataset=ds, params=params, kmer_offsets=load("kmer_offsets", False),
                         kmer_entries=load("kmer_entries", False),
                         kmer_freqs=load("kmer_freqs", False),
                         kmer_entries_compact=load("kmer_entries_compact", False),
                         sketches=load("sketches", False),
                         lsh_bucket_ids=load("lsh_bucket_ids", False),
                         lsh_band_indices=load("lsh_band_indices", False),
                         lsh_seeds=load("lsh_seeds", False), freq_thresholds=ft)


def build_database_index(dataset: SequenceSet, mode: str = "protein", k: int | None = None):
    """k-mer part of clustkit/database.py:build_database (k = 3 for protein)."""
    kk = k if k is not None else (3 if mode == "protein" else 5)
    ko, ke, kf = build_kmer_index(dataset.flat_sequences, dataset.offsets, dataset.lengths, kk,
                                  mode)
    return DatabaseIndex(dataset=dataset, params={"mode": mode, "kmer_size": kk,
                                                  "kmer_index_k": kk},
                         kmer_offsets=ko, kmer_entries=ke, kmer_freqs=kf,
                         kmer_entries_compact=(ke >> 32).astype(np.int32))


# ---------------------------------------------------------------------------
# Stage 1: k-mer scoring
# ---------------------------------------------------------------------------

def batch_score_queries(q_flat, q_offsets, q_lengths, k, alpha_size, kmer_offsets, kmer_entries,
                        compact_entries, kmer_freqs, freq_thresh, num_db, min_total_hits=2,
                        min_diag_hits=2, diag_bin_width=10, max_cands=2000, phase_a_topk=10000):
    """Host-array form of batch_score_queries_c.  Returns (out_targets
    (nq, max_cands), out_counts, out_scores, out_diags)."""
    lib = _lib.require_gpu()
    ql = _c(q_lengths, np.int32)
    nq = len(ql)
    qf = _c(q_flat, np.uint8)
    if qf.size == 0:
        qf = np.zeros(1, np.uint8)
    qo = _c(q_offsets, np.int64)
    ko = _c(kmer_offsets, np.int64)
    ke = _c(kmer_entries, np.int64)
    kf = _c(kmer_freqs, np.int32)
    if ke.size == 0:
        ke = np.zeros(1, np.int64)
    tg = np.zeros((nq, max_cands), np.int32)
    cn = np.zeros(nq, np.int32)
    sc = np.zeros((nq, max_cands), np.int32)
    dg = np.zeros((nq, max_cands), np.int32)
    if nq:
        _lib.check(lib.ck_batch_score_queries(
            qf.ctypes.data, qo.ctypes.data, ql.ctypes.data, nq, int(k), int(alpha_size),
            ko.ctypes.data, ke.ctypes.data, None, kf.ctypes.data, int(freq_thresh), int(num_db),
            int(min_total_hits), int(min_diag_hits), int(diag_bin_width), int(max_cands),
            int(phase_a_topk), tg.ctypes.data, cn.ctypes.data, sc.ctypes.data, dg.ctypes.data))
    return tg, cn, sc, dg


def batch_score_queries_similar(q_flat, q_offsets, q_lengths, kmer_offsets, kmer_entries,
                                kmer_freqs, freq_thresh, num_db, min_total_hits, max_cands,
                                sub_matrix, kmer_score_thresh, db_lengths=None):
    """Host-array form of _batch_score_queries_similar (kmer_index.py:1300-1330):
    BLOSUM-similar 5-mer scoring on the GPU (ck_kmer_score_similar_device).
    Returns (out_targets (nq, max_cands), out_counts)."""
    import torch
    lib = _lib.require_gpu()
    dev = torch.device("cuda", torch.cuda.current_device())
    t = lambda a: _to_device(a, dev)  # noqa: E731
    ql = _c(q_lengths, np.int32)
    nq = len(ql)
    tg = np.zeros((nq, max_cands), np.int32)
    cn = np.zeros(nq, np.int32)
    if nq == 0 or num_db == 0:
        return tg, cn
    if len(kmer_offsets) != 20 ** 5 + 1:
        raise ValueError("similar k-mer scoring needs a k = 5 protein index")
    qf = _c(q_flat, np.uint8)
    ke = _c(kmer_entries, np.int64)
    tid = (ke >> 32).astype(np.int32) if ke.size else np.zeros(1, np.int32)
    dl = (_c(db_lengths, np.int32) if db_lengths is not None
          else np.zeros(max(int(num_db), 1), np.int32))
    sub = _c(sub_matrix, np.int8).reshape(-1)
    if sub.size != 400:
        raise ValueError("sub_matrix must be 20x20")
    d_qf, d_qo, d_ql = t(qf if qf.size else np.zeros(1, np.uint8)), t(_c(q_offsets, np.int64)), t(ql)
    d_ko, d_tid, d_kf, d_dl = (t(_c(kmer_offsets, np.int64)), t(tid),
                               t(_c(kmer_freqs, np.int32)), t(dl))
    out = [torch.empty((nq, max_cands), dtype=torch.int32, device=dev),
           torch.empty(nq, dtype=torch.int32, device=dev),
           torch.empty((nq, max_cands), dtype=torch.int32, device=dev),
           torch.empty((nq, max_cands), dtype=torch.int32, device=dev)]
    s = torch.cuda.current_stream()
    _lib.check(lib.ck_kmer_score_similar_device(
        d_qf.data_ptr(), int(qf.size), d_qo.data_ptr(), d_ql.data_ptr(), nq, int(ql.max()),
        d_ko.data_ptr(), d_tid.data_ptr(), d_kf.data_ptr(), d_dl.data_ptr(), int(num_db),
        int(dl.max(initial=0)), sub.ctypes.data, int(kmer_score_thresh), int(freq_thresh),
        int(min_total_hits), int(max_cands), out[0].data_ptr(), out[1].data_ptr(),
        out[2].data_ptr(), out[3].data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    return out[0].cpu().numpy(), out[1].cpu().numpy()


def tile_offsets(kmer_offsets, kmer_entries, num_db: int) -> np.ndarray:
    """Per-(k-mer, 8192-target tile) posting offsets (ck_kmer_tile_offsets_host),
    vectorised: entries are sorted by (code, target) so one searchsorted over
    code * 2**32 + tile_start finds every boundary."""
    ko = np.asarray(kmer_offsets, np.int64)
    ncodes = len(ko) - 1
    nt = (num_db + 8191) // 8192
    key = (np.repeat(np.arange(ncodes, dtype=np.int64), np.diff(ko)) << 32) | \
        (np.asarray(kmer_entries, np.int64) >> 32)
    bounds = (np.arange(ncodes, dtype=np.int64)[:, None] << 32) + \
        (np.arange(nt + 1, dtype=np.int64) * 8192)[None, :]
    out = np.searchsorted(key, bounds.ravel(), side="left").reshape(ncodes, nt + 1)
    out[:, nt] = ko[1:]
    return out.astype(np.int64)


def _to_device(a, device, chunk_bytes: int = 256 << 20):
    """Host array -> device tensor.  Read-only arrays (memory-mapped .npy
    files) are copied through bounded host chunks, so a large database
    never needs a second full copy in host memory."""
    import torch
    a = np.asarray(a)
    if a.flags.writeable and a.flags.c_contiguous:
        return torch.from_numpy(a).to(device)
    out = torch.empty(a.shape, dtype=torch.from_numpy(np.zeros(0, a.dtype)).dtype, device=device)
    flat_in = a.reshape(-1)
    flat_out = out.reshape(-1)
    step = max(1, chunk_bytes // max(a.itemsize, 1))
    for i in range(0, flat_in.size, step):
        flat_out[i:i + step].copy_(torch.from_numpy(np.array(flat_in[i:i + step])))
    return out


class DeviceKmerIndex:
    """A database and its k-mer index resident in HBM (torch tensors).

    ``DeviceKmerIndex(db)`` uploads a host-built index;
    ``DeviceKmerIndex.build(dataset, k)`` builds it on the device from the
    sequence table (ck_kmer_index_build_device), identical array for array."""

    def __init__(self, db: DatabaseIndex, device="cuda:0", tiled: bool | None = None):
        import torch
        ko = _c(db.kmer_offsets, np.int64)
        ke = _c(db.kmer_entries, np.int64)
        t = lambda a: _to_device(a, device)  # noqa: E731
        tid = t((ke >> 32).astype(np.int32)) if len(ke) else t(np.zeros(1, np.int32))
        pos = t((ke & 0xFFFFFFFF).astype(np.int32)) if len(ke) else t(np.zeros(1, np.int32))
        self._setup(db, device, t(ko), tid, pos, t(_c(db.kmer_freqs, np.int32)), len(ke), tiled,
                    host_entries=ke)

    @classmethod
    def from_disk(cls, db_dir, device="cuda:0", tiled: bool | None = None) -> "DeviceKmerIndex":
        """A database saved in the reference's layout (database.py:223-300),
        loaded into HBM: the arrays are memory-mapped and copied to the
        device; a database saved without its k-mer index gets one built on
        the device from its sequence table."""
        db = load_database(db_dir)
        if db.kmer_offsets is None:
            k = int(db.params.get("kmer_index_k", 3 if db.params["mode"] == "protein" else
                                  db.params["kmer_size"]))
            return cls.build(db.dataset, k=k, mode=db.params["mode"], device=device, tiled=tiled)
        return cls(db, device=device, tiled=tiled)

    @classmethod
    def build(cls, dataset: SequenceSet, k: int | None = None, mode: str = "protein",
              device="cuda:0", tiled: bool | None = None, stream=None) -> "DeviceKmerIndex":
        """Upload the sequence table and build its k-mer index on the device
        (clustkit/kmer_index.py:117-162 layout)."""
        import ctypes

        import torch
        lib = _lib.require_gpu()
        kk = k if k is not None else (3 if mode == "protein" else 5)
        alpha = 20 if mode == "protein" else 4
        ncodes = alpha ** kk
        t = lambda a: _to_device(a, device)  # noqa: E731
        flat = t(dataset.flat_sequences if len(dataset.flat_sequences) else np.zeros(1, np.uint8))
        offs = t(_c(dataset.offsets, np.int64))
        lens = t(_c(dataset.lengths, np.int32))
        cap = int(np.maximum(dataset.lengths.astype(np.int64) - kk + 1, 0).sum())
        ko = torch.empty(ncodes + 1, dtype=torch.int64, device=device)
        kf = torch.empty(ncodes, dtype=torch.int32, device=device)
        tid = torch.empty(max(cap, 1), dtype=torch.int32, device=device)
        pos = torch.empty(max(cap, 1), dtype=torch.int32, device=device)
        total = ctypes.c_int64(0)
        s = stream if stream is not None else torch.cuda.current_stream()
        _lib.check(lib.ck_kmer_index_build_device(
            flat.data_ptr(), offs.data_ptr(), lens.data_ptr(), dataset.num_sequences, kk, alpha,
            cap, ko.data_ptr(), kf.data_ptr(), tid.data_ptr(), pos.data_ptr(),
            ctypes.byref(total), s.cuda_stream))
        n = int(total.value)
        if n < cap:
            tid, pos = tid[:max(n, 1)].clone(), pos[:max(n, 1)].clone()
        db = DatabaseIndex(dataset=dataset,
                           params={"mode": mode, "kmer_size": kk, "kmer_index_k": kk},
                           kmer_offsets=ko.cpu().numpy(), kmer_entries=None,
                           kmer_freqs=kf.cpu().numpy(), kmer_entries_compact=None)
        obj = cls.__new__(cls)
        obj._setup(db, device, ko, tid, pos, kf, n, tiled, uploaded=(flat, offs, lens),
                   stream=s)
        return obj

    def _setup(self, db, device, ko, tid, pos, kf, n_post, tiled, host_entries=None,
               uploaded=None, stream=None):
        import torch
        self.db = db
        self.k = int(db.params.get("kmer_index_k", db.params["kmer_size"]))
        self.alpha = 20 if db.params["mode"] == "protein" else 4
        ds = db.dataset
        self.num_db = ds.num_sequences
        self.max_db_len = int(ds.lengths.max()) if self.num_db else 0
        ncodes = int(ko.shape[0]) - 1
        nt = (self.num_db + 8191) // 8192
        if tiled is None:
            tiled = self.num_db > 0 and ncodes * (nt + 1) <= (32 << 20) and n_post >= ncodes * 64
        self.tiled = bool(tiled)
        self.device = device
        self.n_postings = int(n_post)
        self.kmer_offsets, self.post_tid, self.post_pos, self.kmer_freqs = ko, tid, pos, kf
        t = lambda a: _to_device(a, device)  # noqa: E731
        self.tile_off = None
        if self.tiled:
            if host_entries is not None:
                self.tile_off = t(tile_offsets(db.kmer_offsets, host_entries, self.num_db))
            else:
                lib = _lib.require_gpu()
                self.tile_off = torch.empty((ncodes, nt + 1), dtype=torch.int64, device=device)
                s = stream if stream is not None else torch.cuda.current_stream()
                _lib.check(lib.ck_kmer_tile_offsets_device(
                    ko.data_ptr(), tid.data_ptr(), ncodes, self.num_db, self.tile_off.data_ptr(),
                    s.cuda_stream))
        if uploaded is not None:
            self.db_flat, self.db_offsets, self.db_lengths = uploaded
        else:
            self.db_flat = t(ds.flat_sequences if len(ds.flat_sequences) else
                             np.zeros(1, np.uint8))
            self.db_offsets = t(ds.offsets)
            self.db_lengths = t(ds.lengths)
        self.db_codes = self.db_code_off = None
        if self.num_db and self.k <= 5 and self.alpha ** self.k <= self.CODE_SPACE_MAX:
            self._build_codes(ds.lengths, stream)

    # Code spaces up to this size keep a per-window code array of the
    # database on the device for the target-major Phase B.
    CODE_SPACE_MAX = 8192

    def _build_codes(self, lengths, stream=None):
        """The database's window codes (ck_kmer_db_codes_device): code + 1 per
        window in 16 bits (0: not scored), each sequence's run padded with 0
        to a multiple of 8."""
        import torch
        lib = _lib.require_gpu()
        nw = np.maximum(np.asarray(lengths, np.int64) - self.k + 1, 0)
        off = np.zeros(len(nw) + 1, np.int64)
        np.cumsum((nw + 7) // 8 * 8, out=off[1:])
        self.db_code_off = _to_device(off, self.device)
        self.db_codes = torch.empty(max(int(off[-1]), 8), dtype=torch.int16, device=self.device)
        s = stream if stream is not None else torch.cuda.current_stream()
        _lib.check(lib.ck_kmer_db_codes_device(
            self.db_flat.data_ptr(), self.db_offsets.data_ptr(), self.db_lengths.data_ptr(),
            self.num_db, self.k, self.alpha, self.db_code_off.data_ptr(),
            self.db_codes.data_ptr(), s.cuda_stream))


    # Largest length product the device rank table serves: below 2^24 every
    # product of two lengths is an exact float32 integer, so the table index
    # is the value NumPy takes the log of.
    LOG_TABLE_MAX = 1 << 24

    def log_table(self, max_product: int):
        """Device float32 table t[p] = ln p as NumPy computes it for float32
        (p = 0 .. max_product), grown as needed; None when max_product is
        outside the exact range."""
        import torch
        if max_product >= self.LOG_TABLE_MAX:
            return None
        cur = getattr(self, "_log_table", None)
        if cur is None or cur.numel() <= max_product:
            n = 1 << max(10, int(max_product).bit_length())
            n = min(n, self.LOG_TABLE_MAX)
            with np.errstate(divide="ignore"):
                tab = np.log(np.arange(n, dtype=np.float32))
            self._log_table = torch.from_numpy(tab).to(self.device)
        return self._log_table
    def host_entries(self) -> np.ndarray:
        """kmer_entries in the reference's int64 ``seq << 32 | pos`` form."""
        n = self.n_postings
        tid = self.post_tid[:n].cpu().numpy().astype(np.int64)
        pos = self.post_pos[:n].cpu().numpy().astype(np.int64)
        return (tid << 32) | pos

    def hbm_bytes(self) -> int:
        ts = [self.kmer_offsets, self.post_tid, self.post_pos, self.kmer_freqs, self.db_flat,
              self.db_offsets, self.db_lengths] + ([self.tile_off] if self.tiled else []) + (
              [self.db_codes, self.db_code_off] if self.db_codes is not None else [])
        return int(sum(x.numel() * x.element_size() for x in ts))

    def score(self, q_flat, q_off, q_len, max_qlen: int, freq_thresh: int, out=None,
              stream=None, min_total_hits=2, min_diag_hits=2, diag_bin_width=10,
              max_cands=2000, phase_a_topk=10000, target_major=True, window_codes=True):
        """Enqueue Phase A/B scoring of device-resident queries; returns
        (targets, counts, scores, diags) torch tensors.  ``target_major``
        lets Phase B walk the survivors' windows (dense code spaces) rather
        than re-read the query's posting lists, reading the database's
        precomputed window codes when ``window_codes`` (else its residues);
        all forms give the same result."""
        import torch
        lib = _lib.require_gpu()
        nq = int(q_len.shape[0])
        if out is None:
            out = (torch.empty((nq, max_cands), dtype=torch.int32, device=self.device),
                   torch.empty(nq, dtype=torch.int32, device=self.device),
                   torch.empty((nq, max_cands), dtype=torch.int32, device=self.device),
                   torch.empty((nq, max_cands), dtype=torch.int32, device=self.device))
        s = stream if stream is not None else torch.cuda.current_stream()
        use_codes = bool(target_major and window_codes and self.db_codes is not None)
        _lib.check(lib.ck_kmer_score_device(
            q_flat.data_ptr(), int(q_flat.numel()), q_off.data_ptr(), q_len.data_ptr(), nq,
            int(max_qlen), self.k, self.alpha, self.kmer_offsets.data_ptr(),
            self.post_tid.data_ptr(), self.post_pos.data_ptr(), self.kmer_freqs.data_ptr(),
            self.tile_off.data_ptr() if self.tiled else None, self.db_lengths.data_ptr(),
            self.db_flat.data_ptr() if target_major else None,
            int(self.db_flat.numel()) if target_major else 0,
            self.db_offsets.data_ptr() if target_major else None,
            self.db_codes.data_ptr() if use_codes else None,
            self.db_code_off.data_ptr() if use_codes else None,
            self.num_db, self.max_db_len, int(freq_thresh), int(min_total_hits),
            int(min_diag_hits), int(diag_bin_width), int(max_cands), int(phase_a_topk),
            out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(), out[3].data_ptr(),
            s.cuda_stream))
        return out

    def score_similar(self, q_flat, q_off, q_len, max_qlen: int, freq_thresh: int,
                      kmer_score_thresh: int = 13, out=None, stream=None, min_total_hits=2,
                      max_cands=2000, sub_matrix=None):
        """Enqueue BLOSUM-similar stage-1 scoring (kmer_index.py:1186-1330;
        k = 5 protein indexes): every 5-mer scoring at least
        ``kmer_score_thresh`` against a query 5-mer adds its score to the
        targets of its postings.  Returns (targets, counts, scores, diags)
        torch tensors as ``score`` does (diags 0)."""
        import torch
        if self.k != 5 or self.alpha != 20:
            raise NotImplementedError("similar k-mer scoring is defined for k = 5 protein "
                                      "indexes (kmer_index.py:1206-1207)")
        lib = _lib.require_gpu()
        nq = int(q_len.shape[0])
        if out is None:
            out = (torch.empty((nq, max_cands), dtype=torch.int32, device=self.device),
                   torch.empty(nq, dtype=torch.int32, device=self.device),
                   torch.empty((nq, max_cands), dtype=torch.int32, device=self.device),
                   torch.empty((nq, max_cands), dtype=torch.int32, device=self.device))
        sub = _c(BLOSUM62 if sub_matrix is None else sub_matrix, np.int8).reshape(-1)
        if sub.size != 400:
            raise ValueError("sub_matrix must be 20x20")
        s = stream if stream is not None else torch.cuda.current_stream()
        _lib.check(lib.ck_kmer_score_similar_device(
            q_flat.data_ptr(), int(q_flat.numel()), q_off.data_ptr(), q_len.data_ptr(), nq,
            int(max_qlen), self.kmer_offsets.data_ptr(), self.post_tid.data_ptr(),
            self.kmer_freqs.data_ptr(), self.db_lengths.data_ptr(), self.num_db,
            self.max_db_len, sub.ctypes.data, int(kmer_score_thresh), int(freq_thresh),
            int(min_total_hits), int(max_cands), out[0].data_ptr(), out[1].data_ptr(),
            out[2].data_ptr(), out[3].data_ptr(), s.cuda_stream))
        return out


# ---------------------------------------------------------------------------
# Stage 2/3 helpers (clustkit/search.py:402-479, 727-800; kmer_index.py:2318-2322)
# ---------------------------------------------------------------------------

def _merge_sequences_for_alignment(query: SequenceSet, db: SequenceSet):
    """Queries first (0..nq-1), then database sequences (search.py:402-463)."""
    flat = np.concatenate([query.flat_sequences, db.flat_sequences])
    offsets = np.concatenate([query.offsets, db.offsets + len(query.flat_sequences)])
    return {"lengths": np.concatenate([query.lengths, db.lengths]).astype(np.int32),
            "nq": query.num_sequences, "nd": db.num_sequences,
            "flat_sequences": flat, "offsets": offsets.astype(np.int64)}


def _remap_pairs_to_merged(pairs: np.ndarray, nq_offset: int) -> np.ndarray:
    out = np.empty_like(pairs)
    out[:, 0] = pairs[:, 0]
    out[:, 1] = pairs[:, 1] + np.int32(nq_offset)
    return out


def evalue_rank(scores, q_lengths, t_lengths):
    """0.267 * S - ln(qlen * tlen) in float32 (kmer_index.py:2318-2322)."""
    lam = np.float32(0.267)
    return lam * np.asarray(scores, np.float32) - np.log(
        np.asarray(q_lengths, np.float32) * np.asarray(t_lengths, np.float32))


def _top_k_order(qi, rank, top_k):
    """Indices of the first ``top_k`` entries per query after a stable sort
    by (query, -rank) -- the selection of search.py:727-800, vectorised."""
    order = np.lexsort((-rank, qi))
    sq = qi[order]
    n = len(sq)
    if n == 0:
        return order
    start = np.r_[True, sq[1:] != sq[:-1]]
    first = np.maximum.accumulate(np.where(start, np.arange(n), 0))
    return order[(np.arange(n) - first) < top_k]


def _hits_from_arrays(qi, ti, ident, rank, nq, query: SequenceSet, db: SequenceSet):
    """Per-query SearchHit lists from hit arrays grouped by query, best first."""
    hits = [[] for _ in range(nq)]
    qids, dids = query.ids, db.ids
    for q, d, idn, sc in zip(qi.tolist(), ti.tolist(), ident.tolist(), rank.tolist()):
        hits[q].append(SearchHit(query_id=qids[q], target_id=dids[d], identity=float(idn),
                                 query_length=int(query.lengths[q]),
                                 target_length=int(db.lengths[d]), score=float(sc)))
    return hits


def _collect_top_k_hits(passing_pairs, passing_sims, nq, top_k, query: SequenceSet,
                        db: SequenceSet, passing_scores=None):
    """search.py:727-800: stable lexsort by (query, -rank), first top_k per query."""
    hits = [[] for _ in range(nq)]
    if len(passing_pairs) == 0:
        return hits
    qi = passing_pairs[:, 0]
    di = passing_pairs[:, 1]
    rank = passing_scores if passing_scores is not None else passing_sims
    sel = _top_k_order(qi, rank, top_k)
    qids, dids = query.ids, db.ids
    for q, d, idn, sc in zip(qi[sel].tolist(), di[sel].tolist(), passing_sims[sel].tolist(),
                             rank[sel].tolist()):
        hits[q].append(SearchHit(query_id=qids[q], target_id=dids[d], identity=float(idn),
                                 query_length=int(query.lengths[q]),
                                 target_length=int(db.lengths[d]), score=float(sc)))
    return hits


def auto_band_width(query: SequenceSet, db: SequenceSet) -> int:
    """max(20, int(p95(len) * 0.3)) over queries and database (kmer_index.py:2155-2160)."""
    all_len = np.concatenate([query.lengths, db.lengths])
    return max(20, int(int(np.percentile(all_len, 95)) * 0.3))


_UNSUPPORTED = ("reduced_alphabet", "use_idf", "spaced_seeds", "extra_alphabets",
                "ml_prefilter_model")


def search_kmer_index(db_index: DatabaseIndex, query_dataset: SequenceSet,
                      threshold: float = 0.5, top_k: int = 10, band_width: int | None = None,
                      device: str = "0", min_total_hits: int = 2, min_diag_hits: int = 2,
                      diag_bin_width: int = 10, max_cands_per_query: int = 2000,
                      phase_a_topk: int = 10000, freq_percentile: float = 95.0,
                      min_ungapped_score: int = 0, kmer_score_thresh: int = 13,
                      local_alignment: bool = True, evalue_normalize: bool = True,
                      use_c_sw: bool = False, c_sw_band_width: int = 50,
                      use_c_scoring: bool = False, **unsupported):
    """Standard path of kmer_index.py:1475-2349 on the GPU.

    Stage 1 uses the C scoring rules (kmer_score.c) on device; for k = 5
    protein indexes with kmer_score_thresh > 0 (the reference's default) the
    BLOSUM-similar 5-mer scoring of kmer_index.py:1186-1330 (no diagonal
    hints then, as in the reference).
    Stage 2 is banded SW (mask score > 0): with ``use_c_sw`` the band is
    ``c_sw_band_width``, steered by the Phase-B diagonal hints when
    ``use_c_scoring`` (the reference has hints only from its C scorer,
    kmer_index.py:1627-1653, 1714-1736, 2255-2256); otherwise ``band_width``
    (auto when None) without hints.  With ``local_alignment=False`` (or nucleotide data) it
    is banded global identity (mask identity >= threshold): BLOSUM62 for
    proteins, ranked by the global score; +2/-1 otherwise, ranked by
    identity.  SW ranking: 0.267*S - ln(qlen*tlen) (float32) when
    ``evalue_normalize``, the score otherwise."""
    import time
    from .pairwise import _batch_align_compact, _batch_align_compact_scored, batch_sw_align
    for key in _UNSUPPORTED:
        if unsupported.get(key):
            raise NotImplementedError(f"{key} is not part of the MI355X search path")
    if min_ungapped_score > 0:
        raise NotImplementedError("the ungapped pre-filter is not part of this build")
    t0 = time.perf_counter()
    nq, nd = query_dataset.num_sequences, db_index.dataset.num_sequences
    mode = db_index.params["mode"]
    k = int(db_index.params.get("kmer_index_k", db_index.params["kmer_size"]))
    alpha = 20 if mode == "protein" else 4
    use_similar = k >= 5 and mode == "protein" and kmer_score_thresh > 0
    if use_similar and k != 5:
        raise NotImplementedError("similar k-mer scoring is defined for k = 5 "
                                  "(kmer_index.py:1206-1207)")
    empty = SearchResults([[] for _ in range(nq)], nq, nd, 0, 0, 0.0)
    if nq == 0 or nd == 0:
        empty.runtime_seconds = time.perf_counter() - t0
        return empty
    ft = compute_freq_threshold(db_index.kmer_freqs, nd, freq_percentile)
    if use_similar:
        tg, cn = batch_score_queries_similar(
            query_dataset.flat_sequences, query_dataset.offsets, query_dataset.lengths,
            db_index.kmer_offsets, db_index.kmer_entries, db_index.kmer_freqs, ft, nd,
            min_total_hits, max_cands_per_query, BLOSUM62, kmer_score_thresh,
            db_lengths=db_index.dataset.lengths)
        dg = np.zeros_like(tg)
    else:
        tg, cn, sc, dg = batch_score_queries(
            query_dataset.flat_sequences, query_dataset.offsets, query_dataset.lengths, k,
            alpha, db_index.kmer_offsets, db_index.kmer_entries, db_index.kmer_entries_compact,
            db_index.kmer_freqs, ft, nd, min_total_hits, min_diag_hits, diag_bin_width,
            max_cands_per_query, phase_a_topk)
    total = int(cn.sum())
    if total == 0:
        empty.runtime_seconds = time.perf_counter() - t0
        return empty
    rows = np.repeat(np.arange(nq, dtype=np.int32), cn)
    cols = np.concatenate([tg[q, :cn[q]] for q in range(nq)]).astype(np.int32)
    diags = np.concatenate([dg[q, :cn[q]] for q in range(nq)]).astype(np.int32)
    candidate_pairs = np.stack([rows, cols], axis=1)
    if band_width is None:
        band_width = auto_band_width(query_dataset, db_index.dataset)
    merged = _merge_sequences_for_alignment(query_dataset, db_index.dataset)
    merged_pairs = _remap_pairs_to_merged(candidate_pairs, merged["nq"])
    order = np.argsort(np.minimum(merged_pairs[:, 0], merged_pairs[:, 1]), kind="mergesort")
    merged_pairs = merged_pairs[order]
    candidate_pairs = candidate_pairs[order]
    diags = diags[order]
    use_blosum = mode == "protein"
    use_sw = local_alignment and use_blosum
    scores = None
    if use_sw and use_c_sw:
        sims, scores, _, mask = batch_sw_align(merged_pairs, merged["flat_sequences"],
                                               merged["offsets"], merged["lengths"],
                                               c_sw_band_width, BLOSUM62,
                                               diag_hints=(diags if use_c_scoring and not use_similar
                                                           else None))
    elif use_sw:
        sims, scores, _, mask = batch_sw_align(merged_pairs, merged["flat_sequences"],
                                               merged["offsets"], merged["lengths"],
                                               band_width, BLOSUM62)
    elif use_blosum:
        # Global identity with BLOSUM62 (kmer_index.py:2281-2290).
        sims, scores, mask = _batch_align_compact_scored(
            merged_pairs, merged["flat_sequences"], merged["offsets"], merged["lengths"],
            np.float32(threshold), np.int32(band_width), BLOSUM62)
    else:
        # Global identity, +2/-1 (kmer_index.py:2291-2299).
        sims, mask = _batch_align_compact(
            merged_pairs, merged["flat_sequences"], merged["offsets"], merged["lengths"],
            np.float32(threshold), np.int32(band_width))
    keep = mask.astype(bool)
    pp, ps = candidate_pairs[keep], sims[keep]
    rank = scores[keep].astype(np.float32) if scores is not None else None
    if use_sw and evalue_normalize and len(pp):
        rank = evalue_rank(rank, query_dataset.lengths[pp[:, 0]],
                           db_index.dataset.lengths[pp[:, 1]])
    hits = _collect_top_k_hits(pp, ps, nq, top_k, query_dataset, db_index.dataset,
                               passing_scores=rank)
    return SearchResults(hits, nq, nd, total, len(candidate_pairs),
                         time.perf_counter() - t0)


def search_kmer_index_device(index: DeviceKmerIndex, query_dataset: SequenceSet,
                             threshold: float = 0.5, top_k: int = 10,
                             band_width: int | None = None, min_total_hits: int = 2,
                             min_diag_hits: int = 2, diag_bin_width: int = 10,
                             max_cands_per_query: int = 2000, phase_a_topk: int = 10000,
                             freq_percentile: float = 95.0, evalue_normalize: bool = True,
                             chunk: int = 16384, timings: dict | None = None,
                             as_arrays: bool = False,
                             device_topk: bool = True,
                             local_alignment: bool = True,
                             kmer_score_thresh: int = 13) -> SearchResults:
    """The standard search path with every array stage on the device: the
    database and its index stay in HBM (``index``), queries are uploaded once,
    and per chunk of queries k-mer Phase A/B (ck_kmer_score_device), the
    candidate pairs (ck_search_pairs_device) and banded SW
    (ck_sw_align_device, mask score > 0) run back to back on one stream;
    only candidate ids, identities, scores and masks come back for the
    top-k ranking (search.py:727-800).  Same results as ``search_kmer_index``
    without hints (kmer_index.py:1475-2349).  ``as_arrays``: ``hits`` is a
    dict of arrays (query, target, identity, score; grouped by query, best
    first) instead of per-query lists of SearchHit.  ``device_topk``: rank
    and select the top_k per query on the device (ck_search_topk_device, the
    rank's logarithm read from a table NumPy fills, so ranks are
    bit-identical); off, or for length products of 2^24 and above, the
    host selects from the copied-back candidates.  ``local_alignment=False``
    (or nucleotide data): stage 2 is banded global identity
    (ck_nw_align_device), ranked as ``search_kmer_index`` ranks it.
    ``kmer_score_thresh``: for k = 5 protein indexes, > 0 selects the
    BLOSUM-similar stage 1 as in ``search_kmer_index`` (the reference's
    default); 0 scores exact k-mers."""
    import time

    import torch

    from .device import DeviceSequences, nw_align_device, sw_align_device
    lib = _lib.require_gpu()
    t0 = time.perf_counter()
    nq, nd = query_dataset.num_sequences, index.num_db
    if nq == 0 or nd == 0:
        return SearchResults([[] for _ in range(nq)], nq, nd, 0, 0, time.perf_counter() - t0)
    dev = index.device
    mc = int(max_cands_per_query)
    ft = compute_freq_threshold(index.db.kmer_freqs, nd, freq_percentile)
    # stage 1 as search_kmer_index chooses it (kmer_index.py:1576)
    use_similar = (index.k >= 5 and index.db.params["mode"] == "protein"
                   and kmer_score_thresh > 0)
    if use_similar and index.k != 5:
        raise NotImplementedError("similar k-mer scoring is defined for k = 5 "
                                  "(kmer_index.py:1206-1207)")
    if band_width is None:
        band_width = auto_band_width(query_dataset, index.db.dataset)
    qlens = _c(query_dataset.lengths, np.int32)
    qflat = query_dataset.flat_sequences
    db_flat_len = int(index.db_flat.numel())
    # combined sequence table: database, queries, one empty sentinel
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    flat = torch.cat([index.db_flat, t(qflat if len(qflat) else np.zeros(1, np.uint8))])
    q_off_local = t(_c(query_dataset.offsets, np.int64))
    q_len = t(qlens)
    offs = torch.cat([index.db_offsets, q_off_local + db_flat_len,
                      torch.zeros(1, dtype=torch.int64, device=dev)])
    lens = torch.cat([index.db_lengths, q_len, torch.zeros(1, dtype=torch.int32, device=dev)])
    empty_index = nd + nq
    seqs = DeviceSequences(flat=flat, offsets=offs, lengths=lens,
                           max_len=max(index.max_db_len, int(qlens.max())), n=nd + nq + 1)
    q_flat_dev = flat[db_flat_len:]
    max_qlen = int(qlens.max())
    stream = torch.cuda.current_stream()
    hits = [[] for _ in range(nq)]
    arrs = []
    n_cand = 0
    tk = {"kmer_s": 0.0, "sw_s": 0.0, "host_s": 0.0}
    dlen = index.db.dataset.lengths
    protein = index.db.params["mode"] == "protein"
    use_sw = local_alignment and protein
    # Rank of kmer_index.py:2310-2322: the SW score (E-value-style when
    # evalue_normalize), the global score with a table, else the identity.
    if use_sw:
        rank_mode = _lib.CK_RANK_EVALUE if evalue_normalize else _lib.CK_RANK_SCORE
    else:
        rank_mode = _lib.CK_RANK_SCORE if protein else _lib.CK_RANK_IDENTITY
    logtab = None
    use_dev_topk = device_topk and mc <= 2048
    if use_dev_topk and rank_mode == _lib.CK_RANK_EVALUE:
        logtab = index.log_table(max_qlen * max(index.max_db_len, 1))
        use_dev_topk = logtab is not None
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    for q0 in range(0, nq, chunk):
        n = min(chunk, nq - q0)
        tg = torch.empty((n, mc), dtype=torch.int32, device=dev)
        cn = torch.empty(n, dtype=torch.int32, device=dev)
        sc = torch.empty((n, mc), dtype=torch.int32, device=dev)
        dg = torch.empty((n, mc), dtype=torch.int32, device=dev)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(stream)
        if use_similar:
            index.score_similar(q_flat_dev, q_off_local[q0:q0 + n], q_len[q0:q0 + n],
                                max_qlen, ft, kmer_score_thresh, out=(tg, cn, sc, dg),
                                stream=stream, min_total_hits=min_total_hits, max_cands=mc)
        else:
            index.score(q_flat_dev, q_off_local[q0:q0 + n], q_len[q0:q0 + n], max_qlen, ft,
                        out=(tg, cn, sc, dg), stream=stream, min_total_hits=min_total_hits,
                        min_diag_hits=min_diag_hits, diag_bin_width=diag_bin_width,
                        max_cands=mc, phase_a_topk=phase_a_topk)
        e1.record(stream)
        pairs = torch.empty((n * mc, 2), dtype=torch.int32, device=dev)
        _lib.check(lib.ck_search_pairs_device(tg.data_ptr(), cn.data_ptr(), n, mc, nd + q0,
                                              empty_index, pairs.data_ptr(),
                                              stream.cuda_stream))
        if use_sw:
            out = sw_align_device(seqs, pairs, int(band_width), BLOSUM62, stream=stream)
        else:
            out = nw_align_device(seqs, pairs, int(band_width), BLOSUM62 if protein else None,
                                  threshold=np.float32(threshold), stream=stream)
        e2.record(stream)
        if use_dev_topk:
            o_t = torch.empty((n, top_k), dtype=torch.int32, device=dev)
            o_i = torch.empty((n, top_k), dtype=torch.float32, device=dev)
            o_r = torch.empty((n, top_k), dtype=torch.float32, device=dev)
            o_n = torch.empty(n, dtype=torch.int32, device=dev)
            _lib.check(lib.ck_search_topk_device(
                tg.data_ptr(), cn.data_ptr(), out["sims"].data_ptr(), out["scores"].data_ptr(),
                out["mask"].data_ptr(), q_len[q0:q0 + n].data_ptr(), index.db_lengths.data_ptr(),
                logtab.data_ptr() if logtab is not None else None,
                int(logtab.numel()) if logtab is not None else 0, n, mc, int(top_k),
                rank_mode, o_t.data_ptr(), o_i.data_ptr(), o_r.data_ptr(),
                o_n.data_ptr(), status.data_ptr(), stream.cuda_stream))
            h_cn = cn.cpu().numpy()
            h_n = o_n.cpu().numpy()
            h_t, h_i, h_r = o_t.cpu().numpy(), o_i.cpu().numpy(), o_r.cpu().numpy()
            tk["kmer_s"] += e0.elapsed_time(e1) / 1e3
            tk["sw_s"] += e1.elapsed_time(e2) / 1e3
            th = time.perf_counter()
            n_cand += int(h_cn.sum())
            keep = np.arange(top_k)[None, :] < h_n[:, None]
            qi = np.nonzero(keep)[0].astype(np.int64)
            sel = (h_t[keep], h_i[keep], h_r[keep])
            if as_arrays:
                arrs.append((q0 + qi, sel[0], sel[1], sel[2]))
            else:
                hits[q0:q0 + n] = _hits_from_arrays(
                    qi, sel[0], sel[1], sel[2], n,
                    SequenceSet(query_dataset.flat_sequences, query_dataset.offsets[q0:q0 + n],
                                query_dataset.lengths[q0:q0 + n], query_dataset.ids[q0:q0 + n]),
                    index.db.dataset)
            tk["host_s"] += time.perf_counter() - th
            continue
        h_tg = tg.cpu().numpy()
        h_cn = cn.cpu().numpy()
        h_sims = out["sims"].cpu().numpy().reshape(n, mc)
        h_scores = out["scores"].cpu().numpy().reshape(n, mc)
        h_mask = out["mask"].cpu().numpy().reshape(n, mc)
        tk["kmer_s"] += e0.elapsed_time(e1) / 1e3
        tk["sw_s"] += e1.elapsed_time(e2) / 1e3
        th = time.perf_counter()
        n_cand += int(h_cn.sum())
        valid = (np.arange(mc)[None, :] < h_cn[:, None]) & h_mask.astype(bool)
        qi, ci = np.nonzero(valid)                      # row-major: query, then candidate
        ti = h_tg[qi, ci]
        sims = h_sims[qi, ci]
        rank = h_scores[qi, ci].astype(np.float32) if rank_mode != _lib.CK_RANK_IDENTITY \
            else sims
        if rank_mode == _lib.CK_RANK_EVALUE and len(qi):
            rank = evalue_rank(rank, qlens[q0 + qi], dlen[ti])
        if as_arrays:
            sel = _top_k_order(qi, rank, top_k)
            arrs.append((q0 + qi[sel], ti[sel], sims[sel], rank[sel]))
        else:
            hits[q0:q0 + n] = _collect_top_k_hits(
                np.stack([qi, ti], axis=1), sims, n, top_k,
                SequenceSet(query_dataset.flat_sequences, query_dataset.offsets[q0:q0 + n],
                            query_dataset.lengths[q0:q0 + n], query_dataset.ids[q0:q0 + n]),
                index.db.dataset, passing_scores=rank)
        tk["host_s"] += time.perf_counter() - th
    if use_dev_topk and int(status.item()) != 0:
        raise _lib.HipLibraryError("device top-k: a length product fell outside the rank table")
    if timings is not None:
        timings.update(tk)
    if as_arrays:
        cat = lambda i, dt: (np.concatenate([a[i] for a in arrs]).astype(dt) if arrs  # noqa: E731
                             else np.zeros(0, dt))
        hits = {"query": cat(0, np.int32), "target": cat(1, np.int32),
                "identity": cat(2, np.float32), "score": cat(3, np.float32)}
    return SearchResults(hits, nq, nd, n_cand, n_cand, time.perf_counter() - t0)
