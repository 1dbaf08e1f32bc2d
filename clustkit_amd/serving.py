"""Continuous-batching search serving over an HBM-resident database.

The database sequences, their k-mer index and a per-stream query region stay
in HBM.  Requests (one query each) arrive over time; the scheduler

* admits arrived requests into a FIFO;
* whenever a stream slot is free, forms a batch of up to ``max_batch``
  waiting requests whose workspace fits the HBM budget still unreserved
  (batches already on the GPU hold their reservation until they finish);
* launches the batch on that stream -- query upload, k-mer Phase A/B
  scoring, candidate-pair construction, banded SW, result download -- so
  batches on different streams overlap on the device;
* completes requests when their batch's event has fired: per query, the
  passing candidates are ranked (0.267*S - ln(qlen*tlen), float32) and the
  top-k kept, exactly as clustkit/kmer_index.py:2303-2331 and
  clustkit/search.py:727-800 do.

Results of a request do not depend on the batch it rode in: every device
stage computes per query (per pair for SW) with order-independent exact
integer arithmetic.

The scheduling logic (``Scheduler``) is host code and is tested with a fake
executor on the CPU; ``GpuExecutor`` drives the HIP library.
"""

from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .kmer_search import DeviceKmerIndex, auto_band_width, evalue_rank
from .scoring import BLOSUM62


@dataclass
class Request:
    rid: int
    seq: np.ndarray
    arrival: float
    start: float = 0.0
    done: float = 0.0
    hits: list = field(default_factory=list)   # [(target, identity, rank_score)]
    error: str | None = None


@dataclass
class ServeParams:
    top_k: int = 10
    max_cands: int = 64
    phase_a_topk: int = 10000
    min_total_hits: int = 2
    min_diag_hits: int = 2
    diag_bin_width: int = 10
    freq_percentile: float = 95.0
    band_width: int | None = None
    use_hints: bool = False
    hint_band_width: int = 50
    evalue_normalize: bool = True


class Scheduler:
    """Admission + batching + completion bookkeeping (no GPU calls)."""

    def __init__(self, executor, hbm_budget: int, max_batch: int = 512):
        self.ex = executor
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.waiting: collections.deque[Request] = collections.deque()
        self.inflight = {}      # slot -> (batch, reserved bytes)
        self.reserved = 0
        self.completed: list[Request] = []
        self.rejected: list[Request] = []
        self.batches = 0

    def admit(self, req: Request) -> None:
        self.waiting.append(req)

    def _take_batch(self):
        batch, need = [], 0
        while self.waiting and len(batch) < self.max_batch:
            nxt = self.waiting[0]
            b = self.ex.bytes_for(batch + [nxt])
            if b > self.budget:
                if not batch and not self.inflight:
                    # cannot run even alone on an idle device: reject
                    self.waiting.popleft()
                    nxt.error = f"needs {b} bytes of HBM workspace, budget {self.budget}"
                    self.rejected.append(nxt)
                    continue
                break
            if self.reserved + b > self.budget:
                break
            batch.append(self.waiting.popleft())
            need = b
        return batch, need

    def pump(self, now: float) -> int:
        """Complete finished batches, launch new ones; returns #launched."""
        for slot in list(self.inflight):
            batch, need = self.inflight[slot]
            if self.ex.poll(slot):
                t = self.ex.finish(slot, batch)
                for r in batch:
                    r.done = t
                self.completed.extend(batch)
                self.reserved -= need
                del self.inflight[slot]
        launched = 0
        while self.waiting:
            slot = self.ex.free_slot(set(self.inflight))
            if slot is None:
                break
            batch, need = self._take_batch()
            if not batch:
                break
            for r in batch:
                r.start = now
            self.ex.launch(slot, batch)
            self.inflight[slot] = (batch, need)
            self.reserved += need
            self.batches += 1
            launched += 1
        return launched

    def idle(self) -> bool:
        return not self.waiting and not self.inflight


def serve(scheduler: Scheduler, requests: list[Request], clock=time.perf_counter,
          poll_sleep: float = 0.0002):
    """Replay ``requests`` (arrival times relative to the start) in real time."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = clock()
    for r in reqs:
        r.arrival += t0
    i = 0
    while i < len(reqs) or not scheduler.idle():
        now = clock()
        while i < len(reqs) and reqs[i].arrival <= now:
            scheduler.admit(reqs[i])
            i += 1
        scheduler.pump(now)
        if not scheduler.inflight and not scheduler.waiting and i < len(reqs):
            wait = reqs[i].arrival - clock()
            if wait > 0:
                time.sleep(min(wait, 0.002))
        else:
            time.sleep(poll_sleep)
    return scheduler.completed


def latency_summary(done: list[Request]) -> dict:
    lat = np.array([(r.done - r.arrival) * 1e3 for r in done if r.error is None])
    if lat.size == 0:
        return {"n": 0}
    return {"n": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "max_ms": round(float(lat.max()), 3)}


class GpuExecutor:
    """Runs batches on HIP streams against a DeviceKmerIndex."""

    def __init__(self, index: DeviceKmerIndex, params: ServeParams, n_streams: int = 4,
                 max_batch: int = 512, max_qlen: int = 2000, device="cuda:0",
                 sw_graph: bool = False, prewarm: bool = True, latency_pairs: int = 16384):
        """sw_graph: replay the SW call as a captured hipGraph.  Off by
        default: every batch has a different pair count, so each call would
        re-capture, which costs more host time than launching directly.
        prewarm: run one full batch (max_batch queries of max_qlen residues)
        on every stream at start-up, so the library's per-stream workspaces
        reach their final size before serving (a workspace that grows while
        serving frees and reallocates device memory, which stalls every
        stream).
        latency_pairs: batches of at most this many SW pairs run one pair per
        wavefront (CK_SW_LATENCY), which cannot fill the GPU either way but
        finish each pair about four times sooner; 0 turns this off."""
        import torch
        from .kmer_index import compute_freq_threshold
        self.torch = torch
        self.sw_flags = 0 if sw_graph else _lib.CK_SW_NO_GRAPH
        self.latency_pairs = int(latency_pairs)
        self.host_launch_s = 0.0
        self.host_finish_s = 0.0
        self.idx = index
        self.p = params
        self.n_streams = n_streams
        self.max_batch = max_batch
        self.max_qlen = max_qlen
        self.dev = torch.device(device)
        ds = index.db
        nd = index.num_db
        self.freq_thresh = compute_freq_threshold(ds.kmer_freqs, nd, params.freq_percentile)
        # The reference derives the band from the query set at hand
        # (kmer_index.py:2155-2160); a server does not know future queries, so
        # the band is fixed at start-up: given, or the same rule over the
        # database lengths alone.
        self.band = params.hint_band_width if params.use_hints else (
            params.band_width if params.band_width is not None else
            auto_band_width(_LenOnly(np.zeros(0, np.int32)), ds.dataset))
        # Combined sequence table for SW: database, then one query region per
        # stream, then a zero-length sentinel sequence.
        db_flat_len = int(index.db_flat.numel())
        self.qcap = max_batch * max_qlen
        self.db_flat_len = db_flat_len
        flat = torch.zeros(db_flat_len + n_streams * self.qcap + 1, dtype=torch.uint8,
                           device=self.dev)
        flat[:db_flat_len] = index.db_flat
        n_idx = nd + n_streams * max_batch + 1
        offs = torch.zeros(n_idx, dtype=torch.int64, device=self.dev)
        lens = torch.zeros(n_idx, dtype=torch.int32, device=self.dev)
        offs[:nd] = index.db_offsets
        lens[:nd] = index.db_lengths
        self.empty_index = n_idx - 1
        self.flat, self.offs, self.lens = flat, offs, lens
        self.max_len = max(index.max_db_len, max_qlen)
        mc = params.max_cands
        self.streams = [torch.cuda.Stream(device=self.dev) for _ in range(n_streams)]
        self.slots = []
        for s in range(n_streams):
            pin = lambda n, dt: torch.empty(n, dtype=dt, pin_memory=True)  # noqa: E731
            self.slots.append({
                "h_res": pin(self.qcap, torch.uint8), "h_off": pin(max_batch, torch.int64),
                "h_len": pin(max_batch, torch.int32), "h_goff": pin(max_batch, torch.int64),
                "d_off": torch.empty(max_batch, dtype=torch.int64, device=self.dev),
                "d_len": torch.empty(max_batch, dtype=torch.int32, device=self.dev),
                "tg": torch.empty((max_batch, mc), dtype=torch.int32, device=self.dev),
                "cn": torch.empty(max_batch, dtype=torch.int32, device=self.dev),
                "sc": torch.empty((max_batch, mc), dtype=torch.int32, device=self.dev),
                "dg": torch.empty((max_batch, mc), dtype=torch.int32, device=self.dev),
                "pairs": torch.empty((max_batch * mc, 2), dtype=torch.int32, device=self.dev),
                "sims": torch.empty(max_batch * mc, dtype=torch.float32, device=self.dev),
                "scores": torch.empty(max_batch * mc, dtype=torch.int32, device=self.dev),
                "matches": torch.empty(max_batch * mc, dtype=torch.int32, device=self.dev),
                "mask": torch.empty(max_batch * mc, dtype=torch.uint8, device=self.dev),
                "h_tg": pin(max_batch * mc, torch.int32), "h_cn": pin(max_batch, torch.int32),
                "h_sims": pin(max_batch * mc, torch.float32),
                "h_scores": pin(max_batch * mc, torch.int32),
                "h_mask": pin(max_batch * mc, torch.uint8),
                "event": torch.cuda.Event(), "n": 0,
            })
        lib = _lib.require_gpu()
        self.q_bytes = int(lib.ck_kmer_query_bytes(nd, index.max_db_len, max_qlen, index.k,
                                                   params.diag_bin_width, mc,
                                                   params.phase_a_topk))
        self.sub = np.ascontiguousarray(BLOSUM62, np.int8).reshape(-1)
        if prewarm:
            rng = np.random.default_rng(0)
            fake = [Request(rid=-1, seq=rng.integers(0, 20, max_qlen).astype(np.uint8),
                            arrival=0.0) for _ in range(max_batch)]
            for slot in range(n_streams):
                self.launch(slot, fake)
                self.slots[slot]["event"].synchronize()
                self.finish(slot, fake)
            self.host_launch_s = self.host_finish_s = 0.0

    def bytes_for(self, batch) -> int:
        mc = self.p.max_cands
        # k-mer workspace per query + SW pairs, outputs and queue per candidate
        return len(batch) * (self.q_bytes + mc * (8 + 4 * 6 + 16))

    def free_slot(self, busy):
        for s in range(self.n_streams):
            if s not in busy:
                return s
        return None

    def launch(self, slot: int, batch) -> None:
        t0 = time.perf_counter()
        try:
            self._launch(slot, batch)
        finally:
            self.host_launch_s += time.perf_counter() - t0

    def _launch(self, slot: int, batch) -> None:
        torch = self.torch
        lib = _lib.load()
        st = self.slots[slot]
        n = len(batch)
        if n > self.max_batch:
            raise ValueError("batch larger than max_batch")
        lens = np.array([len(r.seq) for r in batch], np.int32)
        if n and int(lens.max()) > self.max_qlen:
            raise ValueError(f"query longer than max_qlen={self.max_qlen}")
        offs = np.zeros(n, np.int64)
        if n:
            np.cumsum(lens[:-1].astype(np.int64), out=offs[1:])
        total = int(lens.sum())
        res = st["h_res"].numpy()
        for j, r in enumerate(batch):
            res[offs[j]:offs[j] + lens[j]] = r.seq[:lens[j]]
        region = self.db_flat_len + slot * self.qcap
        st["h_off"].numpy()[:n] = offs
        st["h_len"].numpy()[:n] = lens
        st["h_goff"].numpy()[:n] = offs + region
        base = self.idx.num_db + slot * self.max_batch
        stream = self.streams[slot]
        mc = self.p.max_cands
        with torch.cuda.stream(stream):
            self.flat[region:region + total].copy_(st["h_res"][:total], non_blocking=True)
            st["d_off"][:n].copy_(st["h_off"][:n], non_blocking=True)
            st["d_len"][:n].copy_(st["h_len"][:n], non_blocking=True)
            self.offs[base:base + n].copy_(st["h_goff"][:n], non_blocking=True)
            self.lens[base:base + n].copy_(st["h_len"][:n], non_blocking=True)
            qflat = self.flat[region:region + max(total, 1)]
            self.idx.score(qflat, st["d_off"][:n], st["d_len"][:n], int(lens.max(initial=0)),
                           self.freq_thresh,
                           out=(st["tg"][:n], st["cn"][:n], st["sc"][:n], st["dg"][:n]),
                           stream=stream, min_total_hits=self.p.min_total_hits,
                           min_diag_hits=self.p.min_diag_hits,
                           diag_bin_width=self.p.diag_bin_width, max_cands=mc,
                           phase_a_topk=self.p.phase_a_topk)
            _lib.check(lib.ck_search_pairs_device(
                st["tg"].data_ptr(), st["cn"].data_ptr(), n, mc, base, self.empty_index,
                st["pairs"].data_ptr(), stream.cuda_stream))
            m = n * mc
            _lib.check(lib.ck_sw_align_device(
                st["pairs"].data_ptr(), m, self.flat.data_ptr(), self.offs.data_ptr(),
                self.lens.data_ptr(), self.max_len, int(self.band), self.sub.ctypes.data,
                st["dg"].data_ptr() if self.p.use_hints else None, 0.0,
                self.sw_flags | (_lib.CK_SW_LATENCY if m <= self.latency_pairs else 0),
                st["sims"].data_ptr(), st["scores"].data_ptr(), st["matches"].data_ptr(),
                st["mask"].data_ptr(), stream.cuda_stream))
            st["h_tg"][:m].copy_(st["tg"][:n].reshape(-1), non_blocking=True)
            st["h_cn"][:n].copy_(st["cn"][:n], non_blocking=True)
            st["h_sims"][:m].copy_(st["sims"][:m], non_blocking=True)
            st["h_scores"][:m].copy_(st["scores"][:m], non_blocking=True)
            st["h_mask"][:m].copy_(st["mask"][:m], non_blocking=True)
            st["event"].record(stream)
        st["n"] = n

    def poll(self, slot: int) -> bool:
        return self.slots[slot]["event"].query()

    def finish(self, slot: int, batch) -> float:
        t0 = time.perf_counter()
        done = self._finish(slot, batch)
        self.host_finish_s += time.perf_counter() - t0
        return done

    def _finish(self, slot: int, batch) -> float:
        st = self.slots[slot]
        n, mc = st["n"], self.p.max_cands
        tg = st["h_tg"].numpy()[:n * mc].reshape(n, mc)
        cn = st["h_cn"].numpy()[:n]
        sims = st["h_sims"].numpy()[:n * mc].reshape(n, mc)
        scores = st["h_scores"].numpy()[:n * mc].reshape(n, mc)
        mask = st["h_mask"].numpy()[:n * mc].reshape(n, mc)
        dlen = self.idx.db.dataset.lengths
        for j, r in enumerate(batch):
            c = int(cn[j])
            keep = mask[j, :c].astype(bool)
            t = tg[j, :c][keep]
            s = sims[j, :c][keep]
            sc = scores[j, :c][keep].astype(np.float32)
            rank = evalue_rank(sc, np.full(len(t), len(r.seq)), dlen[t]) \
                if self.p.evalue_normalize else sc
            order = np.lexsort((-rank, np.zeros(len(rank))))[:self.p.top_k]
            r.hits = [(int(t[o]), float(s[o]), float(rank[o])) for o in order]
        return time.perf_counter()


class _LenOnly:
    def __init__(self, lengths):
        self.lengths = lengths
