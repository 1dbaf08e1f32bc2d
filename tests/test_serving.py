"""Continuous-batching scheduler (host logic, fake executor on the CPU) and the
GPU serving path: every request's hits are the same alone and in the mix."""

from __future__ import annotations

import time

import numpy as np
import pytest

from clustkit_amd.serving import Request, Scheduler, latency_summary, serve


class FakeExecutor:
    """Completes a batch `delay` pump calls after launch; records batches."""

    def __init__(self, n_slots=2, per_query=100, delay=2):
        self.n_slots, self.per_query, self.delay = n_slots, per_query, delay
        self.t = {}
        self.batches = []

    def bytes_for(self, batch):
        return sum(self.per_query * max(1, len(r.seq) // 100) for r in batch)

    def free_slot(self, busy):
        return next((s for s in range(self.n_slots) if s not in busy), None)

    def launch(self, slot, batch):
        self.t[slot] = self.delay
        self.batches.append([r.rid for r in batch])

    def poll(self, slot):
        self.t[slot] -= 1
        return self.t[slot] <= 0

    def finish(self, slot, batch):
        for r in batch:
            r.hits = [(r.rid, 1.0, 0.0)]
        return time.perf_counter()


def _req(i, n=100):
    return Request(rid=i, seq=np.zeros(n, np.uint8), arrival=0.0)


def test_scheduler_fifo_batches_and_slots():
    ex = FakeExecutor(n_slots=2, per_query=100)
    sch = Scheduler(ex, hbm_budget=10_000, max_batch=4)
    for i in range(10):
        sch.admit(_req(i))
    sch.pump(0.0)
    assert ex.batches == [[0, 1, 2, 3], [4, 5, 6, 7]]       # two slots, FIFO, max_batch
    assert len(sch.inflight) == 2
    for _ in range(5):
        sch.pump(0.0)
    assert sorted(r.rid for r in sch.completed) == list(range(10))
    assert sch.idle() and sch.reserved == 0


def test_scheduler_hbm_budget_defers_and_rejects():
    ex = FakeExecutor(n_slots=3, per_query=100)
    sch = Scheduler(ex, hbm_budget=250, max_batch=8)       # two 100-byte queries fit
    for i in range(5):
        sch.admit(_req(i))
    sch.admit(_req(99, n=400))                              # needs 400 > 250: rejected
    sch.pump(0.0)
    assert ex.batches == [[0, 1]]                           # budget, not slots, limits
    assert sch.reserved == 200
    while not sch.idle():
        sch.pump(0.0)
    assert [b for b in ex.batches] == [[0, 1], [2, 3], [4]]
    assert [r.rid for r in sch.rejected] == [99] and sch.rejected[0].error
    assert sorted(r.rid for r in sch.completed) == [0, 1, 2, 3, 4]


def test_scheduler_admits_while_batches_in_flight():
    ex = FakeExecutor(n_slots=2, per_query=10, delay=3)
    sch = Scheduler(ex, hbm_budget=10_000, max_batch=100)
    sch.admit(_req(0))
    sch.pump(0.0)
    sch.admit(_req(1))
    sch.admit(_req(2))
    sch.pump(0.0)                      # second slot takes the late arrivals
    assert ex.batches == [[0], [1, 2]]
    while not sch.idle():
        sch.pump(0.0)
    assert len(sch.completed) == 3


def test_serve_replays_arrivals_and_reports_latency():
    ex = FakeExecutor(n_slots=2, per_query=1, delay=1)
    sch = Scheduler(ex, hbm_budget=1 << 20, max_batch=16)
    reqs = [Request(rid=i, seq=np.zeros(50, np.uint8), arrival=i * 0.001) for i in range(20)]
    done = serve(sch, reqs)
    assert len(done) == 20
    s = latency_summary(done)
    assert s["n"] == 20 and s["p99_ms"] >= s["p50_ms"] >= 0


@pytest.mark.gpu
def test_gpu_serving_results_same_alone_and_in_mix():
    import torch

    from clustkit_amd.kmer_search import DeviceKmerIndex, SequenceSet, build_database_index
    from clustkit_amd.serving import GpuExecutor, ServeParams
    from clustkit_amd.workloads import FamilySpec, generate_families, mutate
    db = generate_families(FamilySpec(n_roots=1500, n_desc=9, median=300, clip_lo=30,
                                      clip_hi=1500, id_lo=0.3, id_hi=0.9, seed=41))
    ds = SequenceSet.from_arrays(db.flat, db.offsets, db.lengths)
    idx = DeviceKmerIndex(build_database_index(ds, k=3), torch.device("cuda:0"))
    rng = np.random.default_rng(8)
    seqs = []
    for i in range(120):
        if i % 2 == 0:
            r = int(rng.integers(0, len(db.lengths)))
            s = mutate(rng, db.flat[db.offsets[r]:db.offsets[r] + db.lengths[r]],
                       float(rng.uniform(0.3, 0.9)))
            seqs.append(s[:2000] if len(s) >= 50 else np.resize(s, 50))
        else:
            seqs.append(rng.integers(0, 20, int(rng.integers(50, 2000))).astype(np.uint8))
    params = ServeParams(top_k=5, max_cands=64)
    ex = GpuExecutor(idx, params, n_streams=3, max_batch=64, max_qlen=2000)
    print("serving: alone", flush=True)
    alone = {}
    for i, s in enumerate(seqs):
        sch = Scheduler(ex, hbm_budget=1 << 40, max_batch=1)
        done = serve(sch, [Request(rid=i, seq=s, arrival=0.0)])
        alone[i] = done[0].hits
    print("serving: mix", flush=True)
    arr = np.cumsum(rng.exponential(1 / 2000.0, len(seqs)))
    sch = Scheduler(ex, hbm_budget=1 << 40, max_batch=64)
    done = serve(sch, [Request(rid=i, seq=s, arrival=float(a))
                       for i, (s, a) in enumerate(zip(seqs, arr))])
    assert len(done) == len(seqs) and sch.batches < len(seqs)
    for r in done:
        assert r.hits == alone[r.rid], r.rid
    assert any(len(h) for h in alone.values())
