"""Alignment-stage benchmark on one MI355X.

Workload (default ``mixed2m``): 50,000 synthetic proteins (5,000 roots x 10,
median length 350, lognormal sigma 0.5, clipped 30-2000, descendant identity
0.3-0.95), sketched with k=5 / s=128, LSH candidates with the automatic
parameters for threshold 0.3 (512 tables x 1 band) -> ~1.6M candidate
pairs; banded SW with BLOSUM62, gap 11/1, band 100, threshold 0.3.  The
Jaccard pre-filter (floor 0.001, pairwise.py:949-1001) runs on device first;
its 2000-pair RandomState(0) sample estimates ~51% removal, so it applies and
~0.77M pairs enter SW, as they would in the reference.

One "step" = one pass of the SW stage over all candidate pairs already
resident in HBM (work-queue construction + alignment + int32 re-run).
``value`` is the batched form's GCUPS; the plain form (one pair per thread,
int32, one launch per 500,000-pair batch) is timed the same way and reported
beside it.  GPU results are checked bit-exactly against the reference
kernel (oracle/_ref) on the CPU-baseline sample.

    python bench.py [--steps K] [--warmup W] [--workload mixed2m|small]
    python bench.py --probe        # one launch + copy calibration, for rocprofv3 --pmc
    python bench.py --candidates-only   # device sketch + LSH stage only
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SW GCUPS + pairs/s, batched vs plain, 1 MI355X, bit-exact; search p99 ms"
OPS_PER_CELL = 26                       # integer ops per cell of sw_align.c:110-159
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # int32 lane-ops/s: 256 CU x 64 lanes x 2.4 GHz

WORKLOADS = {
    "mixed2m": dict(spec="mixed2m", k=5, sketch=128, threshold=0.3, band=100,
                    desc="50,000 synthetic proteins (5,000 roots x 10, lognormal median 350, "
                         "clip 30-2000, identity 0.3-0.95), k=5 s=128 LSH candidates, "
                         "band 100, t=0.3"),
    "heavy": dict(spec="heavy", k=5, sketch=128, threshold=0.3, band=64,
                  desc="100,000 synthetic proteins (10,000 roots x 10; 90% lognormal median "
                       "250 clip 30-2000, 10% Pareto(1.5) 1000-8000), k=5 s=128 LSH "
                       "candidates, t=0.3, bands 16/32/64/128"),
    "fast200k": dict(spec="fast200k", k=3, sketch=0, threshold=0.3, band=0,
                     desc="200,000 synthetic proteins (20,000 roots x 10, lognormal median 300, "
                          "clip 30-3000, identity 0.25-0.95); every sequence searched against "
                          "all: k-mer index (k=3 and k=5) -> top 64 -> banded SW -> top 10"),
    "small": dict(spec="small", k=5, sketch=128, threshold=0.5, band=100,
                  desc="2,000 synthetic proteins (200 roots x 10, lognormal median 300, "
                       "clip 50-1000, identity 0.3-0.9), k=5 s=128 LSH candidates, "
                       "band 100, t=0.5"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name):
    """Synthetic families (host), then sketches and LSH candidate pairs on
    the device (ck_sketch_device, ck_lsh_candidates_device), each timed.
    Returns (families, sketches uint64 (N, s), pairs (M, 2), lsh params,
    candidate-stage timings)."""
    import torch
    from clustkit_amd.device import DeviceSequences
    from clustkit_amd.lsh import auto_lsh_params, lsh_candidates_device
    from clustkit_amd.sketch import compute_sketches_device
    from clustkit_amd.workloads import SPECS, generate_families
    w = WORKLOADS[name]
    t0 = time.perf_counter()
    fam = generate_families(SPECS[w["spec"]])
    tgen = time.perf_counter() - t0
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths)
    p = auto_lsh_params(w["threshold"], "medium", w["k"])
    timing = {}
    for rep in range(3):           # first pass warms up; the last one is reported
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        dsk = compute_sketches_device(seqs, w["k"], w["sketch"], "protein", 42)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        dpairs = lsh_candidates_device(dsk, p["num_tables"], p["num_bands"], 42)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        timing = {"sketch_ms": round((t2 - t1) * 1e3, 3), "lsh_ms": round((t3 - t2) * 1e3, 3)}
    sk = dsk.cpu().numpy().view(np.uint64)
    pairs = dpairs.cpu().numpy()
    timing.update({"sequences": int(len(fam.lengths)), "residues": int(fam.lengths.sum()),
                   "k": w["k"], "sketch_size": w["sketch"], "tables": p["num_tables"],
                   "bands": p["num_bands"], "candidate_pairs": int(len(pairs))})
    log(f"workload {name}: {len(fam.lengths)} seqs ({tgen:.1f}s host generation), "
        f"{len(pairs)} candidate pairs; device sketch {timing['sketch_ms']} ms, "
        f"LSH {timing['lsh_ms']} ms")
    return fam, sk, pairs, p, timing


def candidates_cpu(fam, w, lshp, sk, pairs):
    """The oracle's vectorised NumPy sketch and LSH (one host thread) on the
    same sequences, timed, and compared word for word with the device
    results."""
    from oracle import sketch_lsh
    t0 = time.perf_counter()
    osk = sketch_lsh.compute_sketches(fam.flat, fam.offsets, fam.lengths, w["k"], w["sketch"],
                                      "protein", seed=42)
    t1 = time.perf_counter()
    opairs = sketch_lsh.lsh_candidates(osk, lshp["num_tables"], lshp["num_bands"], seed=42)
    t2 = time.perf_counter()
    return {"kind": "port", "cores": 1, "sketch_ms": round((t1 - t0) * 1e3, 1),
            "lsh_ms": round((t2 - t1) * 1e3, 1),
            "sketches_equal": bool(np.array_equal(osk, sk)),
            "pairs_equal": bool(np.array_equal(opairs, pairs))}


def cpu_baseline(fam, pairs, band, target_s=12.0):
    """Reference kernel (oracle/_ref, compiled from clustkit/csrc/sw_align.c) or,
    where it is absent, the oracle port, on a bounded random sample."""
    from clustkit_amd.cells import pair_cells
    from clustkit_amd.scoring import BLOSUM62
    from oracle import oracle
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, 16)
    kind = "reference" if oracle.ref_available() else "port"
    rng = np.random.default_rng(1)

    def run(sample, threads=cores):
        t = time.perf_counter()
        if kind == "reference":
            s, sc, mk = oracle.ref_sw_batch(sample, fam.flat, fam.offsets, fam.lengths, band,
                                            BLOSUM62, n_threads=threads)
        else:
            o = oracle.sw_batch(sample, fam.flat, fam.offsets, fam.lengths, band, BLOSUM62,
                                n_threads=threads)
            s, sc, mk = o["sims"], o["scores"], o["mask"]
        return time.perf_counter() - t, (s, sc, mk)

    probe = pairs[rng.choice(len(pairs), size=min(len(pairs), 2000), replace=False)]
    dt, _ = run(probe)
    n = int(min(len(pairs), max(2000, len(probe) * target_s / max(dt, 1e-3))))
    idx = np.sort(rng.choice(len(pairs), size=n, replace=False))
    sample = pairs[idx]
    dt, res = run(sample)
    cells = int(pair_cells(sample, fam.lengths, band).sum())
    # one core, on a sample sized for ~target_s / 2 of work
    n1 = int(max(200, min(n, n * (target_s / 2) / max(dt * cores, 1e-3))))
    s1 = sample[np.sort(rng.choice(len(sample), size=min(n1, len(sample)), replace=False))]
    dt1, _ = run(s1, threads=1)
    cells1 = int(pair_cells(s1, fam.lengths, band).sum())
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": int(cores),
            "kind": kind,
            "sample": f"{n} of {len(pairs)} candidate pairs drawn uniformly, {cells} cells, "
                      f"{dt:.1f}s wall, sw_set_num_threads({cores})",
            "pairs_per_s": round(n / dt, 1), "cpu_model": model,
            "single_core": {"value": round(cells1 / dt1 / 1e9, 4), "unit": "GCUPS",
                            "sample": f"{len(s1)} pairs of the sample above, {cells1} cells, "
                                      f"{dt1:.1f}s wall, sw_set_num_threads(1)"}}, idx, res


NW_OPS_PER_CELL = {"plain": 22, "scored": 26}   # ops per cell of pairwise.py:327-357 / 446-486


def nw_bench(fam, pairs, seqs, d_pairs, band, thr, steps, warmup, with_cpu):
    """Banded global alignment (NW identity, pairwise.py:684-850) over the same
    candidate pairs: the plain (+2/-1, _batch_align_compact) and scored
    (BLOSUM62, _batch_align_compact_scored) forms, at threshold 0 (no early
    stop: every band cell evaluated, the roofline figure) and at the
    workload's threshold (the call as clustering makes it).  Events on the
    launch stream; the oracle on a bounded sample on the host cores."""
    import torch
    from clustkit_amd import _lib
    from clustkit_amd.cells import nw_pair_cells
    from clustkit_amd.device import alloc_outputs, last_paths, nw_align_device
    from clustkit_amd.scoring import BLOSUM62
    stream = torch.cuda.current_stream()
    m = len(pairs)
    out = alloc_outputs(m, d_pairs.device)
    res = {"workload": "the SW workload's candidate pairs after the Jaccard pre-filter; "
                       f"band_width {band} (adaptive half band min(bw, max(10, diff + 10)))",
           "pairs": m}

    def timed(sub, t, flags, k):
        for _ in range(warmup):
            nw_align_device(seqs, d_pairs, band, sub, threshold=t, flags=flags, out=out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            nw_align_device(seqs, d_pairs, band, sub, threshold=t, flags=flags, out=out)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    cells0 = int(nw_pair_cells(pairs, fam.lengths, band, 0.0).sum())
    res["cells_threshold0"] = cells0
    for form, sub in (("plain", None), ("scored", BLOSUM62)):
        ms0 = timed(sub, 0.0, 0, steps)
        single = last_paths()["int32"]
        ms_t = timed(sub, thr, 0, steps)
        ms_pl = timed(sub, 0.0, _lib.CK_NW_PLAIN, max(1, steps // 2))
        g = {k: v.cpu().numpy() for k, v in out.items()}   # threshold-0, plain-kernel run
        nw_align_device(seqs, d_pairs, band, sub, threshold=0.0, out=out)
        torch.cuda.synchronize()
        g16 = {k: v.cpu().numpy() for k, v in out.items()}
        same = bool(np.array_equal(g["sims"].view(np.uint32), g16["sims"].view(np.uint32))
                    and np.array_equal(g["mask"], g16["mask"])
                    and (sub is None or np.array_equal(g["scores"], g16["scores"])))
        gc = cells0 / (ms0 * 1e-3) / 1e9
        ops = NW_OPS_PER_CELL[form]
        ach = cells0 * ops / (ms0 * 1e-3) / 1e12
        res[form] = {
            "threshold0": {"ms": round(ms0, 3), "GCUPS": round(gc, 2),
                           "pairs_per_s": round(m / (ms0 * 1e-3), 1)},
            f"threshold{thr}": {"ms": round(ms_t, 3), "pairs_per_s": round(m / (ms_t * 1e-3), 1)},
            "one_pair_per_thread": {"ms": round(ms_pl, 3),
                                    "GCUPS": round(cells0 / (ms_pl * 1e-3) / 1e9, 2)},
            "sixteen_lane_over_one_pair_per_thread": round(ms_pl / ms0, 2),
            "pairs_one_per_thread_in_default": int(single),
            "both_kernels_equal": same,
            "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": VALU_PEAK_TOPS,
                         "unit": "TOP/s", "frac": round(ach / VALU_PEAK_TOPS, 4),
                         "ops_per_cell": ops, "kernel": "nw_wave_kernel (whole call timed)"},
        }
        log(f"nw {form}: t=0 {ms0:.2f} ms ({gc:.1f} GCUPS), t={thr} {ms_t:.2f} ms, "
            f"one pair per thread {ms_pl:.2f} ms, kernels equal {same}")
        if with_cpu:
            from oracle import oracle
            cores = min(len(os.sched_getaffinity(0)), 16)
            rng = np.random.default_rng(5)
            n = min(m, 200000)
            idx = np.sort(rng.choice(m, size=n, replace=False))
            t0 = time.perf_counter()
            o = oracle.nw_batch(pairs[idx], fam.flat, fam.offsets, fam.lengths, 0.0, band, sub,
                                n_threads=cores)
            dt = time.perf_counter() - t0
            c = int(nw_pair_cells(pairs[idx], fam.lengths, band, 0.0).sum())
            exact = bool(np.array_equal(g16["sims"][idx].view(np.uint32), o[0].view(np.uint32))
                         and np.array_equal(g16["mask"][idx].astype(bool), o[-1])
                         and (sub is None or np.array_equal(g16["scores"][idx], o[1])))
            res[form]["cpu_baseline"] = {
                "value": round(c / dt / 1e9, 3), "unit": "GCUPS", "cores": int(cores),
                "kind": "port",
                "sample": f"{n} of {m} pairs drawn uniformly, threshold 0, {c} cells, "
                          f"{dt:.1f}s wall (oracle/nw_oracle.c, OpenMP)"}
            res[form]["bit_exact_vs_oracle_sample"] = exact
            log(f"nw {form} cpu: {res[form]['cpu_baseline']} exact={exact}")
    return res


def graph_bench(fam, pairs, res, thr, steps):
    """Similarity graph (graph.py:7-39) on the device (ck_graph_csr_device,
    events on the launch stream) against the reference's scipy construction
    on the host: for the SW-passing pairs, and for every candidate pair as an
    edge (the scale figure)."""
    keep = (res["scores"] > 0) & (res["sims"] >= np.float32(thr))
    out = _graph_once(fam, pairs[keep], res["sims"][keep], steps)
    out["input"] = "pairs with SW score > 0 and identity >= threshold"
    out["all_candidates"] = _graph_once(fam, pairs, res["sims"], steps)
    out["all_candidates"]["input"] = "every candidate pair after the Jaccard pre-filter"
    return out


def _graph_once(fam, kp, ks, steps):
    import torch
    from clustkit_amd.device import graph_csr_device
    from oracle.graph_oracle import reference_graph
    kp = np.ascontiguousarray(kp, np.int32)
    ks = np.ascontiguousarray(ks, np.float32)
    n = len(fam.lengths)
    dev = torch.device("cuda:0")
    d_p = torch.from_numpy(kp).to(dev)
    d_s = torch.from_numpy(ks).to(dev)
    bufs = graph_csr_device(d_p, d_s, n)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        graph_csr_device(d_p, d_s, n, *bufs)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    t0 = time.perf_counter()
    ref = reference_graph(n, kp, ks)
    cpu_s = time.perf_counter() - t0
    m = int(bufs[3].item())
    same = bool(m == ref.nnz and np.array_equal(bufs[0].cpu().numpy(), ref.indptr)
                and np.array_equal(bufs[1][:m].cpu().numpy(), ref.indices)
                and np.array_equal(bufs[2][:m].cpu().numpy().view(np.uint32),
                                   ref.data.view(np.uint32)))
    log(f"graph: {len(kp)} edges -> nnz {m}, {ms:.3f} ms on the GPU, scipy {cpu_s * 1e3:.1f} ms, "
        f"identical {same}")
    return {"edges": int(len(kp)), "nnz": m, "sequences": n, "ms": round(ms, 3),
            "edges_per_s": round(len(kp) / (ms * 1e-3), 1),
            "scipy_host_ms": round(cpu_s * 1e3, 1), "identical_to_scipy": same}


def heavy_bench(args):
    """Heavy-tailed workload: one persistent launch per band over all
    candidate pairs (uneven lengths), launch configuration autotuned per
    band, CPU reference on a sample for GCUPS and bit-exactness."""
    import torch

    from clustkit_amd import _lib, autotune
    from clustkit_amd.cells import pair_cells
    from clustkit_amd.device import (DeviceSequences, alloc_outputs, jaccard_device,
                                     last_stats, last_timing, sw_align_device)
    from clustkit_amd.scoring import BLOSUM62
    from oracle import oracle
    w = WORKLOADS["heavy"]
    thr = w["threshold"]
    fam, sk, pairs, lshp, _ = build_workload("heavy")
    dev = torch.device("cuda:0")
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths, dev)
    d_pairs = torch.from_numpy(np.ascontiguousarray(pairs, np.int32)).to(dev)
    d_sk = torch.from_numpy(sk).to(dev)
    m = len(pairs)
    floor = max(0.001, 0.3 * thr ** 5)
    samp = np.random.RandomState(0).choice(m, size=min(2000, m), replace=False)
    js = jaccard_device(d_pairs[torch.from_numpy(samp).to(dev)], d_sk).cpu().numpy()
    if float(np.mean(js < floor)) >= 0.05:
        keep = (jaccard_device(d_pairs, d_sk) >= floor).nonzero().flatten()
        d_pairs = d_pairs[keep].contiguous()
        pairs = d_pairs.cpu().numpy()
    m = len(pairs)
    out = alloc_outputs(m, dev)
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(3)
    cpu_idx = np.sort(rng.choice(m, size=min(m, 20000), replace=False))
    rows, tot_cells, tot_ms = [], 0, 0.0
    for band in [int(b) for b in args.bands.split(",") if b]:
        cells = int(pair_cells(pairs, fam.lengths, band).sum())
        tuned = None if args.no_autotune else autotune.tune(seqs, d_pairs, band)
        sw_align_device(seqs, d_pairs, band, BLOSUM62, threshold=thr, out=out)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            sw_align_device(seqs, d_pairs, band, BLOSUM62, threshold=thr, out=out)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        st = last_stats()
        sw_align_device(seqs, d_pairs, band, BLOSUM62, threshold=thr,
                        flags=_lib.CK_SW_TIMING, out=out)
        q, mm, rr = last_timing()
        got = {k: v.cpu().numpy() for k, v in out.items()}
        exact = None
        if not args.no_cpu:
            sub = pairs[cpu_idx]
            t0 = time.perf_counter()
            s_, sc_, mk_ = oracle.ref_sw_batch(sub, fam.flat, fam.offsets, fam.lengths, band,
                                               BLOSUM62, n_threads=16) \
                if oracle.ref_available() else (None, None, None)
            cdt = time.perf_counter() - t0
            if sc_ is not None:
                exact = bool(np.array_equal(got["scores"][cpu_idx], sc_) and
                             np.array_equal(got["sims"][cpu_idx].view(np.uint32),
                                            s_.view(np.uint32)))
        row = {"band": band, "gcups": round(cells / (ms * 1e-3) / 1e9, 2), "ms_per_step": round(ms, 3),
               "cells": cells, "pairs_per_s": round(m / (ms * 1e-3), 1),
               "phases_ms": {"queue": round(q, 3), "main": round(mm, 3), "rerun": round(rr, 3)},
               "rerun_pairs": st["rerun"],
               "launch_config": ({"lanes_per_pair": tuned["lanes"],
                                  "bin_shift": tuned["bin_shift"],
                                  "pipelined": tuned["pipelined"], "autotuned_ms": tuned["ms"]}
                                 if tuned else "default"),
               "bit_exact_vs_reference_sample": exact}
        if not args.no_cpu and exact is not None:
            sc_cells = int(pair_cells(pairs[cpu_idx], fam.lengths, band).sum())
            row["cpu_reference_gcups_16_threads"] = round(sc_cells / cdt / 1e9, 3)
        log(f"heavy band {band}: {row}")
        rows.append(row)
        tot_cells += cells
        tot_ms += ms
    line = {"metric": METRIC + " (heavy-tail workload)", "value": round(tot_cells / (tot_ms * 1e-3) / 1e9, 2),
            "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": 1,
            "ms_per_step": round(tot_ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int32", "data": "synthetic",
            "config": {"workload": f"heavy: {w['desc']}", "pairs": m, "threshold": thr,
                       "lsh": lshp},
            "bands": rows}
    print(json.dumps(line), flush=True)


def fast_bench(args):
    """Fast path at 200,000 sequences: each sequence is a query against the
    whole set; index build, k-mer Phase A/B, candidate pairs and banded SW
    on the device (search_kmer_index_device), top-10 ranking on the host.
    k = 3 (dense, LDS-tiled) and k = 5 (sparse, exact k-mers).  The reference
    k-mer scorer (oracle/_ref) is timed on a query sample and its candidate
    lists compared with the device's."""
    import torch

    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import (DeviceKmerIndex, SequenceSet, auto_band_width,
                                          search_kmer_index_device)
    from clustkit_amd.workloads import SPECS, generate_families_streams
    from oracle import oracle
    w = WORKLOADS["fast200k"]
    t0 = time.perf_counter()
    fam = generate_families_streams(SPECS["fast200k"], workers=args.gen_workers, progress=log)
    ds = SequenceSet.from_arrays(fam.flat, fam.offsets, fam.lengths)
    log(f"fast200k: {ds.num_sequences} seqs, {fam.flat.size} residues "
        f"({time.perf_counter() - t0:.1f}s host generation)")
    nq = ds.num_sequences
    band = auto_band_width(ds, ds)
    rows = []
    for k in (3, 5):
        torch.cuda.synchronize()
        tb = time.perf_counter()
        idx = DeviceKmerIndex.build(ds, k=k)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - tb
        sub = SequenceSet.from_arrays(fam.flat, fam.offsets[:256], fam.lengths[:256])
        search_kmer_index_device(idx, sub, threshold=w["threshold"], max_cands_per_query=64)
        tm = {}
        torch.cuda.synchronize()
        ts = time.perf_counter()
        res = search_kmer_index_device(idx, ds, threshold=w["threshold"], top_k=10,
                                       max_cands_per_query=64, timings=tm, as_arrays=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - ts
        row = {"k": k, "index_postings": idx.n_postings, "index_tiled": idx.tiled,
               "index_build_s": round(build_s, 3), "band": band, "queries": nq,
               "search_s": round(wall, 3), "queries_per_s": round(nq / wall, 1),
               "kmer_stage_s": round(tm["kmer_s"], 3), "sw_stage_s": round(tm["sw_s"], 3),
               "host_rank_s": round(tm["host_s"], 3), "candidates": res.num_candidates,
               "queries_with_hits": int(len(np.unique(res.hits["query"]))),
               "hits": int(len(res.hits["query"]))}
        if not args.no_cpu and oracle.ref_kmer_available():
            # reference scorer on a query sample: time and candidate equality
            rng = np.random.default_rng(7)
            qs = np.sort(rng.choice(nq, size=400, replace=False))
            qf = np.concatenate([fam.flat[fam.offsets[q]:fam.offsets[q] + fam.lengths[q]]
                                 for q in qs])
            ql = fam.lengths[qs].astype(np.int32)
            qo = np.zeros(len(qs), np.int64)
            qo[1:] = np.cumsum(ql[:-1])
            ke = idx.host_entries()
            ft = compute_freq_threshold(idx.db.kmer_freqs, nq, 95.0)
            tc = time.perf_counter()
            rt, rc, rsc, rd = oracle.ref_kmer_score_batch(
                qf, qo, ql, k, 20, idx.db.kmer_offsets, ke, idx.db.kmer_freqs, ft, nq,
                n_threads=16, max_cands=64)
            cdt = time.perf_counter() - tc
            dt_, dc_, ds_, dd_ = idx.score(
                torch.from_numpy(qf).cuda(), torch.from_numpy(qo).cuda(),
                torch.from_numpy(ql).cuda(), int(ql.max()), ft, max_cands=64)
            dc_ = dc_.cpu().numpy()
            dt_, ds_ = dt_.cpu().numpy(), ds_.cpu().numpy()
            same = bool(np.array_equal(dc_, rc) and all(
                np.array_equal(dt_[i, :rc[i]], rt[i, :rc[i]]) and
                np.array_equal(ds_[i, :rc[i]], rsc[i, :rc[i]]) for i in range(len(qs))))
            row["cpu_reference_kmer_queries_per_s_16_threads"] = round(len(qs) / cdt, 1)
            row["kmer_candidates_equal_reference_sample"] = same
        log(f"fast200k k={k}: {row}")
        rows.append(row)
        del idx
        torch.cuda.empty_cache()
    line = {"metric": METRIC + " (fast path)", "value": rows[0]["queries_per_s"],
            "unit": "queries/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": round(rows[0]["search_s"] * 1e3, 1), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
            "config": {"workload": f"fast200k: {w['desc']}"}, "k_runs": rows}
    print(json.dumps(line), flush=True)


def search_bench(n_db, rates, seconds, dev, gen_workers=16):
    """Search serving over an HBM-resident database: k-mer index built on the
    device, Poisson query arrivals at each offered rate, continuous batching
    (4 streams, batches of up to 512 under an HBM budget), k-mer Phase A/B +
    banded SW of the top 64 candidates + top-10 ranking per query."""
    import dataclasses

    import torch

    from clustkit_amd.kmer_search import DeviceKmerIndex, SequenceSet
    from clustkit_amd.serving import (GpuExecutor, Request, Scheduler, ServeParams,
                                      latency_summary, serve)
    from clustkit_amd.workloads import SPECS, generate_families_streams, make_queries
    spec = SPECS["serve1m"]
    if n_db != 1_000_000:
        spec = dataclasses.replace(spec, n_roots=max(1, n_db // 10))
    t0 = time.perf_counter()
    db = generate_families_streams(spec, workers=gen_workers, progress=log)
    log(f"search db: {len(db.lengths)} seqs, {db.flat.size} residues "
        f"({time.perf_counter() - t0:.1f}s host generation)")
    ds = SequenceSet.from_arrays(db.flat, db.offsets, db.lengths)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    idx = DeviceKmerIndex.build(ds, k=3, device=dev)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"device k-mer index: {idx.n_postings} postings, tiled={idx.tiled}, "
        f"{build_s:.2f}s incl. upload")
    params = ServeParams(top_k=10, max_cands=64)
    ex = GpuExecutor(idx, params, n_streams=4, max_batch=512, max_qlen=2000, device=dev)
    import gc
    gc.collect()
    gc.freeze()             # long-lived objects out of the collector's way while serving
    budget = 32 << 30
    total_q = sum(max(200, int(r * seconds)) for r in rates) + 64 + 32
    qs, src = make_queries(db, total_q, seed=11)
    qi = 0

    def take(n):
        nonlocal qi
        out = qs[qi:qi + n]
        qi += n
        return out

    # warm-up: one full burst
    serve(Scheduler(ex, budget, 512), [Request(rid=i, seq=q, arrival=0.0)
                                       for i, q in enumerate(take(64))])
    rng = np.random.default_rng(12)
    per_rate = []
    mix_hits = {}
    tot_batches = 0
    ex.host_launch_s = ex.host_finish_s = 0.0
    for rate in rates:
        n = max(200, int(rate * seconds))
        arr = np.cumsum(rng.exponential(1.0 / rate, n))
        reqs = [Request(rid=i, seq=q, arrival=float(a)) for i, (q, a) in
                enumerate(zip(take(n), arr))]
        sch = Scheduler(ex, budget, 512)
        t0 = time.perf_counter()
        done = serve(sch, reqs)
        wall = time.perf_counter() - t0
        summ = latency_summary(done)
        first = min(r.arrival for r in done)
        last = max(r.done for r in done)
        summ.update({"offered_qps": rate, "achieved_qps": round(len(done) / (last - first), 1),
                     "batches": sch.batches, "rejected": len(sch.rejected),
                     "wall_s": round(wall, 2)})
        per_rate.append(summ)
        tot_batches += sch.batches
        log(f"search @ {rate} q/s: {summ}")
        mix_hits = {r.rid: (r.seq, r.hits) for r in done[:32]}
    host_ms = {"launch": round(ex.host_launch_s / max(1, tot_batches) * 1e3, 3),
               "finish": round(ex.host_finish_s / max(1, tot_batches) * 1e3, 3)}
    # each of 32 queries of the last mix, run alone, gives the same hits
    same = True
    for rid, (q, hits) in mix_hits.items():
        d = serve(Scheduler(ex, budget, 1), [Request(rid=rid, seq=q, arrival=0.0)])
        same &= d[0].hits == hits
    ok = [r for r in per_rate if r.get("n")]
    head = max(ok, key=lambda r: r["offered_qps"]) if ok else {}
    return {
        "workload": f"serve1m-form database of {len(db.lengths):,} synthetic proteins "
                    "(median 350, clip 30-5000), k=3 index resident in HBM; queries 50-2000 "
                    "aa, 50% planted descendants at 30-90% identity; Poisson arrivals; "
                    "top 64 k-mer candidates -> banded SW -> top 10",
        "db_seqs": int(len(db.lengths)), "db_residues": int(db.flat.size),
        "index_postings": int(idx.n_postings), "index_build_s": round(build_s, 3),
        "band": int(ex.band), "streams": 4, "max_batch": 512,
        "host_ms_per_batch": host_ms,
        "p50_ms": head.get("p50_ms"), "p99_ms": head.get("p99_ms"),
        "qps": head.get("achieved_qps"), "at_offered_qps": head.get("offered_qps"),
        "rates": per_rate, "alone_equals_mix": bool(same),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed2m", choices=sorted(WORKLOADS))
    ap.add_argument("--plain-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch configuration instead of tuning it")
    ap.add_argument("--bands", default="16,32,64,128", help="bands of the heavy workload")
    ap.add_argument("--search-db", type=int, default=1_000_000,
                    help="sequences in the search-serving database (0 = skip serving)")
    ap.add_argument("--search-rates", default="50,200,1000,2000",
                    help="offered Poisson arrival rates, queries/s")
    ap.add_argument("--search-seconds", type=float, default=4.0)
    ap.add_argument("--gen-workers", type=int, default=16,
                    help="worker processes of the synthetic database generator")
    ap.add_argument("--probe", action="store_true",
                    help="one batched launch, then a 1 GiB device copy (counter calibration)")
    ap.add_argument("--no-nw", action="store_true",
                    help="skip the banded global alignment (NW identity) leg")
    ap.add_argument("--candidates-only", action="store_true",
                    help="only the device sketch + LSH stage of the workload (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(the alignment stage shards by pair ranges across GPUs, "
                                   "not exercised here)"}))
        sys.exit(2)

    import torch

    from clustkit_amd import _lib
    from clustkit_amd.cells import pair_cells
    from clustkit_amd.device import (DeviceSequences, alloc_outputs, jaccard_device,
                                     last_stats, last_timing, sw_align_device)
    from clustkit_amd.scoring import BLOSUM62

    _lib.require_gpu()
    if args.workload == "heavy":
        return heavy_bench(args)
    if args.workload == "fast200k":
        return fast_bench(args)
    w = WORKLOADS[args.workload]
    band, thr = w["band"], w["threshold"]
    fam, sk, pairs, lshp, cand = build_workload(args.workload)
    if args.candidates_only:
        print(json.dumps({"candidates": cand, "workload": args.workload}), flush=True)
        return
    if not args.no_cpu:
        cand["cpu_oracle"] = candidates_cpu(fam, w, lshp, sk, pairs)
        log(f"candidate stage on the host: {cand['cpu_oracle']}")
    cells = pair_cells(pairs, fam.lengths, band)
    total_cells = int(cells.sum())
    m = len(pairs)

    dev = torch.device("cuda:0")
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths, dev)
    d_pairs = torch.from_numpy(np.ascontiguousarray(pairs, np.int32)).to(dev)
    d_sk = torch.from_numpy(sk).to(dev)
    out = alloc_outputs(m, dev)

    # Jaccard pre-filter decision (pairwise.py:949-1001), on device.
    floor = max(0.001, 0.3 * thr ** 5)
    samp = np.random.RandomState(0).choice(m, size=min(2000, m), replace=False)
    js = jaccard_device(d_pairs[torch.from_numpy(samp).to(dev)], d_sk).cpu().numpy()
    jrate = float(np.mean(js < floor))
    jaccard_applied = jrate >= 0.05
    if jaccard_applied:
        jall = jaccard_device(d_pairs, d_sk)
        keep = (jall >= floor).nonzero().flatten()
        d_pairs = d_pairs[keep].contiguous()
        pairs = d_pairs.cpu().numpy()
        cells = pair_cells(pairs, fam.lengths, band)
        total_cells = int(cells.sum())
        m = len(pairs)
        out = alloc_outputs(m, dev)
    log(f"jaccard sample removal {jrate:.3%} -> applied={jaccard_applied}; {m} pairs, "
        f"{total_cells / 1e9:.2f} G cells")

    if args.probe:
        # For hardware-counter runs: exactly one launch of each SW kernel, then a
        # plain device-to-device copy of known size whose FETCH_SIZE / WRITE_SIZE
        # readings scale the counters to bytes.
        sw_align_device(seqs, d_pairs, band, BLOSUM62, threshold=thr, out=out)
        torch.cuda.synchronize()
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"probe": args.workload, "pairs": m, "cells": total_cells,
                          "copy_bytes": a.numel() * 4}), flush=True)
        return

    # CPU baseline on a bounded sample of the pairs that enter SW.
    base = None
    if not args.no_cpu:
        base, cpu_idx, cpu_res = cpu_baseline(fam, pairs, band)
        log(f"cpu baseline: {base}")

    stream = torch.cuda.current_stream()

    def timed(flags, steps, warmup, label):
        """K steps of the call as a user makes it (the batched form replays
        its captured hipGraph), bracketed by events on the launch stream;
        then K steps launched directly with per-phase HIP events
        (CK_SW_TIMING) for the phase split and the roofline."""
        for _ in range(warmup):
            sw_align_device(seqs, d_pairs, band, BLOSUM62, threshold=thr, flags=flags, out=out)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            sw_align_device(seqs, d_pairs, band, BLOSUM62, threshold=thr, flags=flags, out=out)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        total_ms = ev0.elapsed_time(ev1)
        st = last_stats()
        main_ms, queue_ms, rerun_ms = [], [], []
        for _ in range(steps):
            sw_align_device(seqs, d_pairs, band, BLOSUM62, threshold=thr,
                            flags=flags | _lib.CK_SW_TIMING, out=out)
            q, mm, r = last_timing()
            queue_ms.append(q); main_ms.append(mm); rerun_ms.append(r)
        torch.cuda.synchronize()
        log(f"{label}: {total_ms / steps:.2f} ms/step (wall {wall / steps * 1e3:.2f}), "
            f"queue {np.mean(queue_ms):.2f} main {np.mean(main_ms):.2f} rerun "
            f"{np.mean(rerun_ms):.2f} ms, stats {st}")
        return total_ms / steps, float(np.mean(main_ms)), st, {
            "queue_ms": round(float(np.mean(queue_ms)), 3),
            "main_ms": round(float(np.mean(main_ms)), 3),
            "rerun_ms": round(float(np.mean(rerun_ms)), 3)}

    tuned = None
    if not args.no_autotune:
        from clustkit_amd import autotune
        tuned = autotune.tune(seqs, d_pairs, band)
        log(f"autotune: {tuned}")

    # Batched form (the headline).
    b_ms, b_main, b_stats, b_phase = timed(0, args.steps, args.warmup, "batched")
    res_b = {k: v.cpu().numpy() for k, v in out.items()}
    # Plain form.
    p_steps = args.plain_steps or max(1, args.steps // 2)
    p_ms, p_main, _, p_phase = timed(_lib.CK_SW_PLAIN, p_steps, min(args.warmup, 1), "plain")
    res_p = {k: v.cpu().numpy() for k, v in out.items()}

    plain_equal = all(np.array_equal(res_b[k], res_p[k]) for k in ("scores", "matches", "mask")) \
        and np.array_equal(res_b["sims"].view(np.uint32), res_p["sims"].view(np.uint32))
    bit_exact = None
    if base is not None:
        s, sc, mk = cpu_res
        bit_exact = bool(np.array_equal(res_b["scores"][cpu_idx], sc)
                         and np.array_equal(res_b["sims"][cpu_idx].view(np.uint32),
                                            s.view(np.uint32))
                         and np.array_equal(res_b["mask"][cpu_idx], mk))

    gcups = total_cells / (b_ms * 1e-3) / 1e9
    gcups_plain = total_cells / (p_ms * 1e-3) / 1e9
    # Dominant kernel: the persistent batched SW kernel, timed with HIP events on
    # the launch stream.  Algorithmic ops per launch = 26 x cells it evaluated.
    rerun_cells = 0
    main_cells = total_cells - rerun_cells
    achieved = main_cells * OPS_PER_CELL / (b_main * 1e-3) / 1e12
    traffic = None
    from clustkit_amd import autotune as _at
    lanes_cfg, _shift_cfg, pipe_cfg = _at.current()
    seg_lanes = lanes_cfg if lanes_cfg in (8, 16) else (8 if 1 <= band <= 32 else 16)
    kernel_name = "sw_pipe_kernel" if (seg_lanes == 16 and pipe_cfg) else "sw_wave_kernel"
    tfile = ROOT / "profiles" / f"traffic_{kernel_name}_mixed2m.txt"
    if args.workload == "mixed2m" and tfile.exists():
        for ln in tfile.read_text().splitlines():
            if ln.startswith("traffic_bytes_per_launch:"):
                traffic = int(ln.split(":")[1])
    line = {
        "metric": METRIC,
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w['desc']}", "pairs": m,
                   "cells": total_cells, "band": band, "threshold": thr,
                   "lsh": lshp, "jaccard_prefilter_applied": jaccard_applied},
        "pairs_per_s": round(m / (b_ms * 1e-3), 1),
        "plain": {"value": round(gcups_plain, 2), "unit": "GCUPS",
                  "pairs_per_s": round(m / (p_ms * 1e-3), 1), "ms_per_step": round(p_ms, 3),
                  "steps": p_steps, "phases": p_phase,
                  "form": "one pair per thread, int32, one launch per 500000-pair batch"},
        "batched_over_plain": round(gcups / gcups_plain, 2),
        "batched_form": "persistent wave kernel over a cost-binned device work queue, "
                        "int32 re-run of saturated pairs, replayed as one captured hipGraph",
        "batched_phases": b_phase,
        "launch_config": ({"lanes_per_pair": tuned["lanes"], "bin_shift": tuned["bin_shift"],
                           "pipelined": tuned["pipelined"], "autotuned_ms": tuned["ms"]}
                          if tuned else "default"),
        "rerun_pairs": b_stats["rerun"],
        "bit_exact_vs_reference_sample": bit_exact,
        "plain_equals_batched": bool(plain_equal),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS,
                     "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                     "traffic": traffic, "kernel": kernel_name,
                     "ops_per_cell": OPS_PER_CELL},
        "cpu_baseline": base,
        "candidates": cand,
    }
    line["graph"] = graph_bench(fam, pairs, res_b, thr, args.steps)
    if not args.no_nw:
        line["nw"] = nw_bench(fam, pairs, seqs, d_pairs, band, thr, args.steps, args.warmup,
                              not args.no_cpu)
    if args.search_db > 0:
        rates = [float(x) for x in args.search_rates.split(",") if x]
        line["search"] = search_bench(args.search_db, rates, args.search_seconds, dev,
                                      args.gen_workers)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
