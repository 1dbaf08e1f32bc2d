"""Alignment-stage benchmark on one MI355X.

Headline workload (``mixed2m``): 50,000 synthetic proteins (5,000 roots x 10,
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

The same JSON line carries the other workloads, each with its own roofline,
CPU baseline and parity check (``configs``): the 2,000-protein workload at
t=0.5 (``small``), the heavy-tailed 100,000-protein workload (``heavy``, band
64; all four bands with ``--workload heavy``), the 200,000-protein k-mer fast
path (``fast200k``, k=3 and k=5), and search serving over a 1,000,000-protein
database (``search``); plus the NW identity (``nw``), similarity graph
(``graph``) and candidate (``candidates``) stages.

    python bench.py [--steps K] [--warmup W]
    python bench.py --workload heavy|fast200k|small   # one workload alone
    python bench.py --probe        # one launch + copy calibration, for rocprofv3 --pmc
    python bench.py --candidates-only   # device sketch + LSH stage only
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SW GCUPS + pairs/s, batched vs plain, 1 MI355X, bit-exact; search p99 ms"
OPS_PER_CELL = 26                       # integer ops per cell of sw_align.c:110-159
HBM_PEAK_GBS = 8000.0                   # MI355X HBM3E, published
TRAFFIC_FILE = "profiles/traffic_sw_pipe_kernel_mixed2m.txt"

WORKLOADS = {
    "mixed2m": dict(spec="mixed2m", k=5, sketch=128, threshold=0.3, band=100,
                    desc="50,000 synthetic proteins (5,000 roots x 10, lognormal median 350, "
                         "clip 30-2000, identity 0.3-0.95), k=5 s=128 LSH candidates, "
                         "band 100, t=0.3"),
    "heavy": dict(spec="heavy", k=5, sketch=128, threshold=0.3, band=64,
                  desc="100,000 synthetic proteins (10,000 roots x 10; 90% lognormal median "
                       "250 clip 30-2000, 10% Pareto(1.5) 1000-8000), k=5 s=128 LSH "
                       "candidates, t=0.3"),
    "fast200k": dict(spec="fast200k", k=3, sketch=0, threshold=0.3, band=0,
                     desc="200,000 synthetic proteins (20,000 roots x 10, lognormal median 300, "
                          "clip 30-3000, identity 0.25-0.95); every sequence searched against "
                          "all: k-mer index (k=3 and k=5) -> top 64 -> banded SW -> top 10"),
    "small": dict(spec="small", k=5, sketch=128, threshold=0.5, band=100,
                  desc="2,000 synthetic proteins (200 roots x 10, lognormal median 300, "
                       "clip 50-1000, identity 0.3-0.9), k=5 s=128 LSH candidates "
                       "(146 tables x 1 band), band 100, t=0.5"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# Host and device facts
# ---------------------------------------------------------------------------

def host_info() -> dict:
    """CPU model and core counts (lscpu), the CPUs this process may use
    (affinity mask) and the thread share the CPU baselines run with."""
    info = {"logical_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "baseline_threads": cpu_threads()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {}
        for ln in out.splitlines():
            if ":" in ln:
                k, v = ln.split(":", 1)
                kv[k.strip()] = v.strip()
        sockets = int(kv.get("Socket(s)", "0") or 0)
        cps = int(kv.get("Core(s) per socket", "0") or 0)
        info.update({"model": kv.get("Model name"), "sockets": sockets,
                     "cores_per_socket": cps,
                     "threads_per_core": int(kv.get("Thread(s) per core", "0") or 0),
                     "physical_cores": sockets * cps})
    except (OSError, ValueError, subprocess.SubprocessError) as e:
        info["lscpu_error"] = repr(e)
    return info


def cpu_threads() -> int:
    """Threads of the CPU baselines: every CPU of the affinity mask, within
    the process's OpenMP share (OMP_NUM_THREADS) when one is set."""
    aff = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(aff, int(share))) if share and share.isdigit() else aff


def device_peaks(props: dict) -> dict:
    """Integer VALU peak (CUs x 64 lanes x engine clock, one 32-bit op per
    lane per cycle) from hipDeviceProp, and the published HBM peak."""
    valu = props["num_cu"] * 64 * props["clock_mhz"] * 1e6 / 1e12
    return {"valu_tops": round(valu, 4), "hbm_gbs": HBM_PEAK_GBS,
            "source": f"hipDeviceProp: {props['num_cu']} CUs x 64 lanes x "
                      f"{props['clock_mhz']:.0f} MHz ({props['arch']}); HBM 8 TB/s published"}


def sw_roofline(cells, kernel_ms, peaks, kernel, traffic=None, traffic_source=None):
    ach = cells * OPS_PER_CELL / (kernel_ms * 1e-3) / 1e12
    r = {"bound": "valu", "achieved": round(ach, 3), "peak": peaks["valu_tops"],
         "unit": "TOP/s", "frac": round(ach / peaks["valu_tops"], 4), "traffic": traffic,
         "kernel": kernel, "ops_per_cell": OPS_PER_CELL,
         "kernel_ms": round(kernel_ms, 3)}
    if traffic_source:
        r["traffic_source"] = traffic_source
    return r


# ---------------------------------------------------------------------------
# Workload construction
# ---------------------------------------------------------------------------

def build_workload(name, gen_workers=1):
    """Synthetic families (host), then sketches and LSH candidate pairs on
    the device (ck_sketch_device, ck_lsh_candidates_device), each timed.
    Returns (families, sketches uint64 (N, s), pairs (M, 2), lsh params,
    candidate-stage timings)."""
    import torch
    from clustkit_amd.device import DeviceSequences
    from clustkit_amd.lsh import auto_lsh_params, lsh_candidates_device
    from clustkit_amd.sketch import compute_sketches_device
    from clustkit_amd.workloads import SPECS, generate_families, generate_families_streams
    w = WORKLOADS[name]
    t0 = time.perf_counter()
    fam = (generate_families_streams(SPECS[w["spec"]], workers=gen_workers)
           if gen_workers > 1 else generate_families(SPECS[w["spec"]]))
    tgen = time.perf_counter() - t0
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths)
    p = auto_lsh_params(w["threshold"], "medium", w["k"])
    timing = {}
    for _ in range(3):           # first passes warm up; the last one is reported
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


def jaccard_on_device(d_pairs, d_sk, thr):
    """The pre-filter decision and mask of pairwise.py:949-1001 on device
    pointers: returns (pairs kept (device), applied, sampled removal rate)."""
    import torch
    from clustkit_amd.device import jaccard_device
    m = int(d_pairs.shape[0])
    floor = max(0.001, 0.3 * thr ** 5)
    if m > 2000:
        samp = np.random.RandomState(0).choice(m, size=2000, replace=False)
        js = jaccard_device(d_pairs[torch.from_numpy(samp).to(d_pairs.device)], d_sk)
        rate = float(np.mean(js.cpu().numpy() < floor))
        if rate < 0.05:
            return d_pairs, False, rate
    else:
        rate = float("nan")
    keep = (jaccard_device(d_pairs, d_sk) >= floor).nonzero().flatten()
    return d_pairs[keep].contiguous(), True, rate


# ---------------------------------------------------------------------------
# SW timing and CPU baselines
# ---------------------------------------------------------------------------

def timed_sw(seqs, d_pairs, band, thr, steps, warmup, flags, out, label):
    """K steps of the call as a user makes it (the batched form replays its
    captured hipGraph), bracketed by events on the launch stream; then K
    steps launched directly with per-phase HIP events (CK_SW_TIMING) for the
    phase split and the kernel time of the roofline."""
    import torch
    from clustkit_amd import _lib
    from clustkit_amd.device import last_stats, last_timing, sw_align_device
    from clustkit_amd.scoring import BLOSUM62
    stream = torch.cuda.current_stream()
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
        queue_ms.append(q)
        main_ms.append(mm)
        rerun_ms.append(r)
    torch.cuda.synchronize()
    log(f"{label}: {total_ms / steps:.3f} ms/step (wall {wall / steps * 1e3:.3f}), "
        f"queue {np.mean(queue_ms):.3f} main {np.mean(main_ms):.3f} rerun "
        f"{np.mean(rerun_ms):.3f} ms, stats {st}")
    return total_ms / steps, float(np.mean(main_ms)), st, {
        "queue_ms": round(float(np.mean(queue_ms)), 3),
        "main_ms": round(float(np.mean(main_ms)), 3),
        "rerun_ms": round(float(np.mean(rerun_ms)), 3)}


def _ref_or_port_sw(sample, fam, band, threads, hints=None):
    from clustkit_amd.scoring import BLOSUM62
    from oracle import oracle
    t = time.perf_counter()
    if oracle.ref_available():
        s, sc, mk = oracle.ref_sw_batch(sample, fam.flat, fam.offsets, fam.lengths, band,
                                        BLOSUM62, hints=hints, n_threads=threads)
    else:
        o = oracle.sw_batch(sample, fam.flat, fam.offsets, fam.lengths, band, BLOSUM62,
                            hints=hints, n_threads=threads)
        s, sc, mk = o["sims"], o["scores"], o["mask"]
    return time.perf_counter() - t, (s, sc, mk)


def sw_cpu_baseline(fam, pairs, band, target_s=12.0, seed=1, single=True):
    """Reference kernel (oracle/_ref, compiled from clustkit/csrc/sw_align.c)
    or, where it is absent, the oracle port, on a bounded random sample sized
    for about ``target_s`` seconds; all baseline threads, and one thread on a
    smaller sample.  Returns (baseline dict, sample indices, its results)."""
    from clustkit_amd.cells import pair_cells
    from oracle import oracle
    threads = cpu_threads()
    kind = "reference" if oracle.ref_available() else "port"
    rng = np.random.default_rng(seed)
    m = len(pairs)
    probe = pairs[rng.choice(m, size=min(m, 2000), replace=False)]
    dt, _ = _ref_or_port_sw(probe, fam, band, threads)
    n = int(min(m, max(2000, len(probe) * target_s / max(dt, 1e-3))))
    idx = np.sort(rng.choice(m, size=n, replace=False))
    sample = pairs[idx]
    dt, res = _ref_or_port_sw(sample, fam, band, threads)
    cells = int(pair_cells(sample, fam.lengths, band).sum())
    base = {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": int(threads),
            "kind": kind,
            "sample": f"{n} of {m} pairs drawn uniformly, {cells} cells, {dt:.2f}s wall, "
                      f"sw_set_num_threads({threads})",
            "pairs_per_s": round(n / dt, 1)}
    if single:
        n1 = int(max(200, min(n, n * (target_s / 2) / max(dt * threads, 1e-3))))
        s1 = sample[np.sort(rng.choice(n, size=min(n1, n), replace=False))]
        dt1, _ = _ref_or_port_sw(s1, fam, band, 1)
        cells1 = int(pair_cells(s1, fam.lengths, band).sum())
        base["single_core"] = {"value": round(cells1 / dt1 / 1e9, 4), "unit": "GCUPS",
                               "sample": f"{len(s1)} pairs of the sample above, {cells1} "
                                         f"cells, {dt1:.2f}s wall, sw_set_num_threads(1)"}
    return base, idx, res


def _bit_exact(got, idx, res):
    s, sc, mk = res
    return bool(np.array_equal(got["scores"][idx], sc)
                and np.array_equal(got["sims"][idx].view(np.uint32), s.view(np.uint32))
                and np.array_equal(got["mask"][idx], mk))


# ---------------------------------------------------------------------------
# Headline extras: end to end from host arrays
# ---------------------------------------------------------------------------

def end_to_end(fam, pairs_all, sk, band, thr, cells, reps=3):
    """Host arrays in, host arrays out, on one stream: H2D of the sequence
    table, candidate pairs and sketches; the Jaccard sample decision and
    mask; compaction; SW; D2H of the four per-pair outputs.  Wall time
    (perf_counter, synchronised) of the last of ``reps`` runs, with an event
    split; and the drop-in call compute_pairwise_alignment(device="0") on the
    same host arrays (host-pointer entry points, host sorts included)."""
    import torch
    from clustkit_amd.device import DeviceSequences, alloc_outputs, sw_align_device
    from clustkit_amd.pairwise import compute_pairwise_alignment
    from clustkit_amd.scoring import BLOSUM62
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    res = {}
    for _ in range(reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record(stream)
        seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths, dev)
        d_all = torch.from_numpy(np.ascontiguousarray(pairs_all, np.int32)).to(dev)
        d_sk = torch.from_numpy(sk.view(np.int64)).to(dev)
        ev[1].record(stream)
        d_pairs, _, _ = jaccard_on_device(d_all, d_sk, thr)
        ev[2].record(stream)
        out = alloc_outputs(int(d_pairs.shape[0]), dev)
        sw_align_device(seqs, d_pairs, band, BLOSUM62, threshold=thr, out=out)
        ev[3].record(stream)
        host = {k: v.cpu().numpy() for k, v in out.items()}
        kept = d_pairs.cpu().numpy()
        ev[4].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        res = {"ms": round(wall * 1e3, 3),
               "gcups_incl_transfers": round(cells / wall / 1e9, 2),
               "split_ms": {"h2d": round(ev[0].elapsed_time(ev[1]), 3),
                            "jaccard_and_compaction": round(ev[1].elapsed_time(ev[2]), 3),
                            "sw": round(ev[2].elapsed_time(ev[3]), 3),
                            "d2h": round(ev[3].elapsed_time(ev[4]), 3)},
               "h2d_bytes": int(fam.flat.nbytes + fam.offsets.nbytes + fam.lengths.nbytes
                                + pairs_all.nbytes + sk.nbytes),
               "d2h_bytes": int(sum(v.nbytes for v in host.values()) + kept.nbytes)}
    kw = dict(band_width=band, sketches=sk, flat_sequences=fam.flat, offsets=fam.offsets,
              use_sw=True, use_c_sw=True)
    for _ in range(2):
        t0 = time.perf_counter()
        fp, fs = compute_pairwise_alignment(pairs_all, None, fam.lengths, thr, device="0", **kw)
        dt = time.perf_counter() - t0
    res["drop_in_call"] = {"ms": round(dt * 1e3, 1), "passing_pairs": int(len(fp)),
                           "call": "compute_pairwise_alignment(candidates, None, lengths, t, "
                                   "band_width, device='0', sketches, flat_sequences, offsets, "
                                   "use_sw=True, use_c_sw=True) from host arrays"}
    log(f"end to end: {res}")
    return res


# ---------------------------------------------------------------------------
# NW identity and similarity graph legs (same candidate pairs)
# ---------------------------------------------------------------------------

NW_OPS_PER_CELL = {"plain": 22, "scored": 26}   # ops per cell of pairwise.py:327-357 / 446-486


def nw_bench(fam, pairs, seqs, d_pairs, band, thr, steps, warmup, with_cpu, peaks):
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
        ms_st = timed(sub, 0.0, _lib.CK_NW_STRIPE, steps)
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
            "one_stripe_at_a_time": {"ms": round(ms_st, 3),
                                     "GCUPS": round(cells0 / (ms_st * 1e-3) / 1e9, 2)},
            "sixteen_lane_over_one_pair_per_thread": round(ms_pl / ms0, 2),
            "pairs_one_per_thread_in_default": int(single),
            "both_kernels_equal": same,
            "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": peaks["valu_tops"],
                         "unit": "TOP/s", "frac": round(ach / peaks["valu_tops"], 4),
                         "ops_per_cell": ops, "kernel": "nw kernels (whole call timed)"},
        }
        log(f"nw {form}: t=0 {ms0:.2f} ms ({gc:.1f} GCUPS), t={thr} {ms_t:.2f} ms, "
            f"one stripe at a time {ms_st:.2f} ms, one pair per thread {ms_pl:.2f} ms, "
            f"kernels equal {same}")
        if with_cpu:
            from oracle import oracle
            threads = cpu_threads()
            rng = np.random.default_rng(5)
            n = min(m, 200000)
            idx = np.sort(rng.choice(m, size=n, replace=False))
            t0 = time.perf_counter()
            o = oracle.nw_batch(pairs[idx], fam.flat, fam.offsets, fam.lengths, 0.0, band, sub,
                                n_threads=threads)
            dt = time.perf_counter() - t0
            c = int(nw_pair_cells(pairs[idx], fam.lengths, band, 0.0).sum())
            exact = bool(np.array_equal(g16["sims"][idx].view(np.uint32), o[0].view(np.uint32))
                         and np.array_equal(g16["mask"][idx].astype(bool), o[-1])
                         and (sub is None or np.array_equal(g16["scores"][idx], o[1])))
            res[form]["cpu_baseline"] = {
                "value": round(c / dt / 1e9, 3), "unit": "GCUPS", "cores": int(threads),
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


# ---------------------------------------------------------------------------
# Configuration legs
# ---------------------------------------------------------------------------

def small_leg(args, peaks):
    """The 2,000-protein workload at t=0.5: device sketch + LSH, Jaccard
    pre-filter, SW at band 100 over every surviving pair (batched and
    one-pair-per-wavefront forms; the GPU cannot fill at this size), every
    pair checked against the reference kernel, which is also the CPU
    baseline on all pairs; and the drop-in pipeline call from host arrays."""
    import torch
    from clustkit_amd import _lib
    from clustkit_amd.cells import pair_cells
    from clustkit_amd.device import DeviceSequences, alloc_outputs
    from clustkit_amd.pairwise import compute_pairwise_alignment
    w = WORKLOADS["small"]
    band, thr = w["band"], w["threshold"]
    fam, sk, pairs_all, lshp, cand = build_workload("small")
    dev = torch.device("cuda:0")
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths, dev)
    d_all = torch.from_numpy(np.ascontiguousarray(pairs_all, np.int32)).to(dev)
    d_sk = torch.from_numpy(sk.view(np.int64)).to(dev)
    d_pairs, applied, rate = jaccard_on_device(d_all, d_sk, thr)
    pairs = d_pairs.cpu().numpy()
    m = len(pairs)
    cells = int(pair_cells(pairs, fam.lengths, band).sum())
    out = alloc_outputs(m, dev)
    steps = max(args.steps, 10)
    b_ms, b_main, _, b_ph = timed_sw(seqs, d_pairs, band, thr, steps, 2, 0, out, "small batched")
    got = {k: v.cpu().numpy() for k, v in out.items()}
    l_ms, l_main, _, l_ph = timed_sw(seqs, d_pairs, band, thr, steps, 2, _lib.CK_SW_LATENCY,
                                     out, "small latency form")
    got_l = {k: v.cpu().numpy() for k, v in out.items()}
    threads = cpu_threads()
    dt, res = _ref_or_port_sw(pairs, fam, band, threads)
    dt1, _ = _ref_or_port_sw(pairs, fam, band, 1)
    idx = np.arange(m)
    best_ms, best_main, form = (b_ms, b_main, "batched") if b_ms <= l_ms else \
        (l_ms, l_main, "one pair per wavefront")
    t0 = time.perf_counter()
    fp, _ = compute_pairwise_alignment(pairs_all, None, fam.lengths, thr, band_width=band,
                                       device="0", sketches=sk, flat_sequences=fam.flat,
                                       offsets=fam.offsets, use_sw=True, use_c_sw=True)
    drop_ms = (time.perf_counter() - t0) * 1e3
    from oracle import oracle
    leg = {"workload": f"small: {w['desc']}", "pairs": m, "cells": cells,
           "candidates": cand, "jaccard_prefilter_applied": applied,
           "jaccard_sample_removal": round(rate, 4),
           "value": round(cells / (best_ms * 1e-3) / 1e9, 2), "unit": "GCUPS",
           "ms_per_step": round(best_ms, 4), "form": form,
           "pairs_per_s": round(m / (best_ms * 1e-3), 1),
           "batched": {"ms": round(b_ms, 4), "phases": b_ph},
           "one_pair_per_wavefront": {"ms": round(l_ms, 4), "phases": l_ph},
           "bit_exact_vs_reference_all_pairs": _bit_exact(got, idx, res) and
           _bit_exact(got_l, idx, res),
           "roofline": sw_roofline(cells, best_main, peaks,
                                   "sw_pipe_kernel (batched)" if form == "batched"
                                   else "sw_pipe_kernel<64> (one pair per wavefront)"),
           "cpu_baseline": {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS",
                            "cores": threads,
                            "kind": "reference" if oracle.ref_available() else "port",
                            "sample": f"all {m} pairs, {cells} cells, {dt:.3f}s wall",
                            "single_core": {"value": round(cells / dt1 / 1e9, 4),
                                            "unit": "GCUPS"}},
           "drop_in_call": {"ms": round(drop_ms, 2), "passing_pairs": int(len(fp))}}
    log(f"small: {leg['value']} GCUPS ({form}), exact {leg['bit_exact_vs_reference_all_pairs']}")
    return leg


def heavy_leg(args, bands, peaks):
    """Heavy-tailed workload: one persistent launch per band over all
    candidate pairs (uneven lengths), launch configuration autotuned per
    band, the reference kernel on a sample for the CPU baseline and
    bit-exactness."""
    import torch
    from clustkit_amd import autotune
    from clustkit_amd.cells import pair_cells
    from clustkit_amd.device import DeviceSequences, alloc_outputs, last_paths
    w = WORKLOADS["heavy"]
    thr = w["threshold"]
    fam, sk, pairs, lshp, cand = build_workload("heavy")
    dev = torch.device("cuda:0")
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths, dev)
    d_all = torch.from_numpy(np.ascontiguousarray(pairs, np.int32)).to(dev)
    d_sk = torch.from_numpy(sk.view(np.int64)).to(dev)
    d_pairs, applied, _ = jaccard_on_device(d_all, d_sk, thr)
    pairs = d_pairs.cpu().numpy()
    m = len(pairs)
    out = alloc_outputs(m, dev)
    rows = []
    for band in bands:
        cells = int(pair_cells(pairs, fam.lengths, band).sum())
        tuned = None if args.no_autotune else autotune.tune(seqs, d_pairs, band)
        ms, main, st, ph = timed_sw(seqs, d_pairs, band, thr, args.steps, 1, 0, out,
                                    f"heavy band {band}")
        paths = last_paths()
        got = {k: v.cpu().numpy() for k, v in out.items()}
        row = {"band": band, "value": round(cells / (ms * 1e-3) / 1e9, 2), "unit": "GCUPS",
               "ms_per_step": round(ms, 3), "cells": cells,
               "pairs_per_s": round(m / (ms * 1e-3), 1), "phases_ms": ph,
               "rerun_pairs": {"wide_score": paths["wide"], "int32": paths["int32"]},
               "launch_config": ({"lanes_per_pair": tuned["lanes"],
                                  "bin_shift": tuned["bin_shift"],
                                  "pipelined": tuned["pipelined"], "autotuned_ms": tuned["ms"]}
                                 if tuned else "default"),
               "roofline": sw_roofline(cells, main, peaks, "sw_pipe_kernel (batched)")}
        if not args.no_cpu:
            base, idx, res = sw_cpu_baseline(fam, pairs, band, target_s=6.0, seed=3,
                                             single=False)
            row["cpu_baseline"] = base
            row["bit_exact_vs_reference_sample"] = _bit_exact(got, idx, res)
        log(f"heavy band {band}: {row}")
        rows.append(row)
    tot_cells = sum(r["cells"] for r in rows)
    tot_ms = sum(r["ms_per_step"] for r in rows)
    return {"workload": f"heavy: {w['desc']}", "pairs": m, "threshold": thr, "lsh": lshp,
            "candidates": cand, "jaccard_prefilter_applied": applied,
            "value": round(tot_cells / (tot_ms * 1e-3) / 1e9, 2), "unit": "GCUPS",
            "bands": rows}


def _postings_visited(flat, offsets, lengths, k, alpha, kmer_offsets, kmer_freqs, ft):
    """Per query, the posting entries the reference walks in Phase A (and
    again in Phase B): every window of k residues below alpha whose k-mer
    frequency is at most the threshold, counted once per position
    (kmer_score.c:91-119, 311-331)."""
    flat = np.asarray(flat, np.uint8)
    n = flat.size - k + 1
    if n <= 0:
        return np.zeros(len(lengths), np.int64)
    code = np.zeros(n, np.int64)
    ok = np.ones(n, bool)
    for j in range(k):
        r = flat[j:j + n].astype(np.int64)
        ok &= r < alpha
        code = code * alpha + np.minimum(r, alpha - 1)
    cnt = np.diff(np.asarray(kmer_offsets, np.int64))
    per_win = np.where(ok & (np.asarray(kmer_freqs)[code] <= ft), cnt[code], 0)
    csum = np.concatenate([[0], np.cumsum(per_win)])
    start = np.asarray(offsets, np.int64)
    end = start + np.maximum(np.asarray(lengths, np.int64) - k + 1, 0)
    return csum[np.minimum(end, n)] - csum[np.minimum(start, n)]


def _reference_search_sample(fam_q, qs, db_lengths, db_flat, db_offsets, index, k, ft, band,
                             mc, threads):
    """Stage 1 and 2 of the reference search on a query sample with its
    compiled kernels: batch_score_queries_c (kmer_score.c:420-472) then
    batch_sw_align_c (sw_align.c:182-266) on the top ``mc`` candidates at
    ``band`` without hints (the GPU path's semantics).  Returns (seconds for
    each stage, (targets, counts, k-mer scores), (sims, scores, mask) per
    pair, pairs, the packed queries)."""
    from clustkit_amd.scoring import BLOSUM62
    from oracle import oracle
    qf = np.concatenate([fam_q.flat[fam_q.offsets[q]:fam_q.offsets[q] + fam_q.lengths[q]]
                         for q in qs])
    ql = fam_q.lengths[qs].astype(np.int32)
    qo = np.zeros(len(qs), np.int64)
    qo[1:] = np.cumsum(ql[:-1].astype(np.int64))
    ke = index.host_entries()
    t0 = time.perf_counter()
    tg, cn, ksc, _ = oracle.ref_kmer_score_batch(qf, qo, ql, k, 20, index.db.kmer_offsets, ke,
                                                 index.db.kmer_freqs, ft, len(db_lengths),
                                                 n_threads=threads, max_cands=mc)
    t1 = time.perf_counter()
    nd = len(db_lengths)
    flat = np.concatenate([db_flat, qf])
    offs = np.concatenate([db_offsets, qo + db_flat.size])
    lens = np.concatenate([db_lengths, ql]).astype(np.int32)
    rows = [np.stack([np.full(int(cn[i]), nd + i), tg[i, :cn[i]]], 1) for i in range(len(qs))]
    pairs = np.concatenate(rows).astype(np.int32) if rows else np.zeros((0, 2), np.int32)
    t2 = time.perf_counter()
    s, sc, mk = oracle.ref_sw_batch(pairs, flat, offs, lens, band, BLOSUM62, n_threads=threads)
    t3 = time.perf_counter()
    return (t1 - t0, t3 - t2), (tg, cn, ksc), (s, sc, mk), pairs, (qf, qo, ql)


def fast_leg(args, peaks):
    """Fast path at 200,000 sequences: each sequence is a query against the
    whole set; index build, k-mer Phase A/B, candidate pairs, banded SW and
    the top-10 selection on the device (search_kmer_index_device).  k = 3
    (dense, LDS-tiled) and k = 5 (sparse, exact k-mers).  Roofline of the
    k-mer stage: 4 B per Phase-A posting visited + 8 B per Phase-B posting
    (every posting is walked twice by the reference), event-timed stage.
    CPU baseline: the reference scorer and SW kernel on a 400-query sample,
    whose candidate lists must equal the device's."""
    import torch
    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import (DeviceKmerIndex, SequenceSet, auto_band_width,
                                          search_kmer_index_device)
    from clustkit_amd.workloads import SPECS, generate_families_streams
    from oracle import oracle
    w = WORKLOADS["fast200k"]
    t0 = time.perf_counter()
    fam = generate_families_streams(SPECS["fast200k"], workers=args.gen_workers)
    ds = SequenceSet.from_arrays(fam.flat, fam.offsets, fam.lengths)
    log(f"fast200k: {ds.num_sequences} seqs, {fam.flat.size} residues "
        f"({time.perf_counter() - t0:.1f}s host generation)")
    nq = ds.num_sequences
    band = auto_band_width(ds, ds)
    threads = cpu_threads()
    rows = []
    for k in (3, 5):
        torch.cuda.synchronize()
        tb = time.perf_counter()
        idx = DeviceKmerIndex.build(ds, k=k)
        torch.cuda.synchronize()
        build_s = time.perf_counter() - tb
        sub = SequenceSet.from_arrays(fam.flat, fam.offsets[:256], fam.lengths[:256])
        # exact k-mers at k = 5 too (kmer_score_thresh=0), the fast-path configuration
        search_kmer_index_device(idx, sub, threshold=w["threshold"], max_cands_per_query=64,
                                 kmer_score_thresh=0)
        tm = {}
        torch.cuda.synchronize()
        ts = time.perf_counter()
        res = search_kmer_index_device(idx, ds, threshold=w["threshold"], top_k=10,
                                       max_cands_per_query=64, timings=tm, as_arrays=True,
                                       kmer_score_thresh=0)
        torch.cuda.synchronize()
        wall = time.perf_counter() - ts
        ft = compute_freq_threshold(idx.db.kmer_freqs, nq, 95.0)
        post = int(_postings_visited(fam.flat, fam.offsets, fam.lengths, k, 20,
                                     idx.db.kmer_offsets, idx.db.kmer_freqs, ft).sum())
        kbytes = 12 * post
        ach = kbytes / tm["kmer_s"] / 1e9
        row = {"k": k, "index_postings": idx.n_postings, "index_tiled": idx.tiled,
               "index_build_s": round(build_s, 3), "band": band, "queries": nq,
               "search_s": round(wall, 3), "value": round(nq / wall, 1), "unit": "queries/s",
               "kmer_stage_s": round(tm["kmer_s"], 3), "sw_stage_s": round(tm["sw_s"], 3),
               "host_rank_s": round(tm["host_s"], 3), "candidates": res.num_candidates,
               "queries_with_hits": int(len(np.unique(res.hits["query"]))),
               "hits": int(len(res.hits["query"])),
               "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                            "traffic": None, "kernel": "k-mer stage (Phase A + selection + "
                            "Phase B), event-timed", "postings_visited": post,
                            "algorithmic_bytes": kbytes}}
        if not args.no_cpu and oracle.ref_kmer_available():
            rng = np.random.default_rng(7)
            qs = np.sort(rng.choice(nq, size=400, replace=False))
            (ck, cs), (rt, rc, rsc), _, _, (qf, qo, ql) = _reference_search_sample(
                fam, qs, fam.lengths, fam.flat, fam.offsets, idx, k, ft, band, 64, threads)
            dt_, dc_, ds_, _ = idx.score(
                torch.from_numpy(qf).cuda(), torch.from_numpy(qo).cuda(),
                torch.from_numpy(ql).cuda(), int(ql.max()), ft, max_cands=64)
            dc_ = dc_.cpu().numpy()
            dt_, ds_ = dt_.cpu().numpy(), ds_.cpu().numpy()
            same = bool(np.array_equal(dc_, rc) and all(
                np.array_equal(dt_[i, :rc[i]], rt[i, :rc[i]]) and
                np.array_equal(ds_[i, :rc[i]], rsc[i, :rc[i]]) for i in range(len(qs))))
            row["cpu_baseline"] = {
                "value": round(len(qs) / (ck + cs), 1), "unit": "queries/s", "cores": threads,
                "kind": "reference",
                "sample": f"{len(qs)} queries drawn uniformly: batch_score_queries_c "
                          f"{ck:.2f}s + batch_sw_align_c on their top-64 candidates "
                          f"(band {band}, no hints) {cs:.2f}s, {threads} threads",
                "kmer_stage_queries_per_s": round(len(qs) / ck, 1)}
            row["kmer_candidates_equal_reference_sample"] = same
        log(f"fast200k k={k}: {row}")
        rows.append(row)
        del idx
        torch.cuda.empty_cache()
    return {"workload": f"fast200k: {w['desc']}", "value": rows[0]["value"],
            "unit": "queries/s", "k_runs": rows}


def search_bench(n_db, rates, seconds, dev, gen_workers=16, with_cpu=True):
    """Search serving over an HBM-resident database: k-mer index built on the
    device, Poisson query arrivals at each offered rate, continuous batching
    (4 streams, batches of up to 512 under an HBM budget), k-mer Phase A/B +
    banded SW of the top 64 candidates + top-10 ranking per query.  Then the
    capacity (every query of a burst admitted at once) and the reference's
    compiled scorer + SW on a query sample (CPU baseline and parity)."""
    import dataclasses

    import torch

    from clustkit_amd.kmer_search import DeviceKmerIndex, SequenceSet, evalue_rank
    from clustkit_amd.scoring import pack_sequences
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
    burst = 4096
    total_q = sum(max(200, int(r * seconds)) for r in rates) + 64 + 32 + burst
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
    mix = []
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
        mix = done
    host_ms = {"launch": round(ex.host_launch_s / max(1, tot_batches) * 1e3, 3),
               "finish": round(ex.host_finish_s / max(1, tot_batches) * 1e3, 3)}
    # capacity: a burst of queries all admitted at t = 0
    reqs = [Request(rid=i, seq=q, arrival=0.0) for i, q in enumerate(take(burst))]
    sch = Scheduler(ex, budget, 512)
    t0 = time.perf_counter()
    serve(sch, reqs)
    cap_s = time.perf_counter() - t0
    capacity = {"queries": burst, "s": round(cap_s, 3), "queries_per_s": round(burst / cap_s, 1),
                "batches": sch.batches}
    log(f"search capacity: {capacity}")
    # each of 32 queries of the last mix, run alone, gives the same hits
    same = True
    for r in mix[:32]:
        d = serve(Scheduler(ex, budget, 1), [Request(rid=r.rid, seq=r.seq, arrival=0.0)])
        same &= d[0].hits == r.hits
    ok = [r for r in per_rate if r.get("n")]
    head = max(ok, key=lambda r: r["offered_qps"]) if ok else {}
    out = {
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
        "rates": per_rate, "capacity": capacity, "alone_equals_mix": bool(same),
    }
    if with_cpu and _ref_search_ready():
        # the burst's first 200 queries through the reference's compiled kernels
        sample = reqs[:200]
        qfam_flat, qfam_off, qfam_len = pack_sequences([r.seq for r in sample])
        qfam = dataclasses.replace(db, flat=qfam_flat, offsets=qfam_off, lengths=qfam_len)
        threads = cpu_threads()
        (ck, cs), (tg, cn, _), (s, sc, mk), _, (_, _, ql) = _reference_search_sample(
            qfam, np.arange(len(sample)), db.lengths, db.flat, db.offsets, idx, 3,
            ex.freq_thresh, ex.band, 64, threads)
        exact = True
        pos = 0
        for i, r in enumerate(sample):
            c = int(cn[i])
            t = tg[i, :c]
            keep = mk[pos:pos + c].astype(bool)
            rank = evalue_rank(sc[pos:pos + c][keep].astype(np.float32),
                               np.full(int(keep.sum()), ql[i]), db.lengths[t[keep]])
            order = np.lexsort((-rank, np.zeros(len(rank))))[:10]
            want = [(int(t[keep][o]), float(s[pos:pos + c][keep][o]), float(rank[o]))
                    for o in order]
            exact &= want == r.hits
            pos += c
        out["cpu_baseline"] = {
            "value": round(len(sample) / (ck + cs), 1), "unit": "queries/s", "cores": threads,
            "kind": "reference",
            "sample": f"first {len(sample)} queries of the capacity burst: "
                      f"batch_score_queries_c {ck:.2f}s + batch_sw_align_c on their top-64 "
                      f"candidates (band {ex.band}, no hints) {cs:.2f}s, {threads} threads"}
        out["hits_equal_reference_sample"] = bool(exact)
        log(f"search cpu baseline: {out['cpu_baseline']} exact={exact}")
    return out


def _ref_search_ready():
    from oracle import oracle
    return oracle.ref_available() and oracle.ref_kmer_available()


# ---------------------------------------------------------------------------
# Entry point
# ---------------------------------------------------------------------------

def _line(metric_suffix, leg, args, extra=None):
    line = {"metric": METRIC + metric_suffix, "value": leg["value"], "unit": leg["unit"],
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic", "config": {"workload": leg["workload"]}}
    line.update(extra or {})
    line["leg"] = leg
    return line


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
    ap.add_argument("--bands", default=None,
                    help="bands of the heavy workload (default 64 in the full run, "
                         "16,32,64,128 with --workload heavy)")
    ap.add_argument("--no-configs", action="store_true",
                    help="headline only: skip the small / heavy / fast200k legs")
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
    from clustkit_amd.device import DeviceSequences, alloc_outputs, sw_align_device
    from clustkit_amd.scoring import BLOSUM62

    _lib.require_gpu()
    props = _lib.device_props(0)
    peaks = device_peaks(props)
    log(f"device: {props}; peaks {peaks}")
    if args.workload == "heavy":
        bands = [int(b) for b in (args.bands or "16,32,64,128").split(",") if b]
        leg = heavy_leg(args, bands, peaks)
        print(json.dumps(_line(" (heavy-tail workload)", leg, args)), flush=True)
        return
    if args.workload == "fast200k":
        leg = fast_leg(args, peaks)
        print(json.dumps(_line(" (fast path)", leg, args)), flush=True)
        return
    if args.workload == "small":
        leg = small_leg(args, peaks)
        print(json.dumps(_line(" (small workload)", leg, args)), flush=True)
        return
    w = WORKLOADS[args.workload]
    band, thr = w["band"], w["threshold"]
    fam, sk, pairs_all, lshp, cand = build_workload(args.workload)
    if args.candidates_only:
        print(json.dumps({"candidates": cand, "workload": args.workload}), flush=True)
        return
    if not args.no_cpu:
        cand["cpu_oracle"] = candidates_cpu(fam, w, lshp, sk, pairs_all)
        log(f"candidate stage on the host: {cand['cpu_oracle']}")

    dev = torch.device("cuda:0")
    seqs = DeviceSequences.upload(fam.flat, fam.offsets, fam.lengths, dev)
    d_all = torch.from_numpy(np.ascontiguousarray(pairs_all, np.int32)).to(dev)
    d_sk = torch.from_numpy(sk.view(np.int64)).to(dev)
    d_pairs, jaccard_applied, jrate = jaccard_on_device(d_all, d_sk, thr)
    pairs = d_pairs.cpu().numpy()
    total_cells = int(pair_cells(pairs, fam.lengths, band).sum())
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
        base, cpu_idx, cpu_res = sw_cpu_baseline(fam, pairs, band)
        base["host"] = host_info()
        log(f"cpu baseline: {base}")

    tuned = None
    if not args.no_autotune:
        from clustkit_amd import autotune
        tuned = autotune.tune(seqs, d_pairs, band)
        log(f"autotune: {tuned}")

    # Batched form (the headline).
    b_ms, b_main, b_stats, b_phase = timed_sw(seqs, d_pairs, band, thr, args.steps,
                                              args.warmup, 0, out, "batched")
    res_b = {k: v.cpu().numpy() for k, v in out.items()}
    # Plain form.
    p_steps = args.plain_steps or max(1, args.steps // 2)
    p_ms, _, _, p_phase = timed_sw(seqs, d_pairs, band, thr, p_steps, min(args.warmup, 1),
                                   _lib.CK_SW_PLAIN, out, "plain")
    res_p = {k: v.cpu().numpy() for k, v in out.items()}

    plain_equal = all(np.array_equal(res_b[k], res_p[k]) for k in ("scores", "matches", "mask")) \
        and np.array_equal(res_b["sims"].view(np.uint32), res_p["sims"].view(np.uint32))
    bit_exact = _bit_exact(res_b, cpu_idx, cpu_res) if base is not None else None

    gcups = total_cells / (b_ms * 1e-3) / 1e9
    gcups_plain = total_cells / (p_ms * 1e-3) / 1e9
    # Dominant kernel: the persistent batched SW kernel, timed with HIP events on
    # the launch stream.  Algorithmic ops per launch = 26 x cells it evaluated.
    from clustkit_amd import autotune as _at
    lanes_cfg, _shift_cfg, pipe_cfg = _at.current()
    seg_lanes = lanes_cfg if lanes_cfg in (8, 16) else (8 if 1 <= band <= 32 else 16)
    kernel_name = "sw_pipe_kernel" if (seg_lanes == 16 and pipe_cfg) else "sw_wave_kernel"
    traffic, tsrc = None, None
    tfile = ROOT / TRAFFIC_FILE
    if args.workload == "mixed2m" and kernel_name == "sw_pipe_kernel" and tfile.exists():
        for ln in tfile.read_text().splitlines():
            if ln.startswith("traffic_bytes_per_launch:"):
                traffic = int(ln.split(":")[1])
                tsrc = (f"read from {TRAFFIC_FILE} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                        "of `bench.py --probe`, scaled by a 1 GiB copy), not measured in "
                        "this run")
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
        "roofline": sw_roofline(total_cells, b_main, peaks, kernel_name, traffic, tsrc),
        "cpu_baseline": base,
        "device": {**props, "peaks": peaks},
        "candidates": cand,
    }
    line["end_to_end"] = end_to_end(fam, pairs_all, sk, band, thr, total_cells)
    line["graph"] = graph_bench(fam, pairs, res_b, thr, args.steps)
    if not args.no_nw:
        line["nw"] = nw_bench(fam, pairs, seqs, d_pairs, band, thr, args.steps, args.warmup,
                              not args.no_cpu, peaks)
    del seqs, d_pairs, d_all, d_sk, out
    torch.cuda.empty_cache()
    if not args.no_configs:
        configs = {"small": small_leg(args, peaks)}
        torch.cuda.empty_cache()
        bands = [int(b) for b in (args.bands or "64").split(",") if b]
        configs["heavy"] = heavy_leg(args, bands, peaks)
        torch.cuda.empty_cache()
        configs["fast200k"] = fast_leg(args, peaks)
        torch.cuda.empty_cache()
        line["configs"] = configs
    if args.search_db > 0:
        rates = [float(x) for x in args.search_rates.split(",") if x]
        line["search"] = search_bench(args.search_db, rates, args.search_seconds, dev,
                                      args.gen_workers, with_cpu=not args.no_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
