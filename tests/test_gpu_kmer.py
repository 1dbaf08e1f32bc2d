"""GPU parity of k-mer index scoring (Phase A/B, top-K) against the reference
golden vectors and the oracle, in both counting modes (LDS tiles, global
atomics), plus the device-resident form."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

KMER_FILES = sorted(GOLDEN.glob("kmer_*.npz"))


def _load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


# Phase B forms: target-major over the database's window codes, over its
# residues, and posting-major.
PHASE_B_FORMS = ((True, True), (True, False), (False, False))


def _rows_equal(a, b, counts, check_diags=True):
    tg, cn, sc, dg = a
    gt, gc, gs, gd = b
    assert np.array_equal(cn, gc)
    for q in range(len(cn)):
        n = cn[q]
        assert np.array_equal(tg[q, :n], gt[q, :n]), q
        assert np.array_equal(sc[q, :n], gs[q, :n]), q
        if check_diags:
            assert np.array_equal(dg[q, :n], gd[q, :n]), q


@pytest.mark.parametrize("path", KMER_FILES, ids=lambda p: p.stem)
def test_gpu_kmer_matches_reference_golden(path):
    from clustkit_amd.kmer_index import build_kmer_index
    from clustkit_amd.kmer_search import batch_score_queries
    g = _load(path)
    k = int(g["k"])
    ko, ke, kf = build_kmer_index(g["db_flat"], g["db_offsets"], g["db_lengths"], k)
    mt, md, w, mc, ta = (int(x) for x in g["params"])
    print(f"kmer golden {path.stem}", flush=True)
    got = batch_score_queries(g["q_flat"], g["q_offsets"], g["q_lengths"], k, 20, ko, ke, None,
                              kf, int(g["freq_thresh"]), len(g["db_lengths"]), mt, md, w, mc, ta)
    _rows_equal(got, (g["targets"], g["counts"], g["scores"], g["diags"]), g["counts"],
                check_diags=md > 1)


@pytest.fixture(scope="module")
def search_db():
    from clustkit_amd.kmer_search import SequenceSet, build_database_index
    from clustkit_amd.workloads import FamilySpec, generate_families, mutate
    db = generate_families(FamilySpec(n_roots=2000, n_desc=9, median=300, clip_lo=30,
                                      clip_hi=1200, id_lo=0.3, id_hi=0.95, seed=31))
    ds = SequenceSet.from_arrays(db.flat, db.offsets, db.lengths)
    rng = np.random.default_rng(3)
    qs = []
    for r in rng.choice(2000, 48, replace=False):
        s = db.flat[db.offsets[r * 10]:db.offsets[r * 10] + db.lengths[r * 10]]
        qs.append(mutate(rng, s, float(rng.uniform(0.3, 0.9))))
    qs += [rng.integers(0, 20, int(n)).astype(np.uint8) for n in rng.integers(50, 2000, 16)]
    from clustkit_amd.scoring import pack_sequences
    qf, qo, ql = pack_sequences(qs)
    return ds, build_database_index(ds, k=3), SequenceSet.from_arrays(qf, qo, ql)


@pytest.mark.parametrize("kw", [dict(), dict(max_cands=64), dict(phase_a_topk=300, max_cands=40),
                                dict(min_diag_hits=1, max_cands=64)],
                         ids=["default", "mc64", "topk300", "phase_a"])
def test_gpu_kmer_device_modes_match_oracle(search_db, oracle_mod, kw):
    import torch
    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import DeviceKmerIndex
    ds, idx, qset = search_db
    nd = ds.num_sequences
    ft = compute_freq_threshold(idx.kmer_freqs, nd, 95.0)
    want = oracle_mod.kmer_score_batch(qset.flat_sequences, qset.offsets, qset.lengths, 3, 20,
                                       idx.kmer_offsets, idx.kmer_entries, idx.kmer_freqs, ft, nd,
                                       n_threads=8, **kw)
    dev = torch.device("cuda:0")
    qf = torch.from_numpy(qset.flat_sequences).to(dev)
    qo = torch.from_numpy(qset.offsets).to(dev)
    ql = torch.from_numpy(qset.lengths).to(dev)
    for tiled in (True, False):
        di = DeviceKmerIndex(idx, dev, tiled=tiled)
        for tm, wc in PHASE_B_FORMS:
            print(f"kmer device tiled={tiled} target_major={tm} window_codes={wc} {kw}",
                  flush=True)
            out = di.score(qf, qo, ql, int(qset.lengths.max()), ft, target_major=tm,
                           window_codes=wc, **kw)
            torch.cuda.synchronize()
            got = tuple(x.cpu().numpy() for x in out)
            _rows_equal(got, want, want[1], check_diags=kw.get("min_diag_hits", 2) > 1)


@pytest.mark.parametrize("mode,k", [("protein", 3), ("protein", 2), ("nucleotide", 5)])
def test_gpu_kmer_target_major_phase_b_edges(oracle_mod, mode, k):
    """Target-major Phase B (survivors' windows against the query's k-mer
    table) against the oracle and the posting-major kernels: unknown residues
    in queries and targets, sequences shorter than k, an empty sequence,
    unaligned sequence starts, bin widths down to 1 (several histogram passes
    per survivor) and on both sides of the limits of the multiply-high bin
    division (3 .. 256), and queries longer than the histogram pass."""
    import torch
    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import DeviceKmerIndex, SequenceSet, build_database_index
    from clustkit_amd.scoring import pack_sequences
    from clustkit_amd.workloads import mutate
    alpha = 20 if mode == "protein" else 4
    rng = np.random.default_rng(17 + k)
    roots = [rng.integers(0, alpha, int(n)).astype(np.uint8) for n in rng.integers(40, 700, 60)]
    seqs = []
    for r in roots:
        seqs.append(r)
        for _ in range(5):
            x = mutate(rng, r, float(rng.uniform(0.4, 0.95))) % alpha
            x = x.astype(np.uint8)
            if rng.random() < 0.3:
                x[rng.integers(0, len(x), max(1, len(x) // 15))] = alpha + 3
            seqs.append(x)
    seqs += [np.zeros(0, np.uint8), np.array([1], np.uint8), np.full(6, alpha, np.uint8),
             rng.integers(0, alpha, 2600).astype(np.uint8)]
    flat, offs, lens = pack_sequences(seqs)
    ds = SequenceSet.from_arrays(flat, offs, lens)
    idx = build_database_index(ds, mode=mode, k=k)
    qs = [seqs[i].copy() for i in rng.choice(len(roots) * 6, 30, replace=False)]
    qs += [seqs[-1][:2400].copy(), np.array([2, 1], np.uint8), seqs[3][::-1].copy()]
    qs[0][::7] = alpha + 1
    qf_, qo_, ql_ = pack_sequences(qs)
    nd = ds.num_sequences
    ft = compute_freq_threshold(idx.kmer_freqs, nd, 95.0)
    dev = torch.device("cuda:0")
    qf = torch.from_numpy(qf_).to(dev)
    qo = torch.from_numpy(qo_).to(dev)
    ql = torch.from_numpy(ql_).to(dev)
    di = DeviceKmerIndex(idx, dev)
    for w in (1, 3, 10, 256, 257):
        kw = dict(diag_bin_width=w, max_cands=50, phase_a_topk=200)
        want = oracle_mod.kmer_score_batch(qf_, qo_, ql_, k, alpha, idx.kmer_offsets,
                                           idx.kmer_entries, idx.kmer_freqs, ft, nd,
                                           n_threads=8, **kw)
        for tm, wc in PHASE_B_FORMS:
            print(f"kmer target-major edges {mode} k={k} w={w} target_major={tm} "
                  f"window_codes={wc}", flush=True)
            out = di.score(qf, qo, ql, int(ql_.max()), ft, target_major=tm, window_codes=wc,
                           **kw)
            torch.cuda.synchronize()
            _rows_equal(tuple(x.cpu().numpy() for x in out), want, want[1])


@pytest.mark.parametrize("mode,k", [("protein", 3), ("protein", 1), ("nucleotide", 5)])
def test_gpu_db_window_codes(mode, k):
    """ck_kmer_db_codes_device against the code definition (code + 1, the
    code being the k residues in base alpha, first most significant; 0 for a
    window holding a residue >= alpha and for the padding of each sequence's
    run to a multiple of 8):
    empty and shorter-than-k sequences, unknown residues, codes 200-255."""
    import torch
    from clustkit_amd.kmer_search import DeviceKmerIndex, SequenceSet, build_database_index
    from clustkit_amd.scoring import pack_sequences
    alpha = 20 if mode == "protein" else 4
    rng = np.random.default_rng(5 + k)
    seqs = [rng.integers(0, alpha, int(n)).astype(np.uint8) for n in rng.integers(0, 300, 200)]
    for s_ in seqs[::5]:
        if len(s_):
            s_[rng.integers(0, len(s_), max(1, len(s_) // 10))] = rng.choice([alpha, 25, 200, 255])
    seqs += [np.zeros(0, np.uint8), np.arange(k - 1, dtype=np.uint8) % alpha,
             np.full(9, alpha - 1, np.uint8)]
    flat, offs, lens = pack_sequences(seqs)
    ds = SequenceSet.from_arrays(flat, offs, lens)
    di = DeviceKmerIndex(build_database_index(ds, mode=mode, k=k), torch.device("cuda:0"))
    print(f"db window codes {mode} k={k}", flush=True)
    torch.cuda.synchronize()
    codes = di.db_codes.cpu().numpy()
    coff = di.db_code_off.cpu().numpy()
    assert coff[0] == 0 and np.all(coff % 8 == 0)
    for s_, a, b in zip(seqs, coff[:-1], coff[1:]):
        nw = max(len(s_) - k + 1, 0)
        assert b - a == (nw + 7) // 8 * 8
        want = np.zeros(b - a, np.int64)
        for p in range(nw):
            w = s_[p:p + k].astype(np.int64)
            if np.all(w < alpha):
                want[p] = int(np.polyval(w, alpha)) + 1
        assert np.array_equal(codes[a:b].astype(np.int64), want)


def test_gpu_search_kmer_index_end_to_end(search_db, oracle_mod):
    """Hits of the GPU search equal those recomputed from oracle stages."""
    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import (_merge_sequences_for_alignment, evalue_rank,
                                          search_kmer_index)
    from clustkit_amd.scoring import BLOSUM62
    ds, idx, qset = search_db
    res = search_kmer_index(idx, qset, threshold=0.3, top_k=5, max_cands_per_query=64)
    nd = ds.num_sequences
    ft = compute_freq_threshold(idx.kmer_freqs, nd, 95.0)
    tg, cn, sc, dg = oracle_mod.kmer_score_batch(qset.flat_sequences, qset.offsets,
                                                 qset.lengths, 3, 20, idx.kmer_offsets,
                                                 idx.kmer_entries, idx.kmer_freqs, ft, nd,
                                                 max_cands=64)
    merged = _merge_sequences_for_alignment(qset, ds)
    bw = max(20, int(int(np.percentile(np.concatenate([qset.lengths, ds.lengths]), 95)) * 0.3))
    for q in range(qset.num_sequences):
        t = tg[q, :cn[q]]
        pairs = np.stack([np.full(len(t), q), t + qset.num_sequences], 1).astype(np.int32)
        o = oracle_mod.sw_batch(pairs, merged["flat_sequences"], merged["offsets"],
                                merged["lengths"], bw, BLOSUM62)
        keep = o["scores"] > 0
        rank = evalue_rank(o["scores"][keep].astype(np.float32),
                           np.full(keep.sum(), qset.lengths[q]), ds.lengths[t[keep]])
        order = np.lexsort((-rank, np.zeros(len(rank))))[:5]
        want = [(int(t[keep][i]), float(o["sims"][keep][i])) for i in order]
        got = [(int(h.target_id[1:]), h.identity) for h in res.hits[q]]
        assert got == want, q


@pytest.mark.parametrize("c_scoring", [True, False])
def test_gpu_search_kmer_index_c_sw_branch(search_db, oracle_mod, c_scoring):
    """use_c_sw=True: SW at c_sw_band_width, steered by the Phase-B diagonal
    hints only when the C scorer produced them (use_c_scoring,
    kmer_index.py:1627-1653, 2255-2256); hits equal those recomputed from
    oracle stages."""
    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import (_merge_sequences_for_alignment, evalue_rank,
                                          search_kmer_index)
    from clustkit_amd.scoring import BLOSUM62
    ds, idx, qset = search_db
    print(f"search c-sw branch use_c_scoring={c_scoring}", flush=True)
    res = search_kmer_index(idx, qset, threshold=0.3, top_k=5, max_cands_per_query=64,
                            use_c_sw=True, c_sw_band_width=30, use_c_scoring=c_scoring)
    nd = ds.num_sequences
    ft = compute_freq_threshold(idx.kmer_freqs, nd, 95.0)
    tg, cn, sc, dg = oracle_mod.kmer_score_batch(qset.flat_sequences, qset.offsets,
                                                 qset.lengths, 3, 20, idx.kmer_offsets,
                                                 idx.kmer_entries, idx.kmer_freqs, ft, nd,
                                                 max_cands=64)
    merged = _merge_sequences_for_alignment(qset, ds)
    for q in range(qset.num_sequences):
        t = tg[q, :cn[q]]
        pairs = np.stack([np.full(len(t), q), t + qset.num_sequences], 1).astype(np.int32)
        hints = dg[q, :cn[q]].astype(np.int32) if c_scoring else None
        o = oracle_mod.sw_batch(pairs, merged["flat_sequences"], merged["offsets"],
                                merged["lengths"], 30, BLOSUM62, hints=hints)
        keep = o["scores"] > 0
        rank = evalue_rank(o["scores"][keep].astype(np.float32),
                           np.full(keep.sum(), qset.lengths[q]), ds.lengths[t[keep]])
        order = np.lexsort((-rank, np.zeros(len(rank))))[:5]
        want = [(int(t[keep][i]), float(o["sims"][keep][i])) for i in order]
        got = [(int(h.target_id[1:]), h.identity) for h in res.hits[q]]
        assert got == want, q


def test_gpu_search_kmer_index_global_alignment(search_db, oracle_mod):
    """local_alignment=False (kmer_index.py:2281-2290): stage 2 is banded
    global identity with BLOSUM62, pass = identity >= threshold, ranked by
    the global score; hits equal those recomputed from oracle stages."""
    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import _merge_sequences_for_alignment, search_kmer_index
    from clustkit_amd.scoring import BLOSUM62
    ds, idx, qset = search_db
    thr = 0.3
    res = search_kmer_index(idx, qset, threshold=thr, top_k=5, max_cands_per_query=64,
                            local_alignment=False)
    nd = ds.num_sequences
    ft = compute_freq_threshold(idx.kmer_freqs, nd, 95.0)
    tg, cn, sc, dg = oracle_mod.kmer_score_batch(qset.flat_sequences, qset.offsets,
                                                 qset.lengths, 3, 20, idx.kmer_offsets,
                                                 idx.kmer_entries, idx.kmer_freqs, ft, nd,
                                                 max_cands=64)
    merged = _merge_sequences_for_alignment(qset, ds)
    bw = max(20, int(int(np.percentile(np.concatenate([qset.lengths, ds.lengths]), 95)) * 0.3))
    n_hits = 0
    for q in range(qset.num_sequences):
        t = tg[q, :cn[q]]
        pairs = np.stack([np.full(len(t), q), t + qset.num_sequences], 1).astype(np.int32)
        sims, scores, mask = oracle_mod.nw_batch(pairs, merged["flat_sequences"],
                                                 merged["offsets"], merged["lengths"], thr, bw,
                                                 BLOSUM62)
        rank = scores[mask].astype(np.float32)
        order = np.lexsort((-rank, np.zeros(len(rank))))[:5]
        want = [(int(t[mask][i]), float(sims[mask][i]), float(rank[i])) for i in order]
        got = [(int(h.target_id[1:]), h.identity, h.score) for h in res.hits[q]]
        assert got == want, q
        n_hits += len(got)
    assert n_hits > 0


@pytest.mark.parametrize("kw", [dict(max_cands_per_query=64, top_k=5),
                                dict(max_cands_per_query=2000, top_k=10, chunk=7),
                                dict(max_cands_per_query=40, min_diag_hits=1, top_k=3),
                                dict(max_cands_per_query=64, top_k=5, evalue_normalize=False),
                                dict(max_cands_per_query=64, top_k=5, local_alignment=False)])
def test_gpu_device_resident_search_equals_search(search_db, kw):
    """search_kmer_index_device (index, scoring, pairs and SW in HBM, host
    ranking only) returns the same hits as search_kmer_index, which the test
    above pins to the oracle stages; also with an index built on device."""
    import torch

    from clustkit_amd.kmer_search import (DeviceKmerIndex, search_kmer_index,
                                          search_kmer_index_device)
    ds, idx, qset = search_db
    print(f"device-resident search {kw}", flush=True)
    kw_host = {k: v for k, v in kw.items() if k != "chunk"}
    want = search_kmer_index(idx, qset, threshold=0.3, **kw_host)
    for dev_idx in (DeviceKmerIndex(idx, torch.device("cuda:0")), DeviceKmerIndex.build(ds, k=3)):
        tm = {}
        got = search_kmer_index_device(dev_idx, qset, threshold=0.3, timings=tm, **kw)
        assert got.num_candidates == want.num_candidates
        for q in range(qset.num_sequences):
            a = [(h.target_id, h.identity, h.score) for h in got.hits[q]]
            b = [(h.target_id, h.identity, h.score) for h in want.hits[q]]
            assert a == b, q
        assert tm["kmer_s"] > 0 and tm["sw_s"] > 0
        for dev_topk in (True, False):
            arr = search_kmer_index_device(dev_idx, qset, threshold=0.3, as_arrays=True,
                                           device_topk=dev_topk, **kw).hits
            flat = [(f"s{t}", i, s) for t, i, s in zip(arr["target"].tolist(),
                                                      arr["identity"].tolist(),
                                                      arr["score"].tolist())]
            assert flat == [(h.target_id, h.identity, h.score) for hq in want.hits for h in hq]
            assert list(arr["query"]) == [q for q, hq in enumerate(want.hits) for _ in hq]


@pytest.mark.parametrize("mc,evalue", [(64, 1), (100, 1), (64, 0), (100, 2)])
def test_gpu_device_topk_matches_stable_lexsort(mc, evalue):
    """ck_search_topk_device against the reference selection
    (search.py:727-800: stable sort by (query, -rank), first top_k) with each
    rank mode (identity, kmer_index.py:2318-2322's rank computed by NumPy,
    the score as float32): many equal ranks
    (repeated scores and lengths), masked and uncounted slots, an empty
    query, a query with fewer passing slots than top_k."""
    import ctypes

    import torch

    from clustkit_amd import _lib
    from clustkit_amd.kmer_search import evalue_rank
    lib = _lib.require_gpu()
    rng = np.random.default_rng(mc + evalue)
    nq, nd, top_k = 6, 50, 10
    tg = rng.integers(0, nd, size=(nq, mc)).astype(np.int32)
    cn = np.array([mc, 0, 5, mc, mc // 2, 3], np.int32)
    sims = rng.choice(np.float32([0.25, 0.5, 0.75, 1.0]), size=(nq, mc)).astype(np.float32)
    scores = rng.integers(1, 6, size=(nq, mc)).astype(np.int32) * 10
    mask = (rng.random((nq, mc)) < 0.8).astype(np.uint8)
    qlens = rng.choice([100, 200], size=nq).astype(np.int32)
    tlens = rng.choice([100, 150, 200], size=nd).astype(np.int32)
    with np.errstate(divide="ignore"):
        logtab = np.log(np.arange(1 << 16, dtype=np.float32))
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    o_t = torch.zeros((nq, top_k), dtype=torch.int32, device="cuda")
    o_i = torch.zeros((nq, top_k), dtype=torch.float32, device="cuda")
    o_r = torch.zeros((nq, top_k), dtype=torch.float32, device="cuda")
    o_n = torch.zeros(nq, dtype=torch.int32, device="cuda")
    st = torch.zeros(1, dtype=torch.int32, device="cuda")
    ins = [d(tg), d(cn), d(sims), d(scores), d(mask), d(qlens), d(tlens), d(logtab)]
    _lib.check(lib.ck_search_topk_device(
        *[x.data_ptr() for x in ins], ins[-1].numel(), nq, mc, top_k, evalue,
        o_t.data_ptr(), o_i.data_ptr(), o_r.data_ptr(), o_n.data_ptr(), st.data_ptr(),
        torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert int(st.item()) == 0
    got_n = o_n.cpu().numpy()
    for q in range(nq):
        slots = np.flatnonzero((np.arange(mc) < cn[q]) & (mask[q] == 1))
        if evalue == 1:      # CK_RANK_EVALUE
            rank = evalue_rank(scores[q, slots], np.full(len(slots), qlens[q]),
                               tlens[tg[q, slots]])
        elif evalue == 2:    # CK_RANK_SCORE
            rank = scores[q, slots].astype(np.float32)
        else:                # CK_RANK_IDENTITY
            rank = sims[q, slots]
        order = np.argsort(-rank, kind="stable")[:top_k]
        assert got_n[q] == len(order), q
        k = len(order)
        assert np.array_equal(o_t[q, :k].cpu().numpy(), tg[q, slots[order]]), q
        assert np.array_equal(o_i[q, :k].cpu().numpy().view(np.uint32),
                              sims[q, slots[order]].view(np.uint32)), q
        assert np.array_equal(o_r[q, :k].cpu().numpy().view(np.uint32),
                              np.asarray(rank[order], np.float32).view(np.uint32)), q


@pytest.mark.parametrize("k", [2, 3])
def test_gpu_device_index_build_matches_host_build(k):
    """ck_kmer_index_build_device reproduces build_kmer_index array for
    array (offsets, frequencies, postings in (sequence, position) order) and
    the tile-offset table, on a database spanning two 8192-target tiles with
    unknown residues, sequences shorter than k and an empty sequence."""
    from clustkit_amd.kmer_search import (DeviceKmerIndex, SequenceSet, build_database_index)
    from clustkit_amd.scoring import pack_sequences
    from clustkit_amd.workloads import FamilySpec, generate_families
    db = generate_families(FamilySpec(n_roots=1000, n_desc=9, median=120, clip_lo=1,
                                      clip_hi=500, seed=23))
    seqs = [db.flat[o:o + n].copy() for o, n in zip(db.offsets, db.lengths)]
    rng = np.random.default_rng(5)
    for s in rng.choice(len(seqs), 300, replace=False):
        x = seqs[s]
        x[rng.integers(0, len(x), max(1, len(x) // 20))] = rng.choice([20, 23, 24], 1)[0]
    seqs += [np.zeros(0, np.uint8), np.array([4], np.uint8), np.array([20, 20, 20, 20], np.uint8)]
    flat, offs, lens = pack_sequences(seqs)
    ds = SequenceSet.from_arrays(flat, offs, lens)
    print(f"device index build k={k} seqs={len(lens)}", flush=True)
    host = build_database_index(ds, k=k)
    dev = DeviceKmerIndex.build(ds, k=k, tiled=True)
    assert np.array_equal(dev.kmer_offsets.cpu().numpy(), host.kmer_offsets)
    assert np.array_equal(dev.kmer_freqs.cpu().numpy(), host.kmer_freqs)
    assert np.array_equal(dev.db.kmer_freqs, host.kmer_freqs)
    assert dev.n_postings == len(host.kmer_entries)
    assert np.array_equal(dev.host_entries(), host.kmer_entries)
    up = DeviceKmerIndex(host, tiled=True)
    assert np.array_equal(dev.tile_off.cpu().numpy(), up.tile_off.cpu().numpy())


def test_gpu_index_from_disk_equals_built(search_db, tmp_path):
    """A database saved in the reference's layout and loaded into HBM
    (DeviceKmerIndex.from_disk) searches exactly like the index built on the
    device; a saved database without its k-mer arrays gets them built on the
    device on load."""
    import os
    from clustkit_amd.kmer_search import (DeviceKmerIndex, save_database,
                                          search_kmer_index_device)
    ds, idx, qset = search_db
    save_database(idx, tmp_path / "db")
    built = DeviceKmerIndex.build(ds, k=3)
    loaded = DeviceKmerIndex.from_disk(tmp_path / "db")
    a = search_kmer_index_device(built, qset, threshold=0.3, top_k=5, max_cands_per_query=64)
    b = search_kmer_index_device(loaded, qset, threshold=0.3, top_k=5, max_cands_per_query=64)
    assert [[(h.target_id, h.identity, h.score) for h in q] for q in a.hits] == \
        [[(h.target_id, h.identity, h.score) for h in q] for q in b.hits]
    for name in ("kmer_offsets", "kmer_entries", "kmer_freqs", "kmer_entries_compact"):
        os.remove(tmp_path / "db" / f"{name}.npy")
    rebuilt = DeviceKmerIndex.from_disk(tmp_path / "db")
    assert np.array_equal(rebuilt.host_entries(), built.host_entries())
    c = search_kmer_index_device(rebuilt, qset, threshold=0.3, top_k=5, max_cands_per_query=64)
    assert [[(h.target_id, h.identity) for h in q] for q in c.hits] == \
        [[(h.target_id, h.identity) for h in q] for q in a.hits]


@pytest.fixture(scope="module")
def similar_db():
    from clustkit_amd.kmer_search import SequenceSet, build_database_index
    from clustkit_amd.scoring import pack_sequences
    from clustkit_amd.workloads import FamilySpec, generate_families, mutate
    db = generate_families(FamilySpec(n_roots=150, n_desc=9, median=250, clip_lo=20,
                                      clip_hi=700, id_lo=0.4, id_hi=0.95, seed=77))
    flat = db.flat.copy()
    flat[::97] = 20                                   # unknown residues
    ds = SequenceSet.from_arrays(flat, db.offsets, db.lengths)
    rng = np.random.default_rng(8)
    qs = [mutate(rng, db.flat[db.offsets[r * 10]:db.offsets[r * 10] + db.lengths[r * 10]], 0.75)
          for r in rng.choice(150, 24, replace=False)]
    qs += [rng.integers(0, 20, int(n)).astype(np.uint8) for n in (3, 5, 60, 400)]
    qf, qo, ql = pack_sequences(qs)
    return ds, build_database_index(ds, k=5), SequenceSet.from_arrays(qf, qo, ql)


@pytest.mark.parametrize("T,mc", [(13, 2000), (13, 12), (11, 64), (20, 64)])
def test_gpu_similar_kmer_scoring_matches_oracle(similar_db, oracle_mod, T, mc):
    """BLOSUM-similar 5-mer scoring on the GPU (ck_kmer_score_similar_device,
    kmer_index.py:1186-1330) equals the oracle: candidate ids in order, their
    count, and the weighted counts; device-resident and host-array forms."""
    import torch
    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import DeviceKmerIndex, batch_score_queries_similar
    from clustkit_amd.scoring import BLOSUM62
    ds, idx, qset = similar_db
    nd = ds.num_sequences
    ft = compute_freq_threshold(idx.kmer_freqs, nd, 95.0)
    want = oracle_mod.kmer_similar_batch(qset.flat_sequences, qset.offsets, qset.lengths,
                                         idx.kmer_offsets, idx.kmer_entries, idx.kmer_freqs, ft,
                                         nd, BLOSUM62, T, min_total_hits=2, max_cands=mc,
                                         n_threads=8)
    print(f"similar k-mers T={T} mc={mc}", flush=True)
    dev = torch.device("cuda:0")
    di = DeviceKmerIndex(idx, dev)
    qf = torch.from_numpy(qset.flat_sequences).to(dev)
    qo = torch.from_numpy(qset.offsets).to(dev)
    ql = torch.from_numpy(qset.lengths).to(dev)
    out = di.score_similar(qf, qo, ql, int(qset.lengths.max()), ft, T, max_cands=mc)
    torch.cuda.synchronize()
    tg, cn, sc, dg = (x.cpu().numpy() for x in out)
    h_tg, h_cn = batch_score_queries_similar(qset.flat_sequences, qset.offsets, qset.lengths,
                                             idx.kmer_offsets, idx.kmer_entries, idx.kmer_freqs,
                                             ft, nd, 2, mc, BLOSUM62, T,
                                             db_lengths=ds.lengths)
    assert np.array_equal(cn, want[1]) and np.array_equal(h_cn, want[1])
    assert cn.sum() > 0
    for q in range(len(cn)):
        n = cn[q]
        assert np.array_equal(tg[q, :n], want[0][q, :n]), q
        assert np.array_equal(h_tg[q, :n], want[0][q, :n]), q
        assert np.array_equal(sc[q, :n], np.minimum(want[2][q, :n], 32767)), q
        assert not dg[q, :n].any()


def test_gpu_search_kmer_index_similar_default(similar_db, oracle_mod):
    """search_kmer_index on a k = 5 protein index with the reference's default
    kmer_score_thresh (13) scores BLOSUM-similar 5-mers (kmer_index.py:1576,
    1654-1668), then banded SW without hints; hits equal those recomputed
    from oracle stages.  search_kmer_index_device agrees."""
    from clustkit_amd.kmer_index import compute_freq_threshold
    from clustkit_amd.kmer_search import (DeviceKmerIndex, _merge_sequences_for_alignment,
                                          evalue_rank, search_kmer_index,
                                          search_kmer_index_device)
    from clustkit_amd.scoring import BLOSUM62
    ds, idx, qset = similar_db
    res = search_kmer_index(idx, qset, threshold=0.3, top_k=5, max_cands_per_query=64)
    nd = ds.num_sequences
    ft = compute_freq_threshold(idx.kmer_freqs, nd, 95.0)
    tg, cn, _ = oracle_mod.kmer_similar_batch(qset.flat_sequences, qset.offsets, qset.lengths,
                                              idx.kmer_offsets, idx.kmer_entries,
                                              idx.kmer_freqs, ft, nd, BLOSUM62, 13,
                                              min_total_hits=2, max_cands=64)
    merged = _merge_sequences_for_alignment(qset, ds)
    bw = max(20, int(int(np.percentile(np.concatenate([qset.lengths, ds.lengths]), 95)) * 0.3))
    for q in range(qset.num_sequences):
        t = tg[q, :cn[q]]
        pairs = np.stack([np.full(len(t), q), t + qset.num_sequences], 1).astype(np.int32)
        o = oracle_mod.sw_batch(pairs, merged["flat_sequences"], merged["offsets"],
                                merged["lengths"], bw, BLOSUM62)
        keep = o["scores"] > 0
        rank = evalue_rank(o["scores"][keep].astype(np.float32),
                           np.full(keep.sum(), qset.lengths[q]), ds.lengths[t[keep]])
        order = np.lexsort((-rank, np.zeros(len(rank))))[:5]
        want = [(int(t[keep][i]), float(o["sims"][keep][i])) for i in order]
        got = [(int(h.target_id[1:]), h.identity) for h in res.hits[q]]
        assert got == want, q
    dres = search_kmer_index_device(DeviceKmerIndex(idx), qset, threshold=0.3, top_k=5,
                                    max_cands_per_query=64, band_width=bw)
    for q in range(qset.num_sequences):
        assert [(h.target_id, h.identity) for h in dres.hits[q]] == \
               [(h.target_id, h.identity) for h in res.hits[q]], q
