"""The benchmark configurations on the GPU, each checked against the oracle.

* C1 (2,000 proteins, k=5 / s=128 sketches, 146 x 1 LSH, t=0.5, band 100) at
  full size, through the call the clustering pipeline makes
  (pipeline.py:675-690 -> pairwise.py:1093-1148): Jaccard pre-filter,
  stable min-index sort, SW, pass mask -- for the GPU branch (device "0")
  and the C and Numba branches (device "cpu").
* C2 (the mixed-length workload) at 5,000 proteins, t=0.3, band 100, with
  the Jaccard pre-filter applied.
* C3 (heavy-tailed lengths: 10% Pareto(1.5) 1000-8000) at 5,000 proteins
  plus one planted near-identical 7,900-residue family, bands 16 / 32 / 64 /
  128 in one launch each, with wide-score re-runs asserted.
* C5 (search serving) at a 20,000-protein database: every request's hits
  from the continuous-batching executor equal k-mer oracle -> SW oracle at
  the executor's band -> E-value rank -> stable top-k.

The oracle composition restates pairwise.py:1093-1208 on top of
oracle.jaccard_prefilter, oracle.sw_batch and oracle.length_ratio_pass."""

from __future__ import annotations

import numpy as np
import pytest

from clustkit_amd.scoring import BLOSUM62

pytestmark = pytest.mark.gpu


def _candidates(fam, k, s, t):
    from clustkit_amd.lsh import auto_lsh_params, lsh_candidates
    from clustkit_amd.sketch import compute_sketches
    sk = compute_sketches(None, fam.lengths, k, s, "protein", 42, device="0",
                          flat_sequences=fam.flat, offsets=fam.offsets)
    p = auto_lsh_params(t, "medium", k)
    pairs = lsh_candidates(sk, p["num_tables"], p["num_bands"], 42, device="0")
    return sk, pairs, p


def _oracle_pipeline(O, fam, pairs, sk, t, band, branch, order=None):
    """pairwise.py:1093-1208 from oracle stages."""
    p = np.ascontiguousarray(pairs, np.int32)
    if sk is not None:
        jm = O.jaccard_prefilter(p, sk, t)
        if jm is not None:
            p = p[jm]
    p = p[np.argsort(np.minimum(p[:, 0], p[:, 1]), kind="mergesort")]
    order = order or ("max_length" if branch == "gpu" else "min_index")
    L = fam.lengths
    if order == "max_length":
        p = p[np.argsort(np.maximum(L[p[:, 0]], L[p[:, 1]]), kind="mergesort")]
    o = O.sw_batch(p, fam.flat, fam.offsets, L, band, BLOSUM62)
    t32 = np.float32(t)
    if branch == "gpu":
        keep = O.length_ratio_pass(p, L, t) & (o["scores"] > 0) & (o["sims"] >= t32)
    elif branch == "c":
        keep = o["sims"] >= t32
    else:
        keep = (o["scores"] > 0) & (o["sims"] >= t32)
    return p[keep], o["sims"][keep], p, o


def _same(got, want, label):
    gp, gs = got
    wp, ws = want
    assert gp.dtype == np.int32 and gs.dtype == np.float32, label
    assert np.array_equal(gp, wp), f"{label}: pair lists differ ({len(gp)} vs {len(wp)})"
    assert np.array_equal(gs.view(np.uint32), ws.view(np.uint32)), f"{label}: identities differ"


@pytest.fixture(scope="module")
def c1():
    from clustkit_amd.workloads import SPECS, generate_families
    fam = generate_families(SPECS["small"])
    sk, pairs, lshp = _candidates(fam, 5, 128, 0.5)
    return fam, sk, pairs, lshp


def test_gpu_c1_pipeline_call_full_size(c1, oracle_mod):
    """C1 at full size: the pipeline's call for each reference branch equals
    the oracle composition; the SW stage also equals the reference kernel
    compiled from its own source on every pair that enters it."""
    from clustkit_amd.pairwise import compute_pairwise_alignment
    fam, sk, pairs, lshp = c1
    assert len(fam.lengths) == 2000 and lshp == {"num_tables": 146, "num_bands": 1}
    t, band = 0.5, 100
    kw = dict(band_width=band, sketches=sk, flat_sequences=fam.flat, offsets=fam.offsets,
              use_sw=True, n_threads=8)
    print(f"C1: {len(pairs)} candidate pairs", flush=True)
    for device, use_c, branch in (("0", True, "gpu"), ("cpu", True, "c"),
                                  ("cpu", False, "numba")):
        print(f"C1 branch {branch}", flush=True)
        got = compute_pairwise_alignment(pairs, None, fam.lengths, t, device=device,
                                         use_c_sw=use_c, **kw)
        wp, ws, entered, o = _oracle_pipeline(oracle_mod, fam, pairs, sk, t, band, branch)
        _same(got, (wp, ws), f"C1 {branch}")
        assert 0 < len(wp) < len(entered) < len(pairs)
    if oracle_mod.ref_available():
        s, sc, mk = oracle_mod.ref_sw_batch(entered, fam.flat, fam.offsets, fam.lengths, band,
                                            BLOSUM62, n_threads=8)
        assert np.array_equal(sc, o["scores"]) and np.array_equal(mk, o["mask"])
        assert np.array_equal(s.view(np.uint32), o["sims"].view(np.uint32))


def test_gpu_c1_pair_order_flag(c1, oracle_mod):
    """pair_order: the GPU branch defaults to the reference GPU branch's
    stable max-length order (pairwise.py:1480-1487, 1553-1555), the CPU
    branches to min-index order; either can be forced, and the set of pairs
    with their identities is the same in both orders."""
    from clustkit_amd.pairwise import compute_pairwise_alignment
    fam, sk, pairs, _ = c1
    kw = dict(band_width=100, sketches=sk, flat_sequences=fam.flat, offsets=fam.offsets,
              use_sw=True, use_c_sw=True)
    g_max = compute_pairwise_alignment(pairs, None, fam.lengths, 0.5, device="0", **kw)
    g_min = compute_pairwise_alignment(pairs, None, fam.lengths, 0.5, device="0",
                                       pair_order="min_index", **kw)
    c_max = compute_pairwise_alignment(pairs, None, fam.lengths, 0.5, device="cpu",
                                       pair_order="max_length", **kw)
    for got, branch, order in ((g_min, "gpu", "min_index"), (c_max, "c", "max_length")):
        wp, ws, _, _ = _oracle_pipeline(oracle_mod, fam, pairs, sk, 0.5, 100, branch, order)
        _same(got, (wp, ws), f"{branch} {order}")
    L = fam.lengths
    mx = np.maximum(L[g_max[0][:, 0]], L[g_max[0][:, 1]])
    assert np.all(np.diff(mx) >= 0) and not np.array_equal(g_max[0], g_min[0])
    key = lambda r: {(int(a), int(b)): int(v) for (a, b), v in  # noqa: E731
                     zip(r[0], r[1].view(np.uint32))}
    assert key(g_max) == key(g_min)
    with pytest.raises(ValueError):
        compute_pairwise_alignment(pairs, None, fam.lengths, 0.5, pair_order="sorted", **kw)


def test_gpu_c2_mixed_lengths_reduced(oracle_mod):
    """C2 form at 5,000 proteins (median 350, clip 30-2000, identity
    0.3-0.95), k=5 / s=128, 512 x 1 LSH, t=0.3, band 100: the Jaccard
    pre-filter applies, and the GPU branch's pairs and identities equal the
    oracle composition."""
    import dataclasses

    from clustkit_amd.pairwise import compute_pairwise_alignment
    from clustkit_amd.workloads import SPECS, generate_families
    fam = generate_families(dataclasses.replace(SPECS["mixed2m"], n_roots=500))
    sk, pairs, lshp = _candidates(fam, 5, 128, 0.3)
    assert lshp == {"num_tables": 512, "num_bands": 1}
    print(f"C2 reduced: {len(pairs)} candidate pairs", flush=True)
    assert oracle_mod.jaccard_prefilter(pairs, sk, 0.3) is not None
    got = compute_pairwise_alignment(pairs, None, fam.lengths, 0.3, band_width=100, device="0",
                                     sketches=sk, flat_sequences=fam.flat, offsets=fam.offsets,
                                     use_sw=True, use_c_sw=True)
    wp, ws, entered, _ = _oracle_pipeline(oracle_mod, fam, pairs, sk, 0.3, 100, "gpu")
    _same(got, (wp, ws), "C2")
    assert len(entered) < len(pairs) and len(wp) > 1000


def test_gpu_c3_heavy_tail_reduced(oracle_mod):
    """C3 form: 5,000 proteins, 10% of root lengths Pareto(1.5) in
    1000-8000, plus a planted 7,900-residue root with three descendants at
    97-99% identity and no indels (best scores above the 16-bit range at
    every band).  Every band runs
    all pairs in one launch; the wide-score pass re-runs some pairs; every
    pair with a sequence of >= 1000 residues and 3,000 others equal the
    oracle."""
    import dataclasses

    from clustkit_amd.device import last_paths
    from clustkit_amd.pairwise import batch_sw_align
    from clustkit_amd.scoring import pack_sequences
    from clustkit_amd.workloads import SPECS, _BG_CDF, _draw, generate_families, mutate
    fam = generate_families(dataclasses.replace(SPECS["heavy"], n_roots=500))
    rng = np.random.default_rng(33)
    seqs = [fam.flat[o:o + n] for o, n in zip(fam.offsets, fam.lengths)]
    root = _draw(rng, _BG_CDF, 7900)
    seqs += [root] + [mutate(rng, root, float(rng.uniform(0.97, 0.99)), 0.0)
                      for _ in range(3)]
    flat, offsets, lengths = pack_sequences(seqs)
    fam = dataclasses.replace(fam, flat=flat, offsets=offsets, lengths=lengths)
    sk, pairs, _ = _candidates(fam, 5, 128, 0.3)
    jm = oracle_mod.jaccard_prefilter(pairs, sk, 0.3)
    if jm is not None:
        pairs = pairs[jm]
    L = lengths
    long_ = np.maximum(L[pairs[:, 0]], L[pairs[:, 1]]) >= 1000
    assert long_.sum() > 200 and (np.maximum(L[pairs[:, 0]], L[pairs[:, 1]]) >= 7000).any()
    idx = np.flatnonzero(long_)
    idx = np.union1d(idx, rng.choice(len(pairs), size=min(3000, len(pairs)), replace=False))
    print(f"C3 reduced: {len(pairs)} pairs, {long_.sum()} with a long sequence", flush=True)
    for band in (16, 32, 64, 128):
        print(f"C3 band {band}", flush=True)
        sims, scores, matches, mask = batch_sw_align(pairs, flat, offsets, lengths, band)
        assert last_paths()["wide"] > 0, band
        o = oracle_mod.sw_batch(pairs[idx], flat, offsets, lengths, band, BLOSUM62)
        assert np.array_equal(scores[idx], o["scores"]), band
        assert np.array_equal(matches[idx], o["matches"]), band
        assert np.array_equal(sims[idx].view(np.uint32), o["sims"].view(np.uint32)), band
        assert np.array_equal(mask[idx], o["mask"]), band
        assert scores.max() >= 32700


def test_gpu_c5_serving_hits_equal_oracle_stages(oracle_mod):
    """C5 form at a 20,000-protein database (median 350, clip 30-5000), k=3
    index built on the device, 96 queries of 50-2000 residues (half planted
    descendants) on Poisson arrivals through the continuous-batching
    executor: every request's hits equal the oracle stages -- k-mer Phase A/B
    (kmer_score.c:198-472), SW of the top 64 at the executor's band, rank
    0.267 S - ln(qlen tlen) in float32, stable top 10 (kmer_index.py:2303-2331,
    search.py:727-800)."""
    import dataclasses

    from clustkit_amd.kmer_search import DeviceKmerIndex, SequenceSet, evalue_rank
    from clustkit_amd.serving import GpuExecutor, Request, Scheduler, ServeParams, serve
    from clustkit_amd.workloads import SPECS, generate_families_streams, make_queries
    db = generate_families_streams(dataclasses.replace(SPECS["serve1m"], n_roots=2000))
    ds = SequenceSet.from_arrays(db.flat, db.offsets, db.lengths)
    idx = DeviceKmerIndex.build(ds, k=3)
    params = ServeParams(top_k=10, max_cands=64)
    ex = GpuExecutor(idx, params, n_streams=4, max_batch=32, max_qlen=2000)
    qs, _ = make_queries(db, 96, seed=5)
    rng = np.random.default_rng(6)
    arr = np.cumsum(rng.exponential(1 / 20000.0, len(qs)))
    sch = Scheduler(ex, hbm_budget=8 << 30, max_batch=32)
    done = serve(sch, [Request(rid=i, seq=q, arrival=float(a))
                       for i, (q, a) in enumerate(zip(qs, arr))])
    assert len(done) == len(qs) and sch.batches < len(qs)
    nd = len(db.lengths)
    ke = idx.host_entries()
    ql = np.array([len(q) for q in qs], np.int32)
    qo = np.zeros(len(qs), np.int64)
    qo[1:] = np.cumsum(ql[:-1].astype(np.int64))
    qf = np.concatenate(qs)
    tg, cn, _, _ = oracle_mod.kmer_score_batch(qf, qo, ql, 3, 20, idx.db.kmer_offsets, ke,
                                               idx.db.kmer_freqs, ex.freq_thresh, nd,
                                               max_cands=64)
    flat = np.concatenate([db.flat, qf])
    offs = np.concatenate([db.offsets, qo + db.flat.size])
    lens = np.concatenate([db.lengths, ql]).astype(np.int32)
    n_hits = 0
    for r in done:
        q = r.rid
        t = tg[q, :cn[q]]
        pairs = np.stack([np.full(len(t), nd + q), t], 1).astype(np.int32)
        o = oracle_mod.sw_batch(pairs, flat, offs, lens, ex.band, BLOSUM62)
        keep = o["scores"] > 0
        rank = evalue_rank(o["scores"][keep].astype(np.float32),
                           np.full(keep.sum(), ql[q]), db.lengths[t[keep]])
        order = np.lexsort((-rank, np.zeros(len(rank))))[:10]
        want = [(int(t[keep][i]), float(o["sims"][keep][i]), float(rank[i])) for i in order]
        assert r.hits == want, q
        n_hits += len(want)
    assert n_hits > 100
