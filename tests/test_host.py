"""Host-side checks that need no GPU: the C-ABI library loads and exports
every symbol the header declares, host data preparation, and error paths."""

from __future__ import annotations

import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared_symbols():
    text = (ROOT / "include" / "clustkit_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(ck_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    from clustkit_amd import _lib
    declared = _declared_symbols()
    assert declared, "no declarations parsed"
    assert sorted(_lib.SIGNATURES) == declared


def test_library_exports_every_declared_symbol():
    from clustkit_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built (run __graft_entry__.build())")
    loaded = _lib.load()
    for name in _declared_symbols():
        assert hasattr(loaded, name), name
    assert loaded.ck_version().startswith(b"clustkit_hip")
    assert loaded.ck_device_count() >= 0


def test_library_shares_torch_hip_runtime():
    """Loading the library before torch must still leave one HIP runtime in
    the process (two runtimes: the second sees no GPU, pointers do not mix)."""
    import subprocess
    import sys
    from clustkit_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from clustkit_amd import _lib; _lib.load(); import torch\n"
            "libs = {l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}\n"
            "print(len(libs))" % str(ROOT))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "1"


def test_entry_points_fail_loudly_without_gpu():
    from clustkit_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    if lib.ck_device_count() > 0:
        pytest.skip("a GPU is visible")
    from clustkit_amd.pairwise import _batch_sw_compact_scored
    from clustkit_amd.scoring import BLOSUM62
    with pytest.raises(_lib.HipLibraryError):
        _batch_sw_compact_scored(np.array([[0, 1]], np.int32), np.zeros(4, np.uint8),
                                 np.array([0, 2]), np.array([2, 2]), 0.5, 10, BLOSUM62)
    # the raw C entry point reports the missing device with an error code
    rc = lib.ck_init(0)
    assert rc == -4 and b"GPU" in lib.ck_last_error()


def test_kmer_score_device_rejects_half_given_database():
    """ck_kmer_score_device takes the database residues as a pair (flat
    bytes and offsets), and the window codes as a pair (codes and code
    offsets, 16-byte aligned), or not at all; a half-given pair or a
    misaligned code array is refused before any device work (dummy non-null
    pointers are never read)."""
    import ctypes
    from clustkit_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    p = ctypes.c_void_p(256)
    for flat, offs, flen in ((p, None, 16), (None, p, 0), (p, p, -1)):
        rc = lib.ck_kmer_score_device(p, 16, p, p, 1, 8, 3, 20, p, p, p, p, None, p, flat, flen,
                                      offs, None, None, 10, 8, 100, 2, 2, 10, 64, 100, p, p, p,
                                      p, None)
        assert rc == -1
        assert b"d_db_flat and d_db_offsets go together" in lib.ck_last_error()
    for codes, coff, msg in ((p, None, b"go together"), (None, p, b"go together"),
                             (ctypes.c_void_p(258), p, b"16-byte aligned")):
        rc = lib.ck_kmer_score_device(p, 16, p, p, 1, 8, 3, 20, p, p, p, p, None, p, None, 0,
                                      None, codes, coff, 10, 8, 100, 2, 2, 10, 64, 100, p, p, p,
                                      p, None)
        assert rc == -1
        assert msg in lib.ck_last_error()


def test_kmer_db_codes_device_rejects_bad_arguments():
    """ck_kmer_db_codes_device refuses code spaces above 8192 codes, k > 5,
    null pointers and a misaligned code array before any device work."""
    import ctypes
    from clustkit_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    lib = _lib.load()
    p = ctypes.c_void_p(256)
    cases = (((p, p, p, 4, 4, 20, p, p), -5),       # 20^4 codes
             ((p, p, p, 4, 6, 4, p, p), -5),        # k = 6
             ((p, p, None, 4, 3, 20, p, p), -1),
             ((p, p, p, 4, 3, 20, p, ctypes.c_void_p(264)), -1),
             ((p, p, p, -1, 3, 20, p, p), -1))
    for args, want in cases:
        assert lib.ck_kmer_db_codes_device(*args, None) == want, args


def test_scoring_table_and_encoding():
    from clustkit_amd.scoring import (BLOSUM62, PROTEIN_ALPHABET, encode_sequence,
                                      pack_sequences)
    # The reference table (pairwise.py:129-151) is symmetric except N/D:
    # B[D][N] = 1 but B[N][D] = 0.  Kept as is: the kernels index
    # sub[row residue * 20 + column residue] exactly like sw_align.c:117.
    asym = {(int(i), int(j)) for i, j in np.argwhere(BLOSUM62 != BLOSUM62.T)}
    assert asym == {(2, 11), (11, 2)}
    assert BLOSUM62[2, 11] == 1 and BLOSUM62[11, 2] == 0
    assert BLOSUM62[18, 18] == 11 and BLOSUM62.min() == -4
    enc = encode_sequence("ACDXwy*", "protein")
    assert list(enc) == [0, 1, 2, 20, 18, 19, 20]
    assert list(encode_sequence("ACGTN", "nucleotide")) == [0, 1, 2, 3, 4]
    flat, off, ln = pack_sequences([enc, enc[:2], np.zeros(0, np.uint8)])
    assert list(off) == [0, 7, 9] and list(ln) == [7, 2, 0] and len(flat) == 9
    assert PROTEIN_ALPHABET == "ACDEFGHIKLMNPQRSTVWY"


def test_auto_lsh_params():
    from clustkit_amd.lsh import auto_lsh_params
    assert auto_lsh_params(0.3, "medium", 5) == {"num_tables": 512, "num_bands": 1}
    assert auto_lsh_params(0.9, "medium", 5)["num_bands"] == 3


def test_top_k_selection_matches_per_query_stable_sort():
    """The vectorised top-k of search.py:727-800 (stable sort by query, then
    rank descending; first top_k per query) against a per-query loop."""
    from clustkit_amd.kmer_search import _top_k_order
    rng = np.random.default_rng(4)
    qi = rng.integers(0, 50, 3000)
    rank = rng.integers(0, 20, 3000).astype(np.float32)        # many ties
    for top_k in (1, 3, 10, 100):
        got = _top_k_order(qi, rank, top_k)
        want = []
        for q in range(50):
            idx = np.flatnonzero(qi == q)
            order = sorted(idx, key=lambda i: -rank[i])          # stable
            want += order[:top_k]
        assert list(got) == want
    assert len(_top_k_order(np.zeros(0, int), np.zeros(0, np.float32), 5)) == 0


def test_nw_pair_cells_brute_force():
    """Cells of the banded global alignment (batch rules of
    clustkit/pairwise.py:684-850) against counting the band cells row by row."""
    from clustkit_amd.cells import nw_pair_cells
    rng = np.random.default_rng(1)
    lens = rng.integers(0, 120, 40).astype(np.int32)
    pairs = rng.integers(0, 40, (200, 2))
    for bwp in (0, 5, 20, 100):
        for t in (0.0, 0.5):
            c = nw_pair_cells(pairs, lens, bwp, t)
            for p, (x, y) in enumerate(pairs):
                la, lb = sorted((int(lens[x]), int(lens[y])))
                if (lb > 0 and np.float32(la) / np.float32(lb) < np.float32(t)) or lb - la > bwp:
                    want = 0
                else:
                    ab = min(bwp, max(10, lb - la + 10))
                    bw = lb if (ab <= 0 or lb <= 50) else ab
                    want = sum(min(lb, i + bw) - max(1, i - bw) + 1 for i in range(1, la + 1))
                assert c[p] == want, (p, bwp, t)


def test_database_save_load_roundtrip(tmp_path):
    """The reference's on-disk layout (clustkit/database.py:223-300 / 304-442):
    every array and the ids/params round-trip; a database written with the
    padded encoded_sequences.npy instead of the compact arrays loads to the
    same compact table."""
    import json
    from clustkit_amd.kmer_search import (SequenceSet, build_database_index, load_database,
                                          save_database)
    rng = np.random.default_rng(2)
    lens = rng.integers(1, 120, 40).astype(np.int32)
    off = np.zeros(40, np.int64)
    off[1:] = np.cumsum(lens)[:-1]
    flat = rng.integers(0, 21, int(lens.sum())).astype(np.uint8)
    ds = SequenceSet.from_arrays(flat, off, lens, ids=[f"p{i}" for i in range(40)])
    db = build_database_index(ds)
    db.sketches = rng.integers(0, 2**63, (40, 16), dtype=np.uint64)
    db.freq_thresholds = {"std": 100}
    save_database(db, tmp_path / "db")
    got = load_database(tmp_path / "db")
    assert got.dataset.ids == ds.ids and got.params["num_sequences"] == 40
    for a, b in ((got.dataset.flat_sequences, flat), (got.dataset.offsets, off),
                 (got.dataset.lengths, lens), (got.kmer_offsets, db.kmer_offsets),
                 (got.kmer_entries, db.kmer_entries), (got.kmer_freqs, db.kmer_freqs),
                 (got.kmer_entries_compact, db.kmer_entries_compact),
                 (got.sketches, db.sketches)):
        assert np.array_equal(a, b)
    assert got.freq_thresholds == {"std": 100} and got.lsh_seeds is None
    # padded form
    pad = tmp_path / "pad"
    pad.mkdir()
    enc = np.full((40, int(lens.max())), 20, np.uint8)
    for i in range(40):
        enc[i, :lens[i]] = flat[off[i]:off[i] + lens[i]]
    np.save(pad / "encoded_sequences.npy", enc)
    np.save(pad / "lengths.npy", lens)
    (pad / "ids.json").write_text(json.dumps(ds.ids))
    (pad / "params.json").write_text(json.dumps({"mode": "protein", "kmer_size": 3}))
    got = load_database(pad)
    assert np.array_equal(got.dataset.flat_sequences, flat)
    assert np.array_equal(got.dataset.offsets, off) and got.kmer_offsets is None
