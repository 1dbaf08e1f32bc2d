"""Residue alphabets, integer encoding and the BLOSUM62 scoring table.

Encoding follows clustkit/io.py:9-18,134-156: protein residues in the order
``ACDEFGHIKLMNPQRSTVWY`` map to 0..19 and anything else to 20; nucleotides
``ACGT`` map to 0..3 and anything else to 4.

``BLOSUM62`` is the reference's table (clustkit/pairwise.py:129-151): the
Henikoff & Henikoff (1992) matrix in the residue order above, with the
reference's one deviation kept for bit-exact parity -- row N, column D is 0
(the published value, and the reference's row D column N, is 1).  The
kernels index it as sub[row residue * 20 + column residue], the row
sequence being the shorter one, as clustkit/csrc/sw_align.c:117 does.  Gap scores are open
-11, extend -1 (a gap of length L costs 10 + L), as in
clustkit/csrc/sw_align.c:45-46.  A residue code >= 20 scores -4 against
anything (sw_align.c:116-119) and two equal codes count as a match.
"""

from __future__ import annotations

import numpy as np

PROTEIN_ALPHABET = "ACDEFGHIKLMNPQRSTVWY"
NUCLEOTIDE_ALPHABET = "ACGT"
PROTEIN_UNKNOWN = len(PROTEIN_ALPHABET)
NUCLEOTIDE_UNKNOWN = len(NUCLEOTIDE_ALPHABET)

GAP_OPEN = -11
GAP_EXTEND = -1
UNKNOWN_SCORE = -4

_ROWS = """
A  4  0 -2 -1 -2  0 -2 -1 -1 -1 -1 -2 -1 -1 -1  1  0  0 -3 -2
C  0  9 -3 -4 -2 -3 -3 -1 -3 -1 -1 -3 -3 -3 -3 -1 -1 -1 -2 -2
D -2 -3  6  2 -3 -1 -1 -3 -1 -4 -3  1 -1  0 -2  0 -1 -3 -4 -3
E -1 -4  2  5 -3 -2  0 -3  1 -3 -2  0 -1  2  0  0 -1 -2 -3 -2
F -2 -2 -3 -3  6 -3 -1  0 -3  0  0 -3 -4 -3 -3 -2 -2 -1  1  3
G  0 -3 -1 -2 -3  6 -2 -4 -2 -4 -3  0 -2 -2 -2  0 -2 -3 -2 -3
H -2 -3 -1  0 -1 -2  8 -3 -1 -3 -2  1 -2  0  0 -1 -2 -3 -2  2
I -1 -1 -3 -3  0 -4 -3  4 -3  2  1 -3 -3 -3 -3 -2 -1  3 -3 -1
K -1 -3 -1  1 -3 -2 -1 -3  5 -2 -1  0 -1  1  2  0 -1 -2 -3 -2
L -1 -1 -4 -3  0 -4 -3  2 -2  4  2 -3 -3 -2 -2 -2 -1  1 -2 -1
M -1 -1 -3 -2  0 -3 -2  1 -1  2  5 -2 -2  0 -1 -1 -1  1 -1 -1
N -2 -3  0  0 -3  0  1 -3  0 -3 -2  6 -2  0  0  1  0 -3 -4 -2
P -1 -3 -1 -1 -4 -2 -2 -3 -1 -3 -2 -2  7 -1 -2 -1 -1 -2 -4 -3
Q -1 -3  0  2 -3 -2  0 -3  1 -2  0  0 -1  5  1  0 -1 -2 -2 -1
R -1 -3 -2  0 -3 -2  0 -3  2 -2 -1  0 -2  1  5 -1 -1 -3 -3 -2
S  1 -1  0  0 -2  0 -1 -2  0 -2 -1  1 -1  0 -1  4  1 -2 -3 -2
T  0 -1 -1 -1 -2 -2 -2 -1 -1 -1 -1  0 -1 -1 -1  1  5  0 -2 -2
V  0 -1 -3 -2 -1 -3 -3  3 -2  1  1 -3 -2 -2 -3 -2  0  4 -3 -1
W -3 -2 -4 -3  1 -2 -2 -3 -3 -2 -1 -4 -4 -2 -3 -3 -2 -3 11  2
Y -2 -2 -3 -2  3 -3  2 -1 -2 -1 -1 -2 -3 -1 -2 -2 -2 -1  2  7
"""


def _parse_matrix(text: str) -> np.ndarray:
    rows = [ln.split() for ln in text.strip().splitlines()]
    assert "".join(r[0] for r in rows) == PROTEIN_ALPHABET
    return np.array([[int(v) for v in r[1:]] for r in rows], dtype=np.int8)


BLOSUM62 = _parse_matrix(_ROWS)
BLOSUM62.setflags(write=False)

# Background amino-acid frequencies of the BLOSUM62 target distribution
# (Henikoff & Henikoff 1992 marginals), in PROTEIN_ALPHABET order.  Used by
# the synthetic workload generator only.
BLOSUM62_BACKGROUND = np.array([
    0.0740, 0.0250, 0.0535, 0.0543, 0.0468, 0.0741, 0.0262, 0.0679, 0.0581, 0.0990,
    0.0248, 0.0448, 0.0387, 0.0342, 0.0516, 0.0572, 0.0508, 0.0728, 0.0130, 0.0323,
], dtype=np.float64)


def encode_sequence(sequence: str, mode: str = "protein") -> np.ndarray:
    """Integer-encode one sequence (clustkit/io.py:134-156)."""
    alphabet = PROTEIN_ALPHABET if mode == "protein" else NUCLEOTIDE_ALPHABET
    unknown = len(alphabet)
    lut = np.full(256, unknown, dtype=np.uint8)
    for code, ch in enumerate(alphabet):
        lut[ord(ch)] = code
        lut[ord(ch.lower())] = code
    raw = np.frombuffer(sequence.encode("latin-1"), dtype=np.uint8)
    return lut[raw]


def pack_sequences(seqs) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Concatenate encoded sequences into the flat layout used by the
    alignment stage: (flat uint8[sum len], offsets int64[N], lengths int32[N])
    -- the compact form of clustkit/io.py:35-131."""
    lengths = np.fromiter((len(s) for s in seqs), dtype=np.int32, count=len(seqs))
    offsets = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs):
        np.cumsum(lengths[:-1], out=offsets[1:])
    flat = np.concatenate([np.asarray(s, dtype=np.uint8) for s in seqs]) if len(seqs) else \
        np.zeros(0, np.uint8)
    return flat, offsets, lengths
