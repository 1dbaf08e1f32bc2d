"""CPU checks of the autotuner's sample sizing (clustkit_amd/autotune.py):
pair costs follow the kernel's band rule, and a sample grows until no pair
in it outlasts its share of the GPU's pair slots."""

import numpy as np

from clustkit_amd.autotune import pair_costs, sample_size


def test_pair_costs_follow_the_band_rule():
    pl = np.array([[100, 120], [33, 40], [17, 300], [300, 17]])
    got = pair_costs(pl, 16)
    # stripes of 16 rows of the shorter side x (2 bw + 1) positions; a longer
    # side of at most 50 takes the whole longer sequence as the band
    assert got.tolist() == [7 * 33, 3 * (2 * 40 + 1), 2 * 33, 2 * 33]
    assert pair_costs(pl[:1], 0).tolist() == [7 * (2 * 120 + 1)]


def test_sample_grows_past_long_pairs_and_stays_otherwise():
    rng = np.random.default_rng(0)
    n, slots = 50_000, 256 * 64
    short = rng.integers(200, 400, n)
    pl = np.stack([short, short + 10], 1)
    assert sample_size(pl, 64, slots, 10 * n) == n          # uniform: no pair dominates
    pl[:20] = 8000                                           # a few 8,000-residue pairs
    grown = sample_size(pl, 64, slots, 100 * n)
    assert n < grown < 100 * n
    cost = pair_costs(pl, 64)
    assert cost.sum() * grown / n / slots >= 2 * cost.max() * 0.999
    assert sample_size(pl, 64, slots, 10 * n) == 10 * n      # capped at the pairs there are
    assert sample_size(pl[:0], 64, slots, 100) == 0
