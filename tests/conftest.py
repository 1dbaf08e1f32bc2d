"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and call the HIP
library through its C-ABI; everything else runs on the CPU."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP library + device)")


def golden_files():
    return sorted(GOLDEN.glob("sw_*.npz"))


def load_golden(path):
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["hints"] = d["hints"] if bool(d["has_hints"]) else None
    d["name"] = Path(path).stem
    return d


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def small_family():
    from clustkit_amd.workloads import FamilySpec, generate_families
    return generate_families(FamilySpec(n_roots=30, n_desc=9, median=260, clip_lo=20,
                                        clip_hi=900, id_lo=0.3, id_hi=0.95, seed=11))
