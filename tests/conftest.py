"""Shared test set-up: import paths, the `gpu` marker, parity helpers."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "astraea-impact-sim_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


def rel_rms(a, ref):
    """sqrt(sum |a - ref|^2 / sum |ref|^2) over all particles (vectors: rows)."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.sqrt(np.sum(ref * ref))
    return float(np.sqrt(np.sum((a - ref) ** 2)) / den) if den > 0 else float(np.sqrt(np.sum(a * a)))


@pytest.fixture(scope="session")
def lib():
    import astraea_amd

    return astraea_amd


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle
