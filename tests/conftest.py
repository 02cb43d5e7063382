"""Test configuration: the package and oracle on sys.path, the gpu marker, shared fixtures."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "fast-er-lmm_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libfasterlmm_amd.so")


@pytest.fixture(scope="session")
def example():
    """The reference's example panel as arrays: 150 strains x 1,500 SNPs on 16 chromosomes, 20 phenotypes."""
    g = np.load(GOLDEN / "example_inputs.npz")
    return {k: g[k] for k in g.files}


def golden(name: str) -> dict:
    g = np.load(GOLDEN / name)
    return {k: g[k] for k in g.files}
