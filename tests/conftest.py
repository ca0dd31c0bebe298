"""Test configuration: repository root on sys.path and the `gpu` marker.

GPU tests (marked `gpu`) need a visible HIP device; they fail loudly without
one rather than skipping, because the product has no CPU fallback.
"""
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
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP kernels")


@pytest.fixture(scope="session")
def golden():
    def load(name: str):
        with np.load(GOLDEN / name, allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    return load


@pytest.fixture(scope="session")
def gpu():
    import hbdia_amd
    if not hbdia_amd.gpu_available():
        raise RuntimeError("no HIP device visible: GPU tests need an AMD GPU")
    return hbdia_amd
