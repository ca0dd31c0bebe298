"""The benchmark entry point's contract on the CPU: one GPU only, and no number without a GPU
(bench.py never falls back to a CPU path)."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

from hbdia_amd import gpu_available

ROOT = Path(__file__).resolve().parents[1]


def run_bench(*args: str) -> subprocess.CompletedProcess:
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], cwd=ROOT, capture_output=True, text=True,
                          timeout=300)


def test_more_than_one_gpu_is_refused():
    p = run_bench("--gpus", "2")
    assert p.returncode == 2
    assert p.stdout == ""
    assert "one GPU" in p.stderr


@pytest.mark.skipif(gpu_available(), reason="checks the behaviour without a HIP device")
def test_no_device_no_number():
    p = run_bench("--steps", "1", "--warmup", "0", "--no-forms", "--no-cpu")
    assert p.returncode != 0
    assert p.stdout == ""  # no JSON line: nothing is measured on the CPU
    assert "no HIP device" in p.stderr
