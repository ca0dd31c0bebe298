"""Writes the golden fixtures of tests/golden/ from the oracle.

    python tests/golden/make_golden.py

The reference itself cannot be built or run here (it needs nvcc, the CUDA
runtime and cuSPARSE), and it ships no fixtures of its own, so the fixtures
are the oracle's outputs on small seeded inputs.  They are checked against
closed forms and brute force in tests/test_oracle.py and carried to the GPU
box, where /root/reference does not exist.

Fixtures (numpy .npz, no pickles):
  config1_y.npz        8^3 stencil, noise 0, bw 32, thr 2, x[i] = i + 1, COO on the CPU:
                       y (512)   -- the reference's simple_hbdia_example input
  noisy16_y.npz        16^3 stencil, noise 0.1, iteration 0, bw 32, thr 2, x[i] = i + 1: y, coo count
  noisy24_y.npz        24^3 stencil, noise 0.02, iteration 1, bw 16, thr 3, x[i] = sin(i): y, coo count
  stencil6_noise.npz   full COO of the 6^3 stencil with noise 0.2, iteration 7 (detects a
                       change of the mt19937 / distribution / sort behaviour)
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

import oracle  # noqa: E402  (test infrastructure)


def spmv_case(n: int, noise: float, iteration: int, bw: int, thr: int, x_kind: str):
    r, c, v, N = oracle.stencil27(n, n, n, noise, iteration)
    hb = oracle.convert(r, c, v, N, N, bw, thr)
    x = np.arange(1, N + 1, dtype=np.float64) if x_kind == "i+1" else np.sin(np.arange(N, dtype=np.float64))
    y = np.zeros(N)
    oracle.hbdia_spmv(hb, x, y, N)
    return y, len(hb["coo_values"])


def main() -> None:
    y, ncoo = spmv_case(8, 0.0, 0, 32, 2, "i+1")
    assert ncoo == 0
    np.savez(HERE / "config1_y.npz", y=y)
    y, ncoo = spmv_case(16, 0.1, 0, 32, 2, "i+1")
    np.savez(HERE / "noisy16_y.npz", y=y, ncoo=np.int64(ncoo))
    y, ncoo = spmv_case(24, 0.02, 1, 16, 3, "sin")
    np.savez(HERE / "noisy24_y.npz", y=y, ncoo=np.int64(ncoo))
    r, c, v, N = oracle.stencil27(6, 6, 6, 0.2, 7)
    np.savez(HERE / "stencil6_noise.npz", rows=r, cols=c, values=v, n=np.int64(N))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
