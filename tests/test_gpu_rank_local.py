"""GPU parity of the rank-local form: one rank of a row-wise partition on one
GPU, its halo columns in the COO part and x's halo segments filled from the
global vector on the host.

Expected agreement: bit-exact against the oracle's rank-local restatement
(same blocked sums over the local segment, same COO order, y = bdia + coo).
Against the global SpMV the rows agree within 1e-12 relative in FP64 (the
halo entries move from the blocked sum to the COO sum, which changes the
summation order) and exactly when every sum is exact (small integers).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hbdia_amd import HBDIA, ExecutionMode, HBDIAError, HBDIAVector, hbdiaSpMV

pytestmark = pytest.mark.gpu


def x_ext_of(loc: dict, x: np.ndarray) -> np.ndarray:
    L, n, r0 = loc["halo_left"], loc["num_rows"], loc["first_row"]
    h = loc["halo_cols"]
    return np.concatenate([x[h[:L]], x[r0:r0 + n], x[h[L:]]]).astype(x.dtype)


def run_rank(G: HBDIA, x: np.ndarray, rank: int, size: int, bw: int, thr: int, cpu: bool, max_coo: int = 10 ** 9):
    mode = ExecutionMode.CPU_COO if cpu else ExecutionMode.GPU_COO
    A = HBDIA.createRankLocalMatrix(G, rank, size, bw, thr, max_coo, mode)
    A.prepareForGPU()
    lay = A.getLocalLayout()
    r0 = lay["first_row"]
    n = A.getNumRows()
    X = HBDIAVector.withHalo(x[r0:r0 + n], A)
    X.fillHalo(A, x)
    Y = HBDIAVector(np.zeros(n, x.dtype))
    assert hbdiaSpMV(A, X, Y, cpu, not cpu)
    A.synchronize()
    return A, Y.download()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("size,bw,thr", [(4, 32, 2), (3, 64, 2), (5, 16, 1), (2, 128, 3)])
@pytest.mark.parametrize("cpu", [False, True])
def test_rank_local_bit_exact(gpu, dtype, size, bw, thr, cpu):
    print(f"case: rank-local 24^3 noise 0.01 size {size} bw {bw} thr {thr} {np.dtype(dtype).name} "
          f"{'cpu' if cpu else 'gpu'} COO", flush=True)
    r, c, v, N = oracle.stencil27(24, 24, 24, 0.01, 2, dtype=dtype)
    x = np.random.default_rng(7).uniform(-1, 1, N).astype(dtype)
    G = HBDIA(r, c, v, N, N)
    hb = oracle.convert(r, c, v, N, N, bw, thr)
    y_global = oracle.hbdia_spmv(hb, x.copy(), np.zeros(N, dtype), N)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for rank in range(size):
        A, y = run_rank(G, x, rank, size, bw, thr, cpu)
        loc = oracle.rank_local(r, c, v, N, N, rank, size, bw, thr)
        want = oracle.rank_local_spmv(loc, x_ext_of(loc, x), np.zeros(loc["num_rows"], dtype))
        np.testing.assert_array_equal(y, want, err_msg=f"rank {rank}")
        r0 = loc["first_row"]
        g = y_global[r0:r0 + loc["num_rows"]]
        scale = np.maximum(np.abs(g), 1.0)
        assert np.max(np.abs(y - g) / scale) <= tol
        assert loc["folded"] > 0 or size == 1


def test_rank_local_exact_sums_equal_global_rows(gpu):
    print("case: rank-local 16^3 integer values, size 4, bw 64", flush=True)
    r, c, _, N = oracle.stencil27(16, 16, 16, 0.02, 4)
    rng = np.random.default_rng(4)
    v = (rng.integers(1, 5, len(r)) * rng.choice([-1.0, 1.0], len(r))).astype(np.float64)
    x = rng.integers(-8, 9, N).astype(np.float64)
    G = HBDIA(r, c, v, N, N)
    y_global = oracle.coo_spmv_longdouble(r, c, v, x, N)
    for rank in range(4):
        A, y = run_rank(G, x, rank, 4, 64, 2, False)
        r0 = rank * (N // 4)
        np.testing.assert_array_equal(y, y_global[r0:r0 + A.getNumRows()])


def test_rank_local_scatter_workload_128(gpu):
    print("case: rank-local 128^3 stencil + 0.5% scatter, rank 3 of 8, bw 64 thr 2", flush=True)
    G = HBDIA.stencil(128, 27, 0.005, seed=0xC2)
    r, c, v = G.getCOO()
    N = G.getNumRows()
    x = np.random.default_rng(11).uniform(-1, 1, N)
    A, y = run_rank(G, x, 3, 8, 64, 2, False)
    loc = oracle.rank_local(r, c, v, N, N, 3, 8, 64, 2)
    want = oracle.rank_local_spmv(loc, x_ext_of(loc, x), np.zeros(loc["num_rows"]))
    np.testing.assert_array_equal(y, want)
    lay = A.getLocalLayout()
    assert lay["folded"] == loc["folded"] > 0
    assert lay["halo_left"] > 0 and lay["halo_right"] > 0


def test_rank_local_rejects_vector_without_halo(gpu):
    print("case: rank-local halo size mismatch", flush=True)
    r, c, v, N = oracle.stencil27(12, 12, 12, 0.0, 0)
    G = HBDIA(r, c, v, N, N)
    A = HBDIA.createRankLocalMatrix(G, 1, 3, 32, 2)
    A.prepareForGPU()
    n = A.getNumRows()
    X = HBDIAVector(np.ones(n))
    Y = HBDIAVector(np.zeros(n))
    with pytest.raises((ValueError, HBDIAError)):
        hbdiaSpMV(A, X, Y, False, True)


@pytest.mark.parametrize("size,rank", [(4, 0), (4, 2), (4, 3), (3, 1)])
def test_halo_exchange_verifies(gpu, size, rank):
    """The reference's self-check of its exchange (x[g] = g, every received entry its global index)
    on the vector filled from the global x."""
    print(f"case: halo verify 16^3 noise 0.02 size {size} rank {rank}", flush=True)
    r, c, v, N = oracle.stencil27(16, 16, 16, 0.02, 1)
    G = HBDIA(r, c, v, N, N)
    A = HBDIA.createRankLocalMatrix(G, rank, size, 32, 2)
    A.prepareForGPU()
    lay = A.getLocalLayout()
    x = np.arange(N, dtype=np.float64)
    r0, n = lay["first_row"], A.getNumRows()
    X = HBDIAVector.withHalo(x[r0:r0 + n], A)
    assert not X.verifyHalo(A) or lay["halo_left"] + lay["halo_right"] == 0  # halo still zero
    X.fillHalo(A, x)
    assert X.verifyHalo(A)
