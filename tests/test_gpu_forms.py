"""GPU tests of the batched and reduced-precision forms of the SpMV.

* Multiple right-hand sides: every column must equal the single-vector
  hbdiaSpMV result bit for bit (same FMA order), and the oracle.
* Reduced-precision storage (FP32 / BF16 values, FP32 accumulation): rows
  recomputed in FP64 must equal the FP64 path bit for bit; every other row must
  satisfy the forward error bound of an FP32 dot product with rounded inputs,
      |y_lp - y| <= (u_store + u_x + gamma_n) * sum |a| |x|,
  u_store = 2^-24 (FP32) or 2^-8 + 2^-24 (BF16, FP64 -> FP32 -> BF16 rounding),
  u_x = 2^-24 (x rounded to FP32), gamma_n = n u / (1 - n u) with u = 2^-24 and
  n = terms in the row (each FMA rounds once).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hbdia_amd import (HBDIA, DeviceArray, HBDIAMixed, HBDIAVector, gen_vector, hbdiaSpMV, hbdiaSpMVMulti)
from hbdia_amd.device import synchronize

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


def oracle_y(A: HBDIA, x: np.ndarray, y0: np.ndarray) -> np.ndarray:
    hb = A.getHBDIA()
    hb["num_blocks"] = A.getNumBlocks()
    hb["block_width"] = A.getBlockWidth()
    return oracle.hbdia_spmv(hb, x.copy(), y0.copy(), A.getNumRows())


def single(A: HBDIA, x: np.ndarray, y0: np.ndarray) -> np.ndarray:
    X, Y = HBDIAVector(x), HBDIAVector(y0)
    hbdiaSpMV(A, X, Y, False, True)
    A.synchronize()
    return Y.download()


@pytest.mark.parametrize("k,scatter", [(1, 0.01), (3, 0.01), (4, 0.01), (8, 0.01), (16, 0.01), (20, 0.01), (32, 0.01),
                                       (9, 0.0), (16, 0.0), (32, 0.0)])
def test_multi_rhs_columns_equal_single_vector(gpu, k, scatter):
    print(f"case: multi-RHS k={k} scatter={scatter}", flush=True)
    n = 20
    A = HBDIA.stencil(n, 27, scatter, seed=21)
    A.convertToHBDIAFormat(64, 2)
    N = n ** 3
    ld = N + 37  # a leading dimension larger than N
    Xh = np.zeros((k, ld))
    for j in range(k):
        Xh[j, :N] = gen_vector(N, 100 + j)
    Y0 = np.full((k, ld), 0.25)
    X, Y = DeviceArray.from_host(Xh.ravel()), DeviceArray.from_host(Y0.ravel())
    hbdiaSpMVMulti(A, X, Y, k, ld, ld)
    A.synchronize()
    Yh = Y.download().reshape(k, ld)
    for j in range(k):
        want = single(A, Xh[j, :N], Y0[j, :N])
        np.testing.assert_array_equal(Yh[j, :N], want, err_msg=f"column {j}")
        np.testing.assert_array_equal(Yh[j, :N], oracle_y(A, Xh[j, :N], Y0[j, :N]))
        assert np.all(Yh[j, N:] == 0.25)  # padding rows untouched


@pytest.mark.parametrize("k", [4, 12])
def test_multi_rhs_without_fallback_keeps_empty_block_rows(gpu, k):
    # rows past the last block keep Y (blocked part only, as hbdia_bdia_spmv)
    rows = np.array([0, 0, 1, 63, 64, 64], np.int32)
    cols = np.array([0, 1, 1, 63, 64, 63], np.int32)
    vals = np.array([2.0, 1.0, 3.0, 4.0, 5.0, 6.0])
    A = HBDIA(rows, cols, vals, 200, 100)  # 200 rows, blocks cover only min(200,100) rows
    A.convertToHBDIAFormat(64, 1)
    assert A.info()["num_coo"] == 0
    N = 200
    Xh = np.vstack([np.arange(N, dtype=np.float64) + j for j in range(k)])
    Y0 = np.full((k, N), -3.0)
    X, Y = DeviceArray.from_host(Xh.ravel()), DeviceArray.from_host(Y0.ravel())
    hbdiaSpMVMulti(A, X, Y, k)
    A.synchronize()
    Yh = Y.download().reshape(k, N)
    for j in range(k):
        np.testing.assert_array_equal(Yh[j], oracle_y(A, Xh[j], Y0[j]))


def test_multi_rhs_float32(gpu):
    A = HBDIA.stencil(16, 27, 0.01, seed=4, dtype=np.float32)
    A.convertToHBDIAFormat(64, 2)
    N, k = 16 ** 3, 5
    Xh = np.vstack([gen_vector(N, 7 + j, dtype=np.float32) for j in range(k)])
    X, Y = DeviceArray.from_host(Xh.ravel()), DeviceArray(k * N, np.float32, zero=True)
    hbdiaSpMVMulti(A, X, Y, k)
    A.synchronize()
    Yh = Y.download().reshape(k, N)
    for j in range(k):
        np.testing.assert_array_equal(Yh[j], oracle_y(A, Xh[j], np.zeros(N, np.float32)))


# ------------------------------------------------------------ reduced precision
def _row_stats(A: HBDIA, x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(sum |a| |x|, terms) per row over the matrix's entries."""
    r, c, v = A.getCOO()
    N = A.getNumRows()
    absax = np.bincount(r, weights=np.abs(v) * np.abs(x[c]), minlength=N)
    terms = np.bincount(r, minlength=N).astype(np.float64)
    return absax, terms


def _bound(absax: np.ndarray, terms: np.ndarray, precision: str) -> np.ndarray:
    """Forward error bound of the FP32 row sums against the FP64 oracle's y: the FP32 bound of the
    module docstring plus the oracle's own FP64 rounding (n 2^-53 sum |a x|)."""
    u_store = U32 if precision == "fp32" else 2.0 ** -8 + U32
    gamma = terms * U32 / (1.0 - terms * U32)
    return (u_store + U32 + gamma + terms * 2.0 ** -53) * absax


def _u_eff(terms: np.ndarray, precision: str) -> np.ndarray:
    """The kernels' per-row unit of the cancellation test: (u_store + (n + 2) 2^-24) widened by
    1.0001 for its FP32 evaluation."""
    u_store = U32 if precision == "fp32" else 2.0 ** -8 + U32
    return (u_store + (terms + 2.0) * U32) * 1.0001


def check_mixed_against_oracle(A: HBDIA, x: np.ndarray, precision: str, ratio, label: str) -> dict:
    """One reduced-precision SpMV checked row by row against the oracle's FP64 y: rows the kernels
    flag for FP64 equal the oracle bit for bit, every other row lies within the forward error bound
    and within the relative-error guarantee u_eff(n) * ratio (rel_error_bound for the longest row)."""
    N = A.getNumRows()
    want = oracle_y(A, x, np.zeros(N))
    M = HBDIAMixed(A, precision, ratio)
    X, Y = HBDIAVector(x), HBDIAVector(np.zeros(N))
    M.spmv(X, Y)
    A.synchronize()
    y = Y.download()
    flagged = M.fp64_mask()
    assert flagged.sum() == M.fp64_rows()
    assert y[flagged].tobytes() == want[flagged].tobytes(), "rows recomputed in FP64 differ from the oracle"
    absax, terms = _row_stats(A, x)
    err = np.abs(y - want)
    bound = _bound(absax, terms, precision)
    assert np.all(err <= bound), f"max excess over the forward bound {np.max(err - bound)}"
    kept = ~flagged & (terms > 0)
    rel = err[kept] / np.abs(want[kept])
    # the guarantee against the exact y, widened by the oracle's FP64 rounding
    fp64 = terms[kept] * 2.0 ** -53 * absax[kept] / np.abs(want[kept])
    lim = _u_eff(terms[kept], precision) * M.cancellation_ratio * (1 + 1e-9) + 2 * fp64
    assert np.all(rel <= lim), f"kept-row guarantee exceeded by {np.max(rel - lim)}"
    assert np.all(rel <= M.rel_error_bound * (1 + 1e-9) + 2 * fp64)
    out = {"max_rel": float(rel.max(initial=0.0)), "rms_rel": float(np.sqrt(np.mean(rel ** 2))) if rel.size else 0.0,
           "fp64_rows": int(flagged.sum()), "ratio": M.cancellation_ratio, "bound": M.rel_error_bound}
    print(f"  {label} {precision}: {out}", flush=True)
    return out


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("n,bw", [(24, 64), (20, 128)])
def test_mixed_precision_against_the_oracle(gpu, precision, n, bw):
    print(f"case: mixed {precision} n={n} bw={bw}", flush=True)
    A = HBDIA.stencil(n, 27, 0.005, seed=0xC3)
    A.convertToHBDIAFormat(bw, 2)
    N = n ** 3
    x = gen_vector(N, 0xC4)
    want = oracle_y(A, x, np.zeros(N))
    r4 = check_mixed_against_oracle(A, x, precision, 1e4, "ratio 1e4")
    # with ratio below 1 every row with terms is recomputed in FP64: the result is the oracle's
    M1 = HBDIAMixed(A, precision, 1e-30)
    Y1 = HBDIAVector(np.zeros(N))
    M1.spmv(HBDIAVector(x), Y1)
    A.synchronize()
    assert Y1.download().tobytes() == want.tobytes()
    assert M1.fp64_rows() == N
    # with an enormous ratio only rows whose error bound reaches |sum| (sign not certain) are
    # recomputed: never more than at ratio 1e4; the forward bound still holds
    r30 = check_mixed_against_oracle(A, x, precision, 1e30, "ratio 1e30")
    assert r30["fp64_rows"] <= r4["fp64_rows"]
    # the default ratio guarantees 1e-2 relative error on every kept row
    rd = check_mixed_against_oracle(A, x, precision, None, "default ratio")
    assert rd["bound"] <= HBDIAMixed.DEFAULT_REL_TOLERANCE * (1 + 1e-12)
    assert rd["max_rel"] <= HBDIAMixed.DEFAULT_REL_TOLERANCE


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_precision_config_192_against_the_oracle(gpu, precision):
    """The precision workload at its configured size: the 192^3 27-point stencil plus 0.5 % random
    scatter, block width 64, threshold 2, x ~ U[-1,1], FP32 / BF16 storage, rows whose cancellation
    ratio exceeds 1e4 recomputed in FP64.  Every row against the oracle's FP64 y (bit-exact for the
    recomputed rows, the proved bounds for the others); BF16 also at ratio 64, where the guarantee
    u_eff * ratio is 0.25."""
    print(f"case: precision 192^3 {precision}", flush=True)
    A = HBDIA.stencil(192, 27, 0.005, seed=0xC3)
    A.convertToHBDIAFormat(64, 2)
    x = gen_vector(192 ** 3, 0xC4)
    r = check_mixed_against_oracle(A, x, precision, 1e4, "192^3 ratio 1e4")
    assert 0 < r["fp64_rows"] < 192 ** 3 // 4
    if precision == "bf16":
        r64 = check_mixed_against_oracle(A, x, precision, 64.0, "192^3 ratio 64")
        assert r64["max_rel"] <= r64["bound"]


def test_mixed_precision_flags_cancelling_rows(gpu):
    # rows built to cancel exactly in FP64: sum a x = 0 with |a x| terms of 1
    n = 300
    rows, cols, vals = [], [], []
    for r in range(n):
        for d, v in ((-1, 1.0), (0, -2.0), (1, 1.0)):
            c = r + d
            if 0 <= c < n:
                rows.append(r)
                cols.append(c)
                vals.append(v)
    A = HBDIA(np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals), n, n)
    A.convertToHBDIAFormat(64, 1)
    x = np.full(n, 1.0 / 3.0)  # interior rows sum to 0 exactly in FP64, not after rounding to BF16 storage
    y64 = oracle_y(A, x, np.zeros(n))
    M = HBDIAMixed(A, "bf16", 1e4)
    X, Y = HBDIAVector(x), HBDIAVector(np.zeros(n))
    M.spmv(X, Y)
    A.synchronize()
    y = Y.download()
    assert M.fp64_rows() >= n - 2
    np.testing.assert_array_equal(y[1:-1], y64[1:-1])


def test_mixed_precision_graph_replay(gpu):
    from hbdia_amd import Graph, Stream
    A = HBDIA.stencil(16, 27, 0.005, seed=9)
    A.convertToHBDIAFormat(64, 2)
    N = 16 ** 3
    M = HBDIAMixed(A, "bf16", 1e4)
    X = DeviceArray.from_host(gen_vector(N, 5))
    Y = DeviceArray(N, np.float64, zero=True)
    M.spmv_raw(X.ptr, Y.ptr)
    A.synchronize()
    eager = Y.download()
    # capture on the matrix's own stream is not exposed; capture a raw launch on a new stream
    s = Stream()
    from hbdia_amd._native import hip, check_hip
    ptrs = A.device_ptrs()
    import ctypes
    ws = DeviceArray(int(hip.hbdia_mixed_workspace_bytes(N)) // 8, np.uint64, zero=True)
    vlp = DeviceArray(A.info()["num_values"], np.uint16)
    check_hip(hip.hbdia_convert_values(ptrs["values"], vlp.ptr, vlp.size, 2, s.handle))
    clp = DeviceArray(max(A.info()["num_coo"], 1), np.uint16)
    check_hip(hip.hbdia_convert_values(ptrs["coo_values"], clp.ptr, A.info()["num_coo"], 2, s.handle))
    s.synchronize()

    def launch():
        check_hip(hip.hbdia_hbdia_spmv_mixed(2, vlp.ptr, ptrs["values"], ptrs["offsets"], ptrs["block_start"],
                                             A.getNumBlocks(), 64, ptrs["coo_rows"], ptrs["coo_cols"], clp.ptr,
                                             ptrs["coo_values"], ptrs["coo_group_ptr"], X.ptr, Y.ptr, N, 1e4, ws.ptr,
                                             s.handle))
    g = Graph.capture(s, launch)
    for rep in range(3):
        Y.upload(np.full(N, float(rep)))
        g.launch(s)
        s.synchronize()
        np.testing.assert_array_equal(Y.download(), eager)
