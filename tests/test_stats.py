"""Statistics and storage report (reference benchmarking/benchmark.cpp:144-230), on the CPU.

Pinned by closed forms for the reference's 8^3 27-point stencil at block width
32, threshold 1 (16 blocks; 4 boundary blocks with 18 offsets and 484 entries,
12 with 27 offsets and 726 entries; 10,648 = 22^3 entries; 27 diagonals), and
for other matrices by the oracle's restatement over the oracle's conversion.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hbdia_amd import HBDIA


def test_reference_stencil_closed_forms():
    A = HBDIA()
    A.create3DStencil27Point(8, 8, 8, 0.0)
    A.convertToHBDIAFormatOnHost(32, 1)
    r = A.storage_report()
    n, nnz = 512, 22 ** 3
    assert (r["rows"], r["nnz"], r["number_diagonals"], r["num_blocks"]) == (n, nnz, 27, 16)
    assert r["histogram"] == [(18, 4, 4 * 484), (27, 12, 12 * 726)]
    assert r["stored_offsets"] == 4 * 18 + 12 * 27 == 396
    assert r["hbdia_bytes"] == 396 * 32 * 8 + 17 * 4 + 396 * 4  # 103,028: no fallback
    assert r["coo_bytes"] == nnz * 16
    assert r["dia_bytes"] == 27 * n * 8 + 27 * 4
    # slices of 32 rows = half a z-plane: the 4 slices of the z = 0 and z = 7 planes have rows of at most
    # 18 entries, the other 12 of 27
    assert r["sell_bytes"] == (4 * 18 + 12 * 27) * 32 * 12 + 16 * 4
    assert r["csr_bytes"] == nnz * 12 + (n + 1) * 4
    assert r["sell_applies"]
    assert r["hbdia_vs_coo"] == pytest.approx(103028 / (nnz * 16))
    assert r["dia_vs_coo"] == pytest.approx(r["dia_bytes"] / r["coo_bytes"])
    # the oracle's restatement meets the same closed forms
    rr, cc, vv, _ = oracle.stencil27(8, 8, 8, 0.0, 0)
    o = oracle.storage_report(oracle.convert(rr, cc, vv, n, n, 32, 1), n, n, nnz, np.bincount(rr, minlength=n), 32, 1)
    for k in ("hbdia_bytes", "coo_bytes", "dia_bytes", "sell_bytes", "csr_bytes", "histogram", "stored_offsets"):
        assert o[k] == r[k], k


@pytest.mark.parametrize("noise,bw,thr", [(0.05, 32, 2), (0.2, 64, 3), (0.0, 16, 1)])
def test_generator_matrices_match_the_oracle(noise, bw, thr):
    n = 10
    A = HBDIA()
    A.create3DStencil27Point(n, n, n, noise, 1)
    A.convertToHBDIAFormatOnHost(bw, thr)
    r, c, v = A.getCOO()
    hb = oracle.convert(r, c, v, n ** 3, n ** 3, bw, thr)
    per_row = np.bincount(r, minlength=n ** 3)  # what the generator counts per row
    want = oracle.storage_report(hb, n ** 3, n ** 3, len(v), per_row, bw, thr)
    got = A.storage_report()
    for k, val in want.items():
        assert got[k] == val, k
    assert got["coo_fallback_entries"] > 0 or noise == 0.0


def test_coo_constructed_matrix_has_no_row_counts_and_float_sizes():
    rng = np.random.default_rng(3)
    r = rng.integers(0, 500, 4000).astype(np.int32)
    c = np.clip(r + rng.integers(-30, 31, 4000), 0, 499).astype(np.int32)
    v = rng.standard_normal(4000).astype(np.float32)
    A = HBDIA(r, c, v, 500, 500)
    A.convertToHBDIAFormatOnHost(64, 2)
    got = A.storage_report()
    hb = oracle.convert(r, c, v, 500, 500, 64, 2)
    want = oracle.storage_report(hb, 500, 500, 4000, [], 64, 2, value_bytes=4)
    for k, val in want.items():
        assert got[k] == val, k
    assert got["value_bytes"] == 4 and got["sell_bytes"] == 8 * 4  # slice pointers only


def test_csv_has_the_reference_columns_and_histogram():
    A = HBDIA()
    A.create3DStencil27Point(8, 8, 8, 0.0)
    A.convertToHBDIAFormatOnHost(32, 1)
    text = A.storage_report_csv("3D27Stencil")
    info, hist = text.split("\n\n")
    head, row = info.strip().split("\n")
    cols = head.split(",")
    assert cols[:4] == ["name", "rows", "cols", "nnz"] and cols[-4:] == ["hbdia_vs_coo", "dia_vs_coo",
                                                                        "csr_vs_coo", "sell_vs_coo"]
    vals = dict(zip(cols, row.split(",")))
    assert vals["name"] == "3D27Stencil" and int(vals["hbdia_cost_bytes"]) == 103028
    assert hist.strip().split("\n") == ["diagonal_count,blocks_with_diagonals,nnz_in_blocks", "18,4,1936",
                                        "27,12,8712"]


def test_report_needs_a_converted_matrix():
    A = HBDIA()
    A.create3DStencil27Point(4, 4, 4, 0.0)
    with pytest.raises(Exception, match="HBDIA format"):
        A.storage_report()
