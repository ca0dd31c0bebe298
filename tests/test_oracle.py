"""Pins of the oracle (CPU restatement of the reference) against closed forms,
hand-derived layouts, invariants, brute force and the committed fixtures.

The reference ships no golden vectors of its own and cannot be built here, so
each oracle function is pinned by checks that a plausible mistake (a dropped
term, a wrong sign or index, a transposed operand) would fail.
"""
from __future__ import annotations

from collections import Counter

import numpy as np
import pytest

import oracle

WEIGHTS = {0: 26.1, 1: -1.1, 2: -0.1, 3: -0.01}


# ---------------------------------------------------------------- generator
@pytest.mark.parametrize("n", [2, 3, 5, 8, 12])
def test_stencil_nnz_closed_form(n):
    r, c, v, N = oracle.stencil27(n, n, n)
    assert N == n ** 3
    assert len(v) == (3 * n - 2) ** 3


def test_stencil_rows_sorted_and_weights():
    n = 7
    r, c, v, N = oracle.stencil27(n, n, n)
    assert np.all(np.diff(r) >= 0)
    for row in range(N):
        sel = r == row
        i, j, k = row % n, (row // n) % n, row // (n * n)
        ci, cj, ck = c[sel] % n, (c[sel] // n) % n, c[sel] // (n * n)
        assert np.all(np.abs(ci - i) <= 1) and np.all(np.abs(cj - j) <= 1) and np.all(np.abs(ck - k) <= 1)
        kind = (ci != i).astype(int) + (cj != j) + (ck != k)
        for kd, val in zip(kind, v[sel]):
            assert val == WEIGHTS[int(kd)]
        if 0 < i < n - 1 and 0 < j < n - 1 and 0 < k < n - 1:
            assert Counter(kind.tolist()) == {0: 1, 1: 6, 2: 12, 3: 8}
    assert set(np.unique(c[r == 0]).tolist()) == {0, 1, n, n + 1, n * n, n * n + 1, n * n + n, n * n + n + 1}


def test_stencil_noise_count_and_distinct_columns():
    n, noise = 10, 0.05
    r, c, v, N = oracle.stencil27(n, n, n, noise, 3)
    extra = len(v) - (3 * n - 2) ** 3
    assert extra == int(27 * N * noise)  # floor(27 N noise), HBDIA.cpp:1459
    pairs = set(zip(r.tolist(), c.tolist()))
    assert len(pairs) == len(v)  # noise never repeats a column within its row
    sv = v[np.abs(v) < 1.0 + 1e-12]
    assert np.all((sv >= -1.0) & (sv < 1.0))


def test_stencil_seed_is_42_plus_iteration():
    a = oracle.stencil27(6, 6, 6, 0.2, 0)
    b = oracle.stencil27(6, 6, 6, 0.2, 0)
    c = oracle.stencil27(6, 6, 6, 0.2, 1)
    assert all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
    assert not np.array_equal(a[1], c[1])


def test_stencil_noise_fixture(golden):
    g = golden("stencil6_noise.npz")
    r, c, v, N = oracle.stencil27(6, 6, 6, 0.2, 7)
    assert N == int(g["n"])
    np.testing.assert_array_equal(r, g["rows"])
    np.testing.assert_array_equal(c, g["cols"])
    np.testing.assert_array_equal(v, g["values"])


def test_stencil_float_values_are_casts():
    r, c, v, N = oracle.stencil27(4, 4, 4, 0.1, 0, dtype=np.float32)
    r2, c2, v2, _ = oracle.stencil27(4, 4, 4, 0.1, 0, dtype=np.float64)
    np.testing.assert_array_equal(r, r2)
    np.testing.assert_array_equal(v, v2.astype(np.float32))


# --------------------------------------------------------------- conversion
def test_config1_layout_closed_form():
    r, c, v, N = oracle.stencil27(8, 8, 8)
    hb = oracle.convert(r, c, v, N, N, 32, 2)
    counts = np.diff(hb["block_start"])
    # 16 blocks of 32 rows = 4 z-planes; planes 0 and 7 lack one z neighbour
    assert hb["num_blocks"] == 16
    assert counts.tolist() == [18, 18] + [27] * 12 + [18, 18]
    assert counts.sum() == 396 and len(hb["values"]) == 12672
    assert len(hb["coo_values"]) == 0
    assert np.all(hb["row_ptr"] == 0)
    assert hb["hist_blocks"][27] == 12 and hb["hist_blocks"][18] == 4
    assert hb["hist_nnz"].sum() == len(v)


def _print_test_matrix():
    """The 20x20 matrix of the reference's print test (tests/run_HBDIA_print_test.cpp:59-92)
    after its COO duplicates are summed and sorted by (row, col), as
    convertToDIAFormat does before that test converts to HBDIA."""
    ent = {}
    def add(i, j, val):
        ent[(i, j)] = ent.get((i, j), 0.0) + val
    for i in range(20):
        add(i, i, 5.0 + i * 0.1)
    for i in range(19):
        add(i, i + 1, 2.0 + i * 0.05)
    for i in range(1, 20):
        add(i, i - 1, 1.0 + i * 0.05)
    for i in range(0, 18, 3):
        add(i, i + 2, 0.5 + i * 0.02)
    for i, j in [(2, 15), (5, 18), (10, 3), (15, 8), (18, 19)]:
        add(i, j, 0.3)
    keys = sorted(ent)
    return (np.array([k[0] for k in keys], np.int32), np.array([k[1] for k in keys], np.int32),
            np.array([ent[k] for k in keys]))


def test_print_test_matrix_hand_derived_layout():
    r, c, v = _print_test_matrix()
    assert len(v) == 68
    hb = oracle.convert(r, c, v, 20, 20, 8, 3)
    offs = [hb["offsets"][hb["block_start"][b]:hb["block_start"][b + 1]].tolist() for b in range(3)]
    assert offs == [[-1, 0, 1, 2], [-1, 0, 1, 2], [-1, 0, 1]]
    coo = list(zip(hb["coo_rows"].tolist(), hb["coo_cols"].tolist(), hb["coo_values"].tolist()))
    assert coo == [(2, 15, 0.3), (5, 18, 0.3), (10, 3, 0.3), (15, 8, 0.3)]
    # value slots [block][offset][lane], filled from the matrix definition
    vals = hb["values"].reshape(-1, 8)
    line = 0
    for b, ob in enumerate(offs):
        for o in ob:
            for lane in range(8):
                row = 8 * b + lane
                want = 0.0
                if row < 20 and 0 <= row + o < 20:
                    want = dict(zip(zip(r.tolist(), c.tolist()), v.tolist())).get((row, row + o), 0.0)
                assert vals[line, lane] == want, (b, o, lane)
            line += 1
    assert vals[4 * 2 + 2, 18 - 16] == pytest.approx(3.2)  # the summed duplicate (18, 19)
    # cpuRowPtr gap-filling: rows 0..2 -> 0, 3..5 -> 1, 6..10 -> 2, 11..15 -> 3, 16..20 -> 4
    assert hb["row_ptr"].tolist() == [0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4]
    assert hb["num_diagonals"] == 6  # -7, -1, 0, 1, 2, 13


def test_conversion_partition_invariant():
    r, c, v, N = oracle.stencil27(9, 9, 9, 0.3, 2)
    for bw, thr in [(1, 1), (7, 2), (32, 5), (64, 40)]:
        hb = oracle.convert(r, c, v, N, N, bw, thr)
        stored = np.count_nonzero(hb["values"])
        assert stored + len(hb["coo_values"]) == len(v)  # no zero values in this input
        # every fallback entry's offset is absent from its block (or its block is past the end)
        for row, col in zip(hb["coo_rows"], hb["coo_cols"]):
            b = row // bw
            if b < hb["num_blocks"]:
                assert (col - row) not in hb["offsets"][hb["block_start"][b]:hb["block_start"][b + 1]]


def test_conversion_max_coo_cap_is_global_and_ordered():
    r, c, v = _print_test_matrix()
    # thr 3 would send 4 entries to the fallback; a cap of 2 stops after the
    # first block's two (blocks ascending, offsets ascending), so block 1
    # keeps offset -7 in the blocked part.
    hb = oracle.convert(r, c, v, 20, 20, 8, 3, max_coo_entries=2)
    offs1 = hb["offsets"][hb["block_start"][1]:hb["block_start"][2]].tolist()
    assert offs1 == [-7, -1, 0, 1, 2]
    assert list(zip(hb["coo_rows"].tolist(), hb["coo_cols"].tolist())) == [(2, 15), (5, 18)]


def test_conversion_duplicates_last_wins_in_blocks():
    rows = np.array([0, 1, 0, 1], np.int32)
    cols = np.array([0, 1, 0, 1], np.int32)
    vals = np.array([1.0, 2.0, 3.0, 4.0])
    hb = oracle.convert(rows, cols, vals, 2, 2, 2, 1)
    assert hb["values"].tolist() == [3.0, 4.0]


# --------------------------------------------------------------------- SpMV
def test_config1_closed_forms():
    r, c, v, N = oracle.stencil27(8, 8, 8)
    hb = oracle.convert(r, c, v, N, N, 32, 2)
    x = np.arange(1, N + 1, dtype=np.float64)
    y = oracle.hbdia_spmv(hb, x, np.zeros(N), N)
    idx = np.arange(N)
    i, j, k = idx % 8, (idx // 8) % 8, idx // 64
    interior = (i > 0) & (i < 7) & (j > 0) & (j < 7) & (k > 0) & (k < 7)
    np.testing.assert_allclose(y[interior], 18.22 * (idx[interior] + 1), rtol=1e-13)
    assert y[0] == pytest.approx(-73.14, rel=1e-13)
    assert y[511] == pytest.approx(11610.51, rel=1e-13)


def test_config1_fixture(golden):
    r, c, v, N = oracle.stencil27(8, 8, 8)
    hb = oracle.convert(r, c, v, N, N, 32, 2)
    y = oracle.hbdia_spmv(hb, np.arange(1, N + 1, dtype=np.float64), np.zeros(N), N)
    np.testing.assert_array_equal(y, golden("config1_y.npz")["y"])


def test_row_sums():
    r, c, v, N = oracle.stencil27(10, 10, 10, 0.05, 1)
    hb = oracle.convert(r, c, v, N, N, 16, 2)
    y = oracle.hbdia_spmv(hb, np.ones(N), np.zeros(N), N)
    want = np.bincount(r, weights=v, minlength=N)
    np.testing.assert_allclose(y, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("bw,thr", [(1, 1), (8, 3), (32, 2), (64, 2), (48, 5)])
def test_spmv_vs_longdouble_brute_force(bw, thr):
    r, c, v, N = oracle.stencil27(12, 12, 12, 0.1, 4)
    hb = oracle.convert(r, c, v, N, N, bw, thr)
    x = np.cos(np.arange(N) * 0.37)
    y = oracle.hbdia_spmv(hb, x, np.zeros(N), N)
    ref = oracle.coo_spmv_longdouble(r, c, v, x, N)
    scale = oracle.coo_spmv_longdouble(r, c, np.abs(v), np.abs(x), N)
    assert np.all(np.abs(y - ref) <= 1e-14 * scale + 1e-300)


def test_noisy_fixtures(golden):
    for name, args, xk in [("noisy16_y.npz", (16, 0.1, 0, 32, 2), "i+1"),
                           ("noisy24_y.npz", (24, 0.02, 1, 16, 3), "sin")]:
        n, noise, it, bw, thr = args
        r, c, v, N = oracle.stencil27(n, n, n, noise, it)
        hb = oracle.convert(r, c, v, N, N, bw, thr)
        x = np.arange(1, N + 1, dtype=np.float64) if xk == "i+1" else np.sin(np.arange(N, dtype=np.float64))
        y = oracle.hbdia_spmv(hb, x, np.zeros(N), N)
        g = golden(name)
        assert len(hb["coo_values"]) == int(g["ncoo"]) > 0
        np.testing.assert_array_equal(y, g["y"])


def test_kernel_semantics_skip_zero_and_out_of_range():
    # one block of width 4, offsets {-1, 0, 5}: -1 is out of range for row 0,
    # 5 is out of range for every row (numRows = 4); a zero value hides a NaN.
    hb = {"values": np.array([9.0, 1.0, 1.0, 1.0,   # offset -1
                              2.0, 0.0, 2.0, 2.0,   # offset 0 (row 1 zero)
                              7.0, 7.0, 7.0, 7.0]),  # offset 5
          "offsets": np.array([-1, 0, 5], np.int32), "block_start": np.array([0, 3], np.int32),
          "num_blocks": 1, "block_width": 4}
    x = np.array([1.0, np.nan, 3.0, 4.0])
    y = oracle.bdia_spmv(hb, x, np.full(4, -5.0), 4)
    assert y[0] == 2.0                     # 9*x[-1] skipped (col < 0)
    assert y[1] == 1.0                     # zero value skips the NaN at x[1]
    assert np.isnan(y[2])                  # 1*x[1] = NaN propagates
    assert y[3] == 1.0 * 3.0 + 2.0 * 4.0


def test_rows_of_empty_blocks_keep_y():
    hb = {"values": np.array([1.0, 1.0]), "offsets": np.array([0], np.int32),
          "block_start": np.array([0, 1, 1], np.int32), "num_blocks": 2, "block_width": 2}
    y = oracle.bdia_spmv(hb, np.array([3.0, 4.0, 5.0, 6.0, 7.0]), np.full(5, -1.0), 5)
    assert y.tolist() == [3.0, 4.0, -1.0, -1.0, -1.0]


def test_cpu_coo_writes_only_rows_with_entries():
    res = np.full(5, 7.0)
    oracle.cpu_coo_spmv(np.array([1, 1, 3]), np.array([0, 4, 2]), np.array([2.0, 3.0, 4.0]),
                        np.array([1.0, 2.0, 3.0, 4.0, 5.0]), 5, res)
    assert res.tolist() == [7.0, 2.0 * 1.0 + 3.0 * 5.0, 7.0, 12.0, 7.0]
