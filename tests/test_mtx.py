"""Matrix Market input (reference src/Format/HBDIA.cpp:161-370), on the CPU.

The loader's rules, each checked on small files written here:
* comment lines start with '%'; "symmetric" in any of them mirrors every
  off-diagonal entry, the mirror right after the original (also for
  "skew-symmetric": the reference does not negate);
* the first other line holds rows, cols, entries; at most that many entry lines
  are read; a line counts when it starts with two integers; a missing value is 1;
* indices are 1-based in the file;
* without dimensions the entries are sorted by (row, col) with std::sort on an
  index array (for fewer than 17 entries that is an insertion sort: equal keys
  keep their order); with dimensions, indices are checked and file order kept.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hbdia_amd import HBDIA


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def coo(A):
    r, c, v = A.getCOO()
    return r.tolist(), c.tolist(), v.tolist()


def test_general_file_is_zero_based_and_sorted(tmp_path):
    f = write(tmp_path, "g.mtx", "%%MatrixMarket matrix coordinate real general\n% a comment\n"
                                 "4 5 5\n3 2 1.5\n1 1 -2\n4 5 3e-1\n1 4 7\n2 2 0.25\n")
    A = HBDIA()
    assert A.loadMTX(f)
    assert (A.getNumRows(), A.getNumCols(), A.getNumNonZeros()) == (4, 5, 5)
    assert coo(A) == ([0, 0, 1, 2, 3], [0, 3, 1, 1, 4], [-2.0, 7.0, 0.25, 1.5, 0.3])


def test_symmetric_mirrors_off_diagonal_and_pattern_values(tmp_path):
    f = write(tmp_path, "s.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 3\n1 1\n3 1\n2 3\n")
    A = HBDIA()
    assert A.loadMTX(f)
    r, c, v = coo(A)
    assert A.getNumNonZeros() == 5
    assert list(zip(r, c)) == [(0, 0), (0, 2), (1, 2), (2, 0), (2, 1)]
    assert v == [1.0] * 5
    g = write(tmp_path, "k.mtx", "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 4\n")
    B = HBDIA()
    B.loadMTX(g)
    assert coo(B) == ([0, 1], [1, 0], [4.0, 4.0])  # mirrored with the same sign, as the reference does


def test_duplicates_keep_file_order_and_junk_lines_do_not_count(tmp_path):
    f = write(tmp_path, "d.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 4\n"
                                 "2 2 1\n\nnot an entry\n2 2 2\n1 3 5\n2 2 3\n9 9 9\n")
    A = HBDIA()
    A.loadMTX(f)
    r, c, v = coo(A)
    # 4 entries counted (blank and text lines skipped, the line after the 4th not read)
    assert (r, c, v) == ([0, 1, 1, 1], [2, 1, 1, 1], [5.0, 1.0, 2.0, 3.0])
    A.removeCOODuplicates()
    assert coo(A) == ([0, 1], [2, 1], [5.0, 6.0])


def test_short_file_and_empty_file(tmp_path):
    f = write(tmp_path, "short.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 10\n1 1 2\n")
    A = HBDIA()
    assert A.loadMTX(f)
    assert A.getNumNonZeros() == 1 and A.getNumRows() == 3
    g = write(tmp_path, "none.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 0\n")
    assert not HBDIA().loadMTX(g)
    assert not HBDIA().loadMTX(str(tmp_path / "missing.mtx"))


def test_with_dimensions_checks_indices_and_keeps_file_order(tmp_path):
    f = write(tmp_path, "g.mtx", "%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n3 1 1\n1 1 2\n2 3 4\n")
    A = HBDIA()
    assert A.loadMTX(f, 10, 12)
    assert (A.getNumRows(), A.getNumCols()) == (10, 12)
    assert coo(A) == ([2, 0, 0, 1, 2], [0, 2, 0, 2, 1], [1.0, 1.0, 2.0, 4.0, 4.0])
    with pytest.raises(IndexError, match="Row index 2"):
        HBDIA().loadMTX(f, 2, 12)
    with pytest.raises(IndexError, match="Column index 2"):
        HBDIA().loadMTX(write(tmp_path, "h.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 1\n1 3 1\n"), 3, 2)
    assert not HBDIA().loadMTX(f, 0, 5)


def test_random_file_against_numpy_and_conversion(tmp_path):
    rng = np.random.default_rng(17)
    n, m = 300, 5000
    keys = rng.choice(n * n, size=m, replace=False)
    r, c = keys // n, keys % n
    v = rng.standard_normal(m)
    lines = "".join(f"{a + 1} {b + 1} {x:.17g}\n" for a, b, x in zip(r, c, v))
    f = write(tmp_path, "rand.mtx", f"%%MatrixMarket matrix coordinate real general\n{n} {n} {m}\n" + lines)
    A = HBDIA()
    assert A.loadMTX(f)
    order = np.lexsort((c, r))
    R, C, V = A.getCOO()
    np.testing.assert_array_equal(R, r[order])
    np.testing.assert_array_equal(C, c[order])
    np.testing.assert_array_equal(V, v[order])  # %.17g round-trips exactly
    # the loaded matrix converts exactly as the oracle converts the same arrays
    A.convertToHBDIAFormatOnHost(64, 2)
    want = oracle.convert(R, C, V, n, n, 64, 2)
    got = A.getHBDIA()
    for key in ("values", "offsets", "block_start", "coo_rows", "coo_cols", "coo_values"):
        np.testing.assert_array_equal(got[key], want[key], err_msg=key)
