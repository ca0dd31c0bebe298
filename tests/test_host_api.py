"""CPU tests of the product's host side against the oracle: the 27-point
generator, the COO constructors and the parallel COO -> HBDIA conversion
(without device upload), the seeded benchmark generators, error behaviour,
and the exported C ABI.  No GPU calls."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from hbdia_amd import HBDIA, ExecutionMode, _native

ROOT = Path(__file__).resolve().parents[1]
KEYS = ["values", "offsets", "block_start", "coo_rows", "coo_cols", "coo_values", "row_ptr", "hist_blocks",
        "hist_nnz"]


def assert_same_layout(A: HBDIA, ob: dict):
    hb = A.getHBDIA()
    for k in KEYS:
        np.testing.assert_array_equal(hb[k], ob[k], err_msg=k)
    assert A.getNumberDiagonals() == ob["num_diagonals"]
    assert A.getNumBlocks() == ob["num_blocks"]


@pytest.mark.parametrize("n,noise,it", [(8, 0.0, 0), (11, 0.07, 5), (16, 0.3, 2), (5, 2.0, 9)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_generator_matches_oracle(n, noise, it, dtype):
    r, c, v, N = oracle.stencil27(n, n, n, noise, it, dtype=dtype)
    A = HBDIA(dtype=dtype)
    A.create3DStencil27Point(n, n, n, noise, it)
    R, C, V = A.getCOO()
    np.testing.assert_array_equal(R, r)
    np.testing.assert_array_equal(C, c)
    np.testing.assert_array_equal(V, v)
    assert A.getNumRows() == N == A.getNumCols()


@pytest.mark.parametrize("bw,thr,maxcoo", [(32, 2, 10 ** 9), (64, 2, 10 ** 9), (1, 1, 10 ** 9), (7, 3, 10 ** 9),
                                           (128, 8, 10 ** 9), (16, 4, 50), (48, 100, 10 ** 9)])
def test_conversion_matches_oracle_stencil(bw, thr, maxcoo):
    r, c, v, N = oracle.stencil27(13, 13, 13, 0.05, 1)
    A = HBDIA(r, c, v, N, N)
    A.convertToHBDIAFormatOnHost(bw, thr, maxcoo)
    assert_same_layout(A, oracle.convert(r, c, v, N, N, bw, thr, maxcoo))


def test_conversion_unsorted_input_with_duplicates():
    rng = np.random.default_rng(5)
    n, m, nnz = 300, 280, 5000
    rows = rng.integers(0, n, nnz).astype(np.int32)
    cols = np.clip(rows + rng.integers(-6, 7, nnz), 0, m - 1).astype(np.int32)
    rows[:100] = rows[100:200]  # force duplicate (row, col) pairs
    cols[:100] = cols[100:200]
    vals = rng.standard_normal(nnz)
    for bw, thr in [(16, 3), (64, 1), (5, 2)]:
        A = HBDIA(rows, cols, vals, n, m)
        A.convertToHBDIAFormatOnHost(bw, thr)
        assert_same_layout(A, oracle.convert(rows, cols, vals, n, m, bw, thr))


def test_conversion_rows_beyond_last_block_and_wide_matrix():
    # numCols < numRows: blocks cover min(N, M) rows only; the rest go to COO
    rows = np.array([0, 1, 5, 9, 9, 3], np.int32)
    cols = np.array([0, 1, 2, 3, 0, 3], np.int32)
    vals = np.arange(1.0, 7.0)
    A = HBDIA(rows, cols, vals, 10, 4)
    A.convertToHBDIAFormatOnHost(2, 1)
    assert_same_layout(A, oracle.convert(rows, cols, vals, 10, 4, 2, 1))


def test_three_argument_constructor_infers_shape():
    A = HBDIA([0, 2, 1], [3, 0, 1], [1.0, 2.0, 3.0])
    assert A.getNumRows() == 3 and A.getNumCols() == 4 and A.getNumNonZeros() == 3


def test_constructor_errors():
    with pytest.raises(ValueError):
        HBDIA([0, 1], [0], [1.0, 2.0], 2, 2)
    with pytest.raises(IndexError):
        HBDIA([0, 5], [0, 1], [1.0, 2.0], 3, 3)
    with pytest.raises(IndexError):
        HBDIA([0, 1], [0, -1], [1.0, 2.0], 3, 3)
    with pytest.raises(ValueError):
        HBDIA([0], [0], [1.0], 0, 3)


def test_remove_duplicates_sums_in_input_order():
    A = HBDIA([1, 0, 1, 1], [2, 0, 2, 0], [0.1, 5.0, 0.2, 3.0], 2, 3)
    A.removeCOODuplicates()
    r, c, v = A.getCOO()
    assert r.tolist() == [0, 1, 1] and c.tolist() == [0, 0, 2]
    assert v.tolist() == [5.0, 3.0, (0.0 + 0.1) + 0.2]


def test_stencil_scatter_generator_properties():
    n, frac = 12, 0.01
    A = HBDIA.stencil(n, 27, frac, seed=123)
    r, c, v = A.getCOO()
    N = n ** 3
    assert np.all(np.diff(r) >= 0)
    assert len(v) == (3 * n - 2) ** 3 + int(frac * 27 * N)
    assert len(set(zip(r.tolist(), c.tolist()))) == len(v)
    diag = r == c
    assert diag.sum() == N
    off = ~diag
    assert np.all((v[off] >= -1.0) & (v[off] < 0.0))
    absum = np.bincount(r[off], weights=np.abs(v[off]), minlength=N)
    np.testing.assert_allclose(v[diag], 1.0 + absum, rtol=1e-14)
    # scatter entries are off-stencil
    ri, rj, rk = r % n, (r // n) % n, r // (n * n)
    ci, cj, ck = c % n, (c // n) % n, c // (n * n)
    near = (np.abs(ri - ci) <= 1) & (np.abs(rj - cj) <= 1) & (np.abs(rk - ck) <= 1)
    assert (~near).sum() == int(frac * 27 * N)
    # same seed, same matrix; other seed, other matrix
    r2, c2, v2 = HBDIA.stencil(n, 27, frac, seed=123).getCOO()
    assert np.array_equal(c, c2) and np.array_equal(v, v2)
    assert not np.array_equal(c, HBDIA.stencil(n, 27, frac, seed=124).getCOO()[1])


def test_seven_point_and_banded_generators():
    A = HBDIA.stencil(6, 7, 0.0, seed=1)
    r, c, v = A.getCOO()
    n = 6
    assert len(v) == 7 * n ** 3 - 6 * n ** 2  # each of the 6 faces loses one neighbour per boundary row
    B = HBDIA.banded(1000, 5, 64, 0.01, seed=3)
    r, c, v = B.getCOO()
    offs = np.unique((c - r)[np.abs(c - r) <= 64])
    assert 0 in offs
    assert np.all(np.diff(r) >= 0)


def test_banded_generator_offband_count_and_batch_shapes():
    # floor(offband * band entries) extra entries, each off the band, none duplicated
    n, k = 3000, 7
    B = HBDIA.banded(n, k, 64, 0.01, seed=12)
    r, c, v = B.getCOO()
    d = c - r
    offs = np.unique(d[np.abs(d) <= 64])
    band_cols = {int(o) for o in offs}
    # offsets of the band are those present on (almost) every row
    counts = {int(o): int((d == o).sum()) for o in offs}
    band = sorted(o for o, cnt in counts.items() if cnt >= n - abs(o) - 1)
    assert len(band) == k and 0 in band
    band_nnz = sum(n - abs(o) for o in band)
    extra = len(v) - band_nnz
    assert extra == int(0.01 * band_nnz)
    keys = r.astype(np.int64) * n + c
    assert len(np.unique(keys)) == len(keys)
    # diagonal dominance: diag = 1 + sum |off-diagonal| of the row
    diag = v[d == 0]
    off_abs = np.bincount(r[d != 0], weights=np.abs(v[d != 0]), minlength=n)
    np.testing.assert_allclose(diag, 1.0 + off_abs, rtol=1e-12)
    del band_cols
    # the batch set: matrix i is banded(shape_i); shapes within the spec
    ms = HBDIA.banded_batch(count=6, min_rows=1000, max_rows=20000, seed=99)
    for i, m in enumerate(ms):
        rows, nk, sd = HBDIA.banded_batch_shape(i, 1000, 20000, seed=99)
        assert 1000 <= rows <= 20000 and 3 <= nk <= 9
        ref = HBDIA.banded(rows, nk, 64, 0.01, seed=sd)
        for a, b in zip(m.getCOO(), ref.getCOO()):
            np.testing.assert_array_equal(a, b)


def test_batch_rejects_unconverted_and_narrow_blocks():
    from hbdia_amd import HBDIABatch, HBDIAError
    A = HBDIA.banded(500, 3, 8, 0.0, seed=1)
    with pytest.raises(HBDIAError, match="not converted"):
        HBDIABatch([A])
    A.convertToHBDIAFormatOnHost(32, 2)
    with pytest.raises(HBDIAError, match="multiples of 64"):
        HBDIABatch([A])
    with pytest.raises(ValueError):
        HBDIABatch([])


def test_generator_thread_count_independent():
    from hbdia_amd import set_num_threads
    set_num_threads(1)
    a = HBDIA.stencil(10, 27, 0.02, seed=9).getCOO()
    set_num_threads(8)
    b = HBDIA.stencil(10, 27, 0.02, seed=9).getCOO()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def _declared(header: Path) -> list[str]:
    text = header.read_text()
    return re.findall(r"^\s*(?:int|size_t|const char\*)\s+(hbdia_\w+)\s*\(", text, flags=re.M)


@pytest.mark.parametrize("header,lib,bound", [("hbdia_hip.h", "libhbdia_hip.so", _native.HIP_FUNCTIONS),
                                              ("hbdia_capi.h", "libhbdia.so", _native.CAPI_FUNCTIONS)])
def test_c_abi_exports_every_declared_symbol(header, lib, bound):
    names = _declared(ROOT / "include" / header)
    assert len(names) > 10
    out = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "hbdia_amd" / "lib" / lib)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(names) == set(bound), set(names) ^ set(bound)


def test_kernel_calls_without_gpu_report_errors_not_crashes():
    if _native.gpu_available():
        pytest.skip("a GPU is visible")
    import ctypes
    p = ctypes.c_void_p()
    assert _native.hip.hbdia_malloc(ctypes.byref(p), 64) != 0
    assert _native.hip.hbdia_last_error()
    # invalid arguments are rejected before any device call
    assert _native.hip.hbdia_bdia_spmv_f64(None, None, None, None, None, 10, 1, 0, None) == 1


@pytest.mark.parametrize("n", [1, 15, 16, 1000, 4096, 1 << 21, 128 ** 3 + 1, 3_000_000])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_multi_ld_pads_power_of_two_strides(n, dtype):
    from hbdia_amd import multi_ld
    size = np.dtype(dtype).itemsize
    ld = multi_ld(n, dtype)
    assert ld >= n and (ld * size) % 128 == 0          # whole 128-byte lines
    assert (ld * size) % 4096 != 0                      # never a multiple of 4 KiB
    assert ld - n < 2 * 128 // size                     # at most one line past the rounding
    if (n * size) % 4096 == 0:
        assert ld == n + 128 // size


def test_batch_tiles_share_an_index_for_equal_offset_lists():
    """hbdia_batch_make_tiles (host code) gives consecutive tiles whose blocks hold equal offset
    lists one offsets index, so the batch kernel can pair them from their records alone; a
    different list of the same length keeps its own index."""
    import ctypes

    class Desc(ctypes.Structure):
        _fields_ = [("values", ctypes.c_int64), ("rows", ctypes.c_int64), ("first_tile", ctypes.c_int64),
                    ("offsets", ctypes.c_int32), ("block_start", ctypes.c_int32), ("coo", ctypes.c_int32),
                    ("coo_nnz", ctypes.c_int32), ("group_ptr", ctypes.c_int32), ("num_rows", ctypes.c_int32),
                    ("num_blocks", ctypes.c_int32), ("block_width", ctypes.c_int32)]

    a = list(range(-1, 10))          # 11 offsets
    b = list(range(0, 11))           # 11 offsets, another list
    lists = [a, a, b, b, a]          # blocks 0..4, width 64: one tile each
    offs = np.array(sum(lists, []) + [0] * 16, np.int32)
    starts = np.cumsum([0] + [len(l) for l in lists]).astype(np.int32)
    d = Desc(values=0, rows=0, first_tile=0, offsets=0, block_start=0, coo=0, coo_nnz=0, group_ptr=0,
             num_rows=320, num_blocks=5, block_width=64)
    tiles = np.zeros((5, 20), np.int32)
    rc = _native.hip.hbdia_batch_make_tiles(ctypes.byref(d), 1, offs.ctypes.data, starts.ctypes.data, None, 8,
                                            tiles.ctypes.data, 5)
    assert rc == 0, _native.hip.hbdia_last_error()
    at = tiles[:, 6]                 # the offsets index of each tile
    assert list(at) == [0, 0, 22, 22, 44]
    assert list(tiles[:, 5]) == [11] * 5
    for t, l in zip(tiles, lists):   # every index names the tile's own list
        assert list(offs[t[6]:t[6] + 11]) == l
        assert list(t[10:20]) == l[:10]
