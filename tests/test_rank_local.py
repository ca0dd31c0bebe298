"""CPU tests of the rank-local form: one rank of a row-wise partition, run on
one GPU with its halo columns in the COO part (no halo exchange).

The oracle's restatement is pinned by what the mathematics fixes: with small
integer values and x every sum is exact, so each rank's y must equal the
global y on the rank's rows exactly (a dropped, doubled or misplaced term, or
a wrong halo index, changes it); the blocked part plus the rank-local COO
part must hold exactly the rank's entries; every halo index must land on the
global column it names.  The product's host state is then compared with the
oracle's, array for array.  No GPU calls.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hbdia_amd import HBDIA
from hbdia_amd import HBDIAError

INT_MIN = -(2 ** 31)


def integer_stencil(n: int, noise: float, seed: int):
    """The reference generator's pattern with small non-zero integer values."""
    r, c, _, N = oracle.stencil27(n, n, n, noise, seed)
    rng = np.random.default_rng(seed)
    v = rng.integers(1, 5, len(r)) * rng.choice([-1.0, 1.0], len(r))
    return r, c, v.astype(np.float64), N


def x_ext_of(loc: dict, x: np.ndarray) -> np.ndarray:
    """[left halo | local | right halo] gathered from the global x."""
    L, n, r0 = loc["halo_left"], loc["num_rows"], loc["first_row"]
    h = loc["halo_cols"]
    return np.concatenate([x[h[:L]], x[r0:r0 + n], x[h[L:]]]).astype(x.dtype)


CASES = [(2, 8, 2), (3, 32, 1), (5, 64, 3), (4, 16, 2), (1, 32, 2), (7, 5, 2)]


@pytest.mark.parametrize("size,bw,thr", CASES)
def test_oracle_rank_local_exact_against_global(size, bw, thr):
    r, c, v, N = integer_stencil(12, 0.03, 1)
    x = np.random.default_rng(2).integers(-8, 9, N).astype(np.float64)
    hb = oracle.convert(r, c, v, N, N, bw, thr)
    y_global = oracle.hbdia_spmv(hb, x, np.zeros(N), N)
    np.testing.assert_array_equal(y_global, oracle.coo_spmv_longdouble(r, c, v, x, N))
    for rank in range(size):
        loc = oracle.rank_local(r, c, v, N, N, rank, size, bw, thr)
        y = oracle.rank_local_spmv(loc, x_ext_of(loc, x), np.zeros(loc["num_rows"]))
        r0 = loc["first_row"]
        np.testing.assert_array_equal(y, y_global[r0:r0 + loc["num_rows"]], err_msg=f"rank {rank}")


@pytest.mark.parametrize("size,bw,thr", CASES[:4])
def test_oracle_rank_local_holds_exactly_the_rank_entries(size, bw, thr):
    r, c, v, N = integer_stencil(10, 0.05, 3)
    parts = oracle.partition_row_wise(r, c, v, N, size)
    # partition invariant: ranks in order give back the input, each in input order
    rows = np.concatenate([p["rows"] + p["first_row"] for p in parts])
    order = np.argsort(np.concatenate([np.nonzero((r >= p["first_row"]) & (r < p["first_row"] + p["num_rows"]))[0]
                                       for p in parts]), kind="stable")
    np.testing.assert_array_equal(rows[order], r)
    for rank in range(size):
        loc = oracle.rank_local(r, c, v, N, N, rank, size, bw, thr)
        n, r0, L = loc["num_rows"], loc["first_row"], loc["halo_left"]
        hb = loc["hb"]
        got = []
        bs = hb["block_start"]
        for b in range(len(bs) - 1):
            for k in range(bs[b], bs[b + 1]):
                for lane in range(bw):
                    row = b * bw + lane
                    val = hb["values"][k * bw + lane]
                    col = row + int(hb["offsets"][k])
                    if row < n and val != 0 and 0 <= col < n:
                        got.append((row + r0, col + r0, val))
        ext = np.concatenate([loc["halo_cols"][:L], np.arange(r0, r0 + n), loc["halo_cols"][L:]])
        for row, col, val in zip(loc["coo_rows"], loc["coo_cols"], loc["coo_values"]):
            got.append((int(row) + r0, int(ext[int(col) + L]), val))
        p = loc["part"]
        want = sorted(zip((p["rows"] + r0).tolist(), p["cols"].tolist(), p["values"].tolist()))
        assert sorted(got) == want, f"rank {rank}"
        assert np.all(np.diff(loc["coo_rows"]) >= 0)


def test_oracle_request_lists_and_vector_offsets():
    r, c, v, N = integer_stencil(9, 0.08, 4)
    size = 3
    for rank in range(size):
        loc = oracle.rank_local(r, c, v, N, N, rank, size, 16, 2)
        q = N // size
        owner_lo = [p * q for p in range(size)]
        owner_hi = [(p + 1) * q if p < size - 1 else N for p in range(size)]
        covered = []
        for p, lst in enumerate(loc["process_data_ranges"]):
            for s, e in lst:
                assert owner_lo[p] <= s < e <= owner_hi[p]
                covered.extend(range(s, e))
        assert covered == [x for s, e in loc["ranges"] for x in range(s, e)]
        assert loc["total_buffer_size"] == len(covered)
        assert set(loc["part"]["cols"].tolist()) <= set(covered)
        n, r0, L = loc["num_rows"], loc["first_row"], loc["halo_left"]
        ext = np.concatenate([loc["halo_cols"][:L], np.arange(r0, r0 + n), loc["halo_cols"][L:]])
        vo = oracle.vector_offsets(loc["hb"], loc["process_data_ranges"], rank, size, N, n)
        hb = loc["hb"]
        bs = hb["block_start"]
        for b in range(len(bs) - 1):
            for k in range(bs[b], bs[b + 1]):
                g = r0 + b * 16 + int(hb["offsets"][k])
                if vo[k] == INT_MIN:
                    assert g not in set(covered)
                elif g < 0:
                    assert vo[k] == g
                else:
                    assert ext[vo[k]] == g


def assert_same_local(A: HBDIA, loc: dict, x_dtype=np.float64):
    hb = A.getHBDIA()
    ob = loc["hb"]
    for k in ["values", "offsets", "block_start", "coo_rows", "coo_cols", "coo_values", "row_ptr", "hist_blocks",
              "hist_nnz"]:
        np.testing.assert_array_equal(hb[k], ob[k], err_msg=k)
    lay = A.getLocalLayout()
    assert lay["first_row"] == loc["first_row"]
    assert lay["halo_left"] == loc["halo_left"] and lay["halo_right"] == loc["halo_right"]
    assert lay["folded"] == loc["folded"]
    np.testing.assert_array_equal(lay["halo_cols"], loc["halo_cols"])
    np.testing.assert_array_equal(lay["data_ranges"], np.asarray(loc["ranges"], np.int32).reshape(-1, 2))
    np.testing.assert_array_equal(lay["coo_rows"], loc["coo_rows"])
    np.testing.assert_array_equal(lay["coo_cols"], loc["coo_cols"])
    np.testing.assert_array_equal(lay["coo_values"], loc["coo_values"])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("size,bw,thr", [(4, 32, 2), (3, 64, 1), (5, 8, 3), (2, 7, 2)])
def test_product_rank_local_layout_matches_oracle(dtype, size, bw, thr):
    r, c, v, N = oracle.stencil27(14, 14, 14, 0.02, 3, dtype=dtype)
    G = HBDIA(r, c, v, N, N)
    for rank in range(size):
        A = HBDIA.createRankLocalMatrix(G, rank, size, bw, thr)
        loc = oracle.rank_local(r, c, v, N, N, rank, size, bw, thr)
        R, C, V = A.getCOO()
        np.testing.assert_array_equal(R, loc["part"]["rows"])
        np.testing.assert_array_equal(C, loc["part"]["cols"])
        np.testing.assert_array_equal(V, loc["part"]["values"])
        A.buildLocalCOO()
        assert_same_local(A, loc)
        vo = oracle.vector_offsets(loc["hb"], loc["process_data_ranges"], rank, size, N, loc["num_rows"])
        np.testing.assert_array_equal(A.getVectorOffsets(), vo)


def test_product_rank_local_max_coo_and_remainder_rank():
    r, c, v, N = oracle.stencil27(11, 11, 11, 0.05, 7)
    G = HBDIA(r, c, v, N, N)
    size = 6  # 1331 rows: the last rank takes the remainder
    for rank in (0, size - 1):
        A = HBDIA.createRankLocalMatrix(G, rank, size, 16, 2, 40)
        A.buildLocalCOO()
        assert_same_local(A, oracle.rank_local(r, c, v, N, N, rank, size, 16, 2, 40))


def test_rank_local_errors():
    r, c, v, N = oracle.stencil27(6, 6, 6, 0.0, 0)
    G = HBDIA(r, c, v, N, N)
    with pytest.raises((ValueError, HBDIAError)):
        HBDIA.createRankLocalMatrix(G, 4, 4, 32, 2)
    with pytest.raises((ValueError, HBDIAError)):
        HBDIA.createRankLocalMatrix(G, 0, N + 1, 32, 2)
    G.convertToHBDIAFormatOnHost(32, 2)
    with pytest.raises(Exception):
        G.buildLocalCOO()  # not a partial matrix
