"""GPU tests of the COO -> HBDIA conversion on the device (hbdia_convert_coo).

convertToHBDIAFormatOnDevice must leave exactly the state of the host
conversion (reference src/Format/HBDIA.cpp:507-718): every host array, the
counts and histograms, and the device arrays prepareForGPU would upload; the
oracle's conversion pins both.  The cases cover what the rules depend on: the
threshold, the global max_coo_entries cap (0, small, large), duplicate entries
(the last in input order wins), unsorted input, rows past the last block (more
rows than columns), block widths that do and do not divide the row count,
float and double, and -- through the C ABI directly -- a global row map.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hbdia_amd import HBDIA, HBDIAVector, gen_vector, hbdiaSpMV
from hbdia_amd import kernels as K
from hbdia_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

HOST_KEYS = ("values", "offsets", "block_start", "coo_rows", "coo_cols", "coo_values", "row_ptr", "hist_blocks",
             "hist_nnz")


def fetch(ptr, n, dtype):
    return K._fetch(ptr, n, dtype) if n else np.empty(0, dtype)


def device_state(A: HBDIA) -> dict:
    """The device arrays of a converted matrix, downloaded."""
    i, p = A.info(), A.device_ptrs()
    nb, groups = i["num_blocks"], K.num_groups(i["rows"])
    out = {
        "values": fetch(p["values"], i["num_values"], A.dtype),
        "offsets": fetch(p["offsets"], i["num_offsets"] + K.OFFSETS_PAD, np.int32),
        "block_start": fetch(p["block_start"], nb + 1, np.int32),
        "block_sizes": fetch(p["block_sizes"], nb, np.int32),
        "device_coo": bool(i["device_coo"]),
    }
    if i["device_coo"]:
        out.update(coo_rows=fetch(p["coo_rows"], i["num_coo"], np.int32),
                   coo_cols=fetch(p["coo_cols"], i["num_coo"], np.int32),
                   coo_values=fetch(p["coo_values"], i["num_coo"], A.dtype),
                   coo_row_ptr=fetch(p["coo_row_ptr"], i["rows"] + 1, np.int32),
                   coo_group_ptr=fetch(p["coo_group_ptr"], groups + 1, np.int32))
    return out


def check_same(r, c, v, nr, nc, bw, thr, mc, gpu_coo=True):
    mode = 1 if gpu_coo else 0
    A = HBDIA(r, c, v, nr, nc)
    A.convertToHBDIAFormatOnHost(bw, thr, mc, False, mode)
    A.prepareForGPU()
    B = HBDIA(r, c, v, nr, nc)
    B.convertToHBDIAFormatOnDevice(bw, thr, mc, False, mode)
    ia, ib = A.info(), B.info()
    for k in ("num_blocks", "num_values", "num_offsets", "num_coo", "num_diagonals", "hist_len", "runs3_blocks",
              "device_coo", "block_width", "threshold", "has_hbdia"):
        assert ia[k] == ib[k], k
    ha, hb = A.getHBDIA(), B.getHBDIA()
    for k in HOST_KEYS:
        np.testing.assert_array_equal(hb[k], ha[k], err_msg=k)
        assert hb[k].dtype == ha[k].dtype
    # bit patterns, not just values (signed zeros, NaN payloads)
    assert hb["values"].tobytes() == ha["values"].tobytes()
    assert hb["coo_values"].tobytes() == ha["coo_values"].tobytes()
    da, db = device_state(A), device_state(B)
    for k in da:
        np.testing.assert_array_equal(db[k], da[k], err_msg="device " + k)
    # both against the oracle
    want = oracle.convert(r, c, v, nr, nc, bw, thr, mc)
    for k in HOST_KEYS:
        np.testing.assert_array_equal(hb[k], want[k], err_msg="oracle " + k)
    assert ib["num_diagonals"] == want["num_diagonals"]
    return A, B


def random_coo(nr, nc, nnz, seed, dtype=np.float64, dup=0.0, band=None):
    rng = np.random.default_rng(seed)
    r = rng.integers(0, nr, nnz).astype(np.int32)
    if band is None:
        c = rng.integers(0, nc, nnz).astype(np.int32)
    else:
        c = np.clip(r + rng.integers(-band, band + 1, nnz), 0, nc - 1).astype(np.int32)
    v = rng.standard_normal(nnz).astype(dtype)
    if dup:
        k = int(nnz * dup)
        src = rng.integers(0, nnz, k)
        r = np.concatenate([r, r[src]])
        c = np.concatenate([c, c[src]])
        v = np.concatenate([v, rng.standard_normal(k).astype(dtype)])
        perm = rng.permutation(len(v))
        r, c, v = r[perm], c[perm], v[perm]
    return r, c, v


@pytest.mark.parametrize("n,bw,thr", [(8, 32, 1), (12, 64, 2), (16, 128, 3), (20, 64, 1), (11, 7, 2)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_stencil_matches_host_and_oracle(gpu, n, bw, thr, dtype):
    print(f"case: convert stencil27 n={n} bw={bw} thr={thr} {np.dtype(dtype).name}", flush=True)
    r, c, v, _ = oracle.stencil27(n, n, n, 0.01 if n > 8 else 0.0, 3, dtype=dtype)
    A, B = check_same(r, c, v, n ** 3, n ** 3, bw, thr, 1_000_000_000)
    if (n, bw, thr) == (8, 32, 1):  # the reference's 8^3 configuration: closed-form counts
        i = B.info()
        assert (i["num_blocks"], i["num_offsets"], i["num_values"], i["num_coo"]) == (16, 396, 12672, 0)


@pytest.mark.parametrize("mc", [0, 1, 40, 400, 1_000_000_000])
@pytest.mark.parametrize("thr", [1, 3, 6])
def test_threshold_and_global_cap(gpu, mc, thr):
    print(f"case: convert banded cap={mc} thr={thr}", flush=True)
    r, c, v = random_coo(1500, 1500, 20000, seed=thr * 7 + mc % 97, band=40)
    check_same(r, c, v, 1500, 1500, 64, thr, mc)


@pytest.mark.parametrize("bw", [1, 5, 64, 192])
def test_duplicates_unsorted_input_last_wins(gpu, bw):
    print(f"case: convert duplicates bw={bw}", flush=True)
    r, c, v = random_coo(700, 700, 6000, seed=bw, dup=0.3, band=12)
    A, B = check_same(r, c, v, 700, 700, bw, 2, 1_000_000_000)
    check_same(r, c, v, 700, 700, bw, 2, 25)  # the cap reached part way


@pytest.mark.parametrize("nr,nc", [(900, 500), (500, 900), (333, 333)])
def test_rectangular_and_ragged_blocks(gpu, nr, nc):
    print(f"case: convert rectangular {nr}x{nc}", flush=True)
    r, c, v = random_coo(nr, nc, 9000, seed=nr + nc, band=30)
    for bw in (64, 100):
        check_same(r, c, v, nr, nc, bw, 2, 1_000_000_000)


def test_cpu_coo_mode_and_single_entry(gpu):
    print("case: convert cpu-coo mode, tiny inputs", flush=True)
    r, c, v = random_coo(400, 400, 3000, seed=5, band=50)
    A, B = check_same(r, c, v, 400, 400, 64, 4, 1_000_000_000, gpu_coo=False)
    assert not B.info()["device_coo"]
    check_same(np.array([3], np.int32), np.array([2], np.int32), np.array([1.5]), 10, 10, 4, 1, 1_000_000_000)
    check_same(np.array([3], np.int32), np.array([2], np.int32), np.array([1.5]), 10, 10, 4, 2, 1_000_000_000)
    # every entry past the last block (1 column => 1 block of 64 rows)
    rr = np.arange(64, 200, dtype=np.int32)
    check_same(rr, np.zeros_like(rr), np.ones(len(rr)), 200, 1, 64, 1, 1_000_000_000)


def test_special_values_are_copied_bit_for_bit(gpu):
    print("case: convert special values", flush=True)
    r, c, v = random_coo(300, 300, 2000, seed=9, band=3)
    v[::7] = -0.0
    v[1::11] = np.inf
    v[2::13] = np.nan
    check_same(r, c, v, 300, 300, 64, 2, 1_000_000_000)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_after_device_conversion(gpu, dtype):
    print(f"case: spmv after device conversion {np.dtype(dtype).name}", flush=True)
    A = HBDIA.stencil(24, 27, 0.01, seed=4, dtype=dtype)
    r, c, v = A.getCOO()
    B = HBDIA(r, c, v, A.getNumRows(), A.getNumCols())
    B.convertToHBDIAFormatOnDevice(64, 2)
    assert B.info()["runs3_blocks"] > 0
    A.convertToHBDIAFormat(64, 2)
    x = gen_vector(A.getNumRows(), 11, dtype=dtype)
    ys = []
    for M in (A, B):
        X, Y = HBDIAVector(x), HBDIAVector(np.zeros_like(x))
        hbdiaSpMV(M, X, Y, False, True)
        M.synchronize()
        ys.append(Y.download())
    np.testing.assert_array_equal(ys[1], ys[0])
    hb = B.getHBDIA()
    hb["num_blocks"], hb["block_width"] = B.getNumBlocks(), 64
    np.testing.assert_array_equal(ys[1], oracle.hbdia_spmv(hb, x.copy(), np.zeros_like(x), A.getNumRows()))


def test_c_abi_global_rows_and_sorted_fallback(gpu):
    """hbdia_convert_coo with a global row map: offsets are col - global_rows[row].  Pinned by the oracle
    on columns shifted so that col' - row = col - global_rows[row] (blocks, placement and fallback order
    depend only on rows and offsets), with the fallback's columns shifted back."""
    print("case: convert C ABI global rows", flush=True)
    rng = np.random.default_rng(21)
    nr, nc = 600, 5000
    G = (np.arange(nr) * 3 + 1000).astype(np.int32)
    r = rng.integers(0, nr, 8000).astype(np.int32)
    c = np.clip(G[r] + rng.integers(-20, 21, len(r)), 0, nc - 1).astype(np.int32)
    v = rng.standard_normal(len(r))
    got = K.convert_coo(DeviceArray.from_host(r), DeviceArray.from_host(c), DeviceArray.from_host(v), nr, nc,
                        64, 20, 1_000_000_000, global_rows=DeviceArray.from_host(G))
    shifted = (c.astype(np.int64) - G[r] + r).astype(np.int32)
    want = oracle.convert(r, shifted, v, nr, nc, 64, 20)  # about 20 entries per pair: about half fall back
    np.testing.assert_array_equal(got["values"], want["values"])
    np.testing.assert_array_equal(got["offsets"][:got["num_pairs"]], want["offsets"])
    assert not got["offsets"][got["num_pairs"]:].any()
    np.testing.assert_array_equal(got["block_start"], want["block_start"])
    np.testing.assert_array_equal(got["coo_rows"], want["coo_rows"])
    np.testing.assert_array_equal(got["coo_cols"], want["coo_cols"] - want["coo_rows"] + G[want["coo_rows"]])
    np.testing.assert_array_equal(got["coo_values"], want["coo_values"])
    assert got["number_diagonals"] == want["num_diagonals"]
    # the input is unsorted, so is the fallback: the sorted copy is its stable sort by row
    assert got["num_coo"] > 1000 and not got["coo_sorted"]
    order = np.argsort(got["coo_rows"], kind="stable")
    np.testing.assert_array_equal(got["coo_rows_sorted"], got["coo_rows"][order])
    np.testing.assert_array_equal(got["coo_cols_sorted"], got["coo_cols"][order])
    np.testing.assert_array_equal(got["coo_values_sorted"], got["coo_values"][order])
    # per-block sizes, and placed entries = the block's entries minus its fallback entries
    bw = 64
    sizes = np.diff(want["block_start"])
    np.testing.assert_array_equal(got["block_sizes"][:-1], sizes)
    nb = len(sizes)
    placed = np.bincount(r // bw, minlength=nb + 2)[:nb] - np.bincount(want["coo_rows"] // bw, minlength=nb + 2)[:nb]
    np.testing.assert_array_equal(got["block_placed"], placed)
    bits = K.runs3_bits(want["offsets"], want["block_start"], bw)
    assert (got["runs3_blocks"] is None) == (bits is None)


def test_c_abi_runs3_bits_and_errors(gpu):
    print("case: convert C ABI runs3 bits", flush=True)
    r, c, v, _ = oracle.stencil27(20, 20, 20, 0.0, 0)
    got = K.convert_coo(DeviceArray.from_host(r), DeviceArray.from_host(c), DeviceArray.from_host(v), 8000, 8000,
                        64, 1)
    want = oracle.convert(r, c, v, 8000, 8000, 64, 1)
    np.testing.assert_array_equal(got["runs3_blocks"], K.runs3_bits(want["offsets"], want["block_start"], 64))
    assert got["runs3_count"] == int(sum(bin(int(w)).count("1") for w in got["runs3_blocks"]))
    assert got["coo_sorted"] and got["num_coo"] == 0
    none = K.convert_coo(DeviceArray.from_host(r), DeviceArray.from_host(c), DeviceArray.from_host(v), 8000, 8000,
                         64, 1, runs3=False)
    assert none["runs3_blocks"] is None and none["runs3_count"] == 0
    from hbdia_amd import HBDIAError
    with pytest.raises(HBDIAError, match="block_width"):
        K.convert_coo(DeviceArray.from_host(r), DeviceArray.from_host(c), DeviceArray.from_host(v), 8000, 8000, 0, 1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_device_only_conversion_defers_host_values(gpu, dtype):
    """hostValues=False: the device arrays are those of the full device conversion, the SpMV is bit-identical,
    size-only queries (info, storage report) leave the values on the device, and the first host access reads
    back exactly the host conversion's arrays."""
    print(f"case: device-only conversion 20^3 noise 0.05 {np.dtype(dtype).name}", flush=True)
    r, c, v, N = oracle.stencil27(20, 20, 20, 0.05, 3, dtype=dtype)
    A = HBDIA(r, c, v, N, N)
    A.convertToHBDIAFormatOnDevice(64, 2)
    B = HBDIA(r, c, v, N, N)
    B.convertToHBDIAFormatOnDevice(64, 2, hostValues=False)
    assert B.info()["host_values"] == 0 and A.info()["host_values"] == 1
    da, db = device_state(A), device_state(B)
    for k in da:
        np.testing.assert_array_equal(db[k], da[k], err_msg="device " + k)
    x = np.random.default_rng(5).uniform(-1, 1, N).astype(dtype)
    ya, yb = HBDIAVector(np.zeros(N, dtype)), HBDIAVector(np.zeros(N, dtype))
    X = HBDIAVector(x)
    hbdiaSpMV(A, X, ya, False, True)
    hbdiaSpMV(B, X, yb, False, True)
    A.synchronize()
    B.synchronize()
    assert yb.download().tobytes() == ya.download().tobytes()
    ra, rb = A.storage_report(), B.storage_report()
    assert ra["hbdia_data_size"] == rb["hbdia_data_size"] and ra["hbdia_bytes"] == rb["hbdia_bytes"]
    assert B.info()["num_values"] == A.info()["num_values"]
    assert B.info()["host_values"] == 0
    ha, hb = A.getHBDIA(), B.getHBDIA()  # the first host access reads the values back
    assert B.info()["host_values"] == 1
    for k in HOST_KEYS:
        np.testing.assert_array_equal(hb[k], ha[k], err_msg=k)
    assert hb["values"].tobytes() == ha["values"].tobytes()


def test_device_only_conversion_then_prepare_and_mixed(gpu):
    print("case: device-only conversion, prepareForGPU and FP32 storage", flush=True)
    from hbdia_amd import HBDIAMixed
    r, c, v, N = oracle.stencil27(16, 16, 16, 0.02, 1)
    A = HBDIA(r, c, v, N, N)
    A.convertToHBDIAFormatOnDevice(64, 2)
    B = HBDIA(r, c, v, N, N)
    B.convertToHBDIAFormatOnDevice(64, 2, hostValues=False)
    x = np.random.default_rng(6).uniform(-1, 1, N)
    X = HBDIAVector(x)
    ya, yb = HBDIAVector(np.zeros(N)), HBDIAVector(np.zeros(N))
    ma, mb = HBDIAMixed(A, "fp32"), HBDIAMixed(B, "fp32")
    assert B.info()["host_values"] == 0  # the reduced-precision copy is made on the device
    ma.spmv(X, ya)
    mb.spmv(X, yb)
    A.synchronize()
    B.synchronize()
    assert yb.download().tobytes() == ya.download().tobytes()
    ma.close()
    mb.close()
    B.prepareForGPU()  # re-uploads from the host values it reads back first
    assert B.info()["host_values"] == 1
    hbdiaSpMV(A, X, ya, False, True)
    hbdiaSpMV(B, X, yb, False, True)
    A.synchronize()
    B.synchronize()
    assert yb.download().tobytes() == ya.download().tobytes()
