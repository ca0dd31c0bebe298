"""The oracle: a CPU restatement of the reference's HBDIA SpMV path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package, and only as the checker
(or, in bench.py, as the timed CPU baseline).  Product code never imports it.

The restatement lives in oracle/hbdia_oracle.cpp (built by oracle/Makefile into
oracle/build/libhbdia_oracle.so); every function there cites the reference
file:line it follows.  This module is a thin numpy/ctypes front end.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_int, c_int64, c_void_p
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIB_PATH = _DIR / "build" / "libhbdia_oracle.so"


def _load():
    if not _LIB_PATH.exists():
        raise RuntimeError(f"{_LIB_PATH} missing; run `make -C oracle`")
    lib = ctypes.CDLL(str(_LIB_PATH), mode=os.RTLD_NOW)
    P = c_void_p
    sig = {
        "orc_stencil27_f64": (P, [c_int, c_int, c_int, c_double, c_int]),
        "orc_stencil27_f32": (P, [c_int, c_int, c_int, c_double, c_int]),
        "orc_coo_nnz_f64": (c_int64, [P]), "orc_coo_nnz_f32": (c_int64, [P]),
        "orc_coo_nrows_f64": (c_int, [P]), "orc_coo_nrows_f32": (c_int, [P]),
        "orc_coo_get_f64": (None, [P, P, P, P]), "orc_coo_get_f32": (None, [P, P, P, P]),
        "orc_coo_free_f64": (None, [P]), "orc_coo_free_f32": (None, [P]),
        "orc_convert_f64": (P, [P, P, P, c_int64, c_int, c_int, c_int, c_int, c_int]),
        "orc_convert_f32": (P, [P, P, P, c_int64, c_int, c_int, c_int, c_int, c_int]),
        "orc_hb_sizes_f64": (None, [P, P]), "orc_hb_sizes_f32": (None, [P, P]),
        "orc_hb_get_f64": (None, [P] * 10), "orc_hb_get_f32": (None, [P] * 10),
        "orc_hb_free_f64": (None, [P]), "orc_hb_free_f32": (None, [P]),
        "orc_bdia_spmv_f64": (None, [P, P, P, P, P, c_int64, c_int, c_int]),
        "orc_bdia_spmv_f32": (None, [P, P, P, P, P, c_int64, c_int, c_int]),
        "orc_cpu_coo_spmv_f64": (None, [P, P, P, c_int64, P, c_int, P]),
        "orc_cpu_coo_spmv_f32": (None, [P, P, P, c_int64, P, c_int, P]),
        "orc_hbdia_spmv_f64": (None, [P, P, P, c_int, c_int, P, P, P, c_int64, P, P, P, c_int64]),
        "orc_hbdia_spmv_f32": (None, [P, P, P, c_int, c_int, P, P, P, c_int64, P, P, P, c_int64]),
        "orc_coo_spmv_ld": (None, [P, P, P, c_int64, P, c_int64, P]),
        "orc_max_threads": (c_int, []),
        "orc_set_threads": (None, [c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = _load()


def _sfx(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _ptr(a: np.ndarray):
    return a.ctypes.data if a.size else None


def set_threads(n: int) -> None:
    _lib.orc_set_threads(int(n))


def max_threads() -> int:
    return int(_lib.orc_max_threads())


def stencil27(nx: int, ny: int, nz: int, noise: float = 0.0, iteration: int = 0, dtype=np.float64):
    """Reference create3DStencil27Point output: (rows, cols, values, numRows)."""
    s = _sfx(dtype)
    h = getattr(_lib, "orc_stencil27_" + s)(nx, ny, nz, float(noise), iteration)
    try:
        nnz = getattr(_lib, "orc_coo_nnz_" + s)(h)
        n = getattr(_lib, "orc_coo_nrows_" + s)(h)
        r = np.empty(nnz, np.int32)
        c = np.empty(nnz, np.int32)
        v = np.empty(nnz, np.dtype(dtype))
        getattr(_lib, "orc_coo_get_" + s)(h, r.ctypes.data, c.ctypes.data, v.ctypes.data)
    finally:
        getattr(_lib, "orc_coo_free_" + s)(h)
    return r, c, v, n


def convert(rows, cols, values, num_rows: int, num_cols: int, block_width: int, threshold: int,
            max_coo_entries: int = 1_000_000_000) -> dict:
    """Reference convertToHBDIAFormat + prepareForGPU flattening."""
    v = np.ascontiguousarray(values)
    s = _sfx(v.dtype)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    h = getattr(_lib, "orc_convert_" + s)(_ptr(r), _ptr(c), _ptr(v), len(v), int(num_rows), int(num_cols),
                                         int(block_width), int(threshold), int(max_coo_entries))
    try:
        sz = np.zeros(7, np.int64)
        getattr(_lib, "orc_hb_sizes_" + s)(h, sz.ctypes.data)
        nb, nval, noff, ncoo, ndiag, hl, rpl = (int(q) for q in sz)
        out = {
            "values": np.empty(nval, v.dtype), "offsets": np.empty(noff, np.int32),
            "block_start": np.empty(nb + 1, np.int32), "coo_rows": np.empty(ncoo, np.int32),
            "coo_cols": np.empty(ncoo, np.int32), "coo_values": np.empty(ncoo, v.dtype),
            "row_ptr": np.empty(rpl, np.int32), "hist_blocks": np.empty(hl, np.int32),
            "hist_nnz": np.empty(hl, np.int32),
        }
        getattr(_lib, "orc_hb_get_" + s)(h, *(_ptr(a) for a in out.values()))
        out["num_blocks"] = nb
        out["num_diagonals"] = ndiag
        out["block_width"] = int(block_width)
    finally:
        getattr(_lib, "orc_hb_free_" + s)(h)
    return out


def bdia_spmv(hb: dict, x: np.ndarray, y: np.ndarray, num_rows: int) -> np.ndarray:
    """Reference bdia_spmv_kernel on the host; y is updated in place and returned."""
    s = _sfx(hb["values"].dtype)
    getattr(_lib, "orc_bdia_spmv_" + s)(_ptr(hb["values"]), _ptr(hb["offsets"]), _ptr(hb["block_start"]),
                                        x.ctypes.data, y.ctypes.data, int(num_rows), int(hb["num_blocks"]),
                                        int(hb["block_width"]))
    return y


def cpu_coo_spmv(rows, cols, values, x: np.ndarray, num_rows: int, res: np.ndarray) -> np.ndarray:
    s = _sfx(np.asarray(values).dtype)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    v = np.ascontiguousarray(values)
    getattr(_lib, "orc_cpu_coo_spmv_" + s)(_ptr(r), _ptr(c), _ptr(v), len(v), x.ctypes.data, int(num_rows),
                                           res.ctypes.data)
    return res


def hbdia_spmv(hb: dict, x: np.ndarray, y: np.ndarray, num_rows: int, coo_buf: np.ndarray | None = None) -> np.ndarray:
    """Reference hbdiaSpMV (BDIA, COO into a persistent buffer, y += coo); y in/out."""
    s = _sfx(hb["values"].dtype)
    if coo_buf is None:
        coo_buf = np.zeros(num_rows, hb["values"].dtype)
    getattr(_lib, "orc_hbdia_spmv_" + s)(
        _ptr(hb["values"]), _ptr(hb["offsets"]), _ptr(hb["block_start"]), int(hb["num_blocks"]),
        int(hb["block_width"]), _ptr(hb["coo_rows"]), _ptr(hb["coo_cols"]), _ptr(hb["coo_values"]),
        len(hb["coo_values"]), x.ctypes.data, y.ctypes.data, coo_buf.ctypes.data, int(num_rows))
    return y


def coo_spmv_longdouble(rows, cols, values, x: np.ndarray, num_rows: int) -> np.ndarray:
    """Independent brute-force y = A x over the raw COO in long double."""
    r = np.ascontiguousarray(rows, dtype=np.int32)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    v = np.ascontiguousarray(values, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(num_rows, np.float64)
    _lib.orc_coo_spmv_ld(_ptr(r), _ptr(c), _ptr(v), len(v), xx.ctypes.data, int(num_rows), y.ctypes.data)
    return y


def storage_report(hb: dict, num_rows: int, num_cols: int, nnz: int, nnz_per_row, block_width: int,
                   threshold: int, value_bytes: int = 8) -> dict:
    """Restates printMatrixInfo's storage costs (benchmarking/benchmark.cpp:144-230) over an oracle
    conversion `hb` (convert()'s dict): HBDIA (:176-186), COO (:189-190), DIA (:193-194), SELL over
    slices of block_width rows with the per-row counts (:197-215; empty counts leave only the slice
    pointers), CSR (:218-220); Python integers, so nothing overflows."""
    hist_blocks = [int(v) for v in hb["hist_blocks"]]
    hist_nnz = [int(v) for v in hb["hist_nnz"]]
    ncoo = len(hb["coo_values"])
    offsets = sum(i * h for i, h in enumerate(hist_blocks))
    nb = int(hb["num_blocks"])
    nd = int(hb["num_diagonals"])
    hbdia = len(hb["values"]) * value_bytes + (nb + 1) * 4 + offsets * 4 + ncoo * 4 * 2 + ncoo * value_bytes
    coo = nnz * 4 * 2 + nnz * value_bytes
    dia = nd * min(num_rows, num_cols) * value_bytes + nd * 4
    sell, mx = 0, 0
    per = [int(v) for v in nnz_per_row]
    for i, k in enumerate(per):
        if i % block_width == 0 and i > 0:
            sell += mx * block_width * value_bytes + mx * block_width * 4
            mx = 0
        mx = max(mx, k)
    sell += mx * block_width * value_bytes + mx * block_width * 4
    sell += (num_rows + block_width - 1) // block_width * 4
    csr = nnz * value_bytes + nnz * 4 + (num_rows + 1) * 4
    hist = [(i, hist_blocks[i], hist_nnz[i] if i < len(hist_nnz) else 0) for i in range(len(hist_blocks))
            if hist_blocks[i] > 0 or (i < len(hist_nnz) and hist_nnz[i] > 0)]
    return {"hbdia_bytes": hbdia, "coo_bytes": coo, "dia_bytes": dia, "sell_bytes": sell, "csr_bytes": csr,
            "stored_offsets": offsets, "number_diagonals": nd, "num_blocks": nb, "coo_fallback_entries": ncoo,
            "hbdia_data_size": len(hb["values"]), "sell_applies": threshold == 1, "histogram": hist}
