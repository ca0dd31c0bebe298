"""Python callers of the device kernels exported by libhbdia_hip.so.

Each function launches one C-ABI entry point on DeviceArray operands; the
semantics are documented in include/hbdia_hip.h.  Launches are asynchronous
on `stream` (None = null stream).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._native import check_hip, hip
from .device import DeviceArray, Stream

ROW_GROUP = 64
OFFSETS_PAD = 32  # HBDIA_OFFSETS_PAD of include/hbdia_hip.h


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "f64"
    if dt == np.float32:
        return "f32"
    raise TypeError(f"unsupported value dtype {dt}")


def _h(stream):
    return stream.handle if stream is not None else None


def num_groups(num_rows: int) -> int:
    return (int(num_rows) + ROW_GROUP - 1) // ROW_GROUP


def bdia_spmv(values: DeviceArray, offsets: DeviceArray, block_start: DeviceArray, x: DeviceArray,
              y: DeviceArray, num_rows: int, num_blocks: int, block_width: int,
              stream: Stream | None = None) -> None:
    fn = getattr(hip, "hbdia_bdia_spmv_" + _sfx(values.dtype))
    check_hip(fn(values.ptr, offsets.ptr, block_start.ptr, x.ptr, y.ptr, int(num_rows), int(num_blocks),
                 int(block_width), _h(stream)))


def hbdia_spmv(values: DeviceArray, offsets: DeviceArray, block_start: DeviceArray, num_blocks: int,
               block_width: int, coo_rows: DeviceArray, coo_cols: DeviceArray, coo_values: DeviceArray,
               coo_group_ptr: DeviceArray, x: DeviceArray, y: DeviceArray, num_rows: int,
               stream: Stream | None = None) -> None:
    fn = getattr(hip, "hbdia_hbdia_spmv_" + _sfx(values.dtype))
    check_hip(fn(values.ptr, offsets.ptr, block_start.ptr, int(num_blocks), int(block_width), coo_rows.ptr,
                 coo_cols.ptr, coo_values.ptr, coo_group_ptr.ptr, x.ptr, y.ptr, int(num_rows), _h(stream)))


def runs3_bits(offsets: np.ndarray, block_start: np.ndarray, block_width: int) -> np.ndarray | None:
    """The runs3_blocks hint: bit b set when block b's offsets are runs of three consecutive values
    (a multiple of three offsets, each triple consecutive); None when no block qualifies or the block
    width is neither a multiple of 64 nor 32."""
    nb = len(block_start) - 1
    if (block_width % ROW_GROUP and block_width != ROW_GROUP // 2) or nb <= 0:
        return None
    bits = np.zeros((nb + 31) // 32, np.uint32)
    for b in range(nb):
        o = np.asarray(offsets[block_start[b]:block_start[b + 1]])
        if len(o) and len(o) % 3 == 0 and np.all(o[1::3] == o[0::3] + 1) and np.all(o[2::3] == o[1::3] + 1):
            bits[b >> 5] |= np.uint32(1 << (b & 31))
    return bits if bits.any() else None


def bdia_spmv_ex(values: DeviceArray, offsets: DeviceArray, block_start: DeviceArray, x: DeviceArray,
                 y: DeviceArray, num_rows: int, num_blocks: int, block_width: int, runs3: DeviceArray | None,
                 stream: Stream | None = None) -> None:
    fn = getattr(hip, "hbdia_bdia_spmv_ex_" + _sfx(values.dtype))
    check_hip(fn(values.ptr, offsets.ptr, block_start.ptr, x.ptr, y.ptr, int(num_rows), int(num_blocks),
                 int(block_width), runs3.ptr if runs3 is not None else None, _h(stream)))


def hbdia_spmv_ex(values: DeviceArray, offsets: DeviceArray, block_start: DeviceArray, num_blocks: int,
                  block_width: int, coo_rows: DeviceArray, coo_cols: DeviceArray, coo_values: DeviceArray,
                  coo_group_ptr: DeviceArray, x: DeviceArray, y: DeviceArray, num_rows: int,
                  runs3: DeviceArray | None, stream: Stream | None = None) -> None:
    fn = getattr(hip, "hbdia_hbdia_spmv_ex_" + _sfx(values.dtype))
    check_hip(fn(values.ptr, offsets.ptr, block_start.ptr, int(num_blocks), int(block_width), coo_rows.ptr,
                 coo_cols.ptr, coo_values.ptr, coo_group_ptr.ptr, x.ptr, y.ptr, int(num_rows),
                 runs3.ptr if runs3 is not None else None, _h(stream)))


def hbdia_spmm(values: DeviceArray, offsets: DeviceArray, block_start: DeviceArray, num_blocks: int,
               block_width: int, coo, X: DeviceArray, ldx: int, Y: DeviceArray, ldy: int, nrhs: int, num_rows: int,
               stream: Stream | None = None, runs3: DeviceArray | None = None) -> None:
    """Many right-hand sides; coo = (rows, cols, values, group_ptr) DeviceArrays or None; runs3 the
    runs-of-three block bitset (None: no hints)."""
    fn = getattr(hip, "hbdia_hbdia_spmm_ex_" + _sfx(values.dtype))
    cr, cc, cv, gp = (a.ptr for a in coo) if coo is not None else (None, None, None, None)
    check_hip(fn(values.ptr, offsets.ptr, block_start.ptr, int(num_blocks), int(block_width), cr, cc, cv, gp, X.ptr,
                 int(ldx), Y.ptr, int(ldy), int(nrhs), int(num_rows), runs3.ptr if runs3 is not None else None,
                 _h(stream)))


def hbdia_spmm_rm(values: DeviceArray, offsets: DeviceArray, block_start: DeviceArray, num_blocks: int,
                  block_width: int, coo, X: DeviceArray, Y: DeviceArray, nrhs: int, num_rows: int,
                  stream: Stream | None = None) -> None:
    """Many right-hand sides, X and Y row-major (row r's nrhs values adjacent); coo as for hbdia_spmm."""
    fn = getattr(hip, "hbdia_hbdia_spmm_rm_" + _sfx(values.dtype))
    cr, cc, cv, gp = (a.ptr for a in coo) if coo is not None else (None, None, None, None)
    check_hip(fn(values.ptr, offsets.ptr, block_start.ptr, int(num_blocks), int(block_width), cr, cc, cv, gp, X.ptr,
                 Y.ptr, int(nrhs), int(num_rows), _h(stream)))


def coo_spmv(coo_rows: DeviceArray, coo_cols: DeviceArray, coo_values: DeviceArray, coo_group_ptr: DeviceArray,
             x: DeviceArray, res: DeviceArray, num_rows: int, stream: Stream | None = None) -> None:
    fn = getattr(hip, "hbdia_coo_spmv_" + _sfx(coo_values.dtype))
    check_hip(fn(coo_rows.ptr, coo_cols.ptr, coo_values.ptr, coo_group_ptr.ptr, x.ptr, res.ptr, int(num_rows),
                 _h(stream)))


def coo_group_ptr(coo_rows: DeviceArray, nnz: int, num_rows: int, out: DeviceArray,
                  stream: Stream | None = None) -> None:
    if out.size < num_groups(num_rows) + 1:
        raise ValueError("group pointer buffer too small")
    check_hip(hip.hbdia_coo_group_ptr(coo_rows.ptr if coo_rows is not None else None, int(nnz), int(num_rows),
                                      out.ptr, _h(stream)))


def vector_add(dst: DeviceArray, src: DeviceArray, n: int, stream: Stream | None = None) -> None:
    fn = getattr(hip, "hbdia_vector_add_" + _sfx(dst.dtype))
    check_hip(fn(dst.ptr, src.ptr, int(n), _h(stream)))


class ConvertArgs(ctypes.Structure):
    """hbdia_convert_args_t of include/hbdia_hip.h."""
    _fields_ = [("rows", ctypes.c_void_p), ("cols", ctypes.c_void_p), ("values", ctypes.c_void_p),
                ("global_rows", ctypes.c_void_p), ("nnz", ctypes.c_int64), ("num_rows", ctypes.c_int32),
                ("num_cols", ctypes.c_int32), ("block_width", ctypes.c_int32), ("threshold", ctypes.c_int32),
                ("max_coo_entries", ctypes.c_int32), ("elem_size", ctypes.c_int32),
                ("want_row_sorted_coo", ctypes.c_int32), ("want_runs3", ctypes.c_int32)]


class ConvertResult(ctypes.Structure):
    """hbdia_convert_result_t of include/hbdia_hip.h."""
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "values", "offsets", "block_start", "block_sizes", "block_placed", "runs3_blocks", "coo_rows", "coo_cols",
        "coo_values", "coo_rows_sorted", "coo_cols_sorted", "coo_values_sorted")] + [
        ("num_pairs", ctypes.c_int64), ("num_coo", ctypes.c_int64), ("num_blocks", ctypes.c_int32),
        ("number_diagonals", ctypes.c_int32), ("runs3_count", ctypes.c_int32), ("coo_sorted", ctypes.c_int32)]


def _fetch(ptr, count: int, dtype) -> np.ndarray:
    out = np.empty(int(count), dtype)
    if out.nbytes:
        check_hip(hip.hbdia_memcpy(out.ctypes.data, ptr, out.nbytes))
    return out


def convert_coo(rows: DeviceArray, cols: DeviceArray, values: DeviceArray, num_rows: int, num_cols: int,
                block_width: int, threshold: int, max_coo_entries: int = 1_000_000_000,
                global_rows: DeviceArray | None = None, row_sorted_coo: bool = True, runs3: bool = True,
                stream: Stream | None = None) -> dict:
    """COO -> HBDIA on the device (hbdia_convert_coo); returns every result array downloaded to the host
    (None where the result holds NULL) and the counts, and releases the device arrays."""
    a = ConvertArgs(rows.ptr, cols.ptr, values.ptr, global_rows.ptr if global_rows is not None else None,
                    int(values.size), int(num_rows), int(num_cols), int(block_width), int(threshold),
                    int(max_coo_entries), values.dtype.itemsize, int(bool(row_sorted_coo)), int(bool(runs3)))
    r = ConvertResult()
    check_hip(hip.hbdia_convert_coo(ctypes.byref(a), ctypes.byref(r), _h(stream)))
    try:
        nb, pairs, ncoo, vt = r.num_blocks, r.num_pairs, r.num_coo, values.dtype
        get = lambda ptr, n, dt: _fetch(ptr, n, dt) if ptr else None  # noqa: E731
        return {
            "values": get(r.values, pairs * block_width, vt) if pairs else np.empty(0, vt),
            "offsets": get(r.offsets, pairs + OFFSETS_PAD, np.int32),
            "block_start": get(r.block_start, nb + 1, np.int32),
            "block_sizes": get(r.block_sizes, nb + 1, np.int32),
            "block_placed": get(r.block_placed, nb, np.int32) if nb else np.empty(0, np.int32),
            "runs3_blocks": get(r.runs3_blocks, (nb + 31) // 32, np.uint32),
            "coo_rows": get(r.coo_rows, ncoo, np.int32) if ncoo else np.empty(0, np.int32),
            "coo_cols": get(r.coo_cols, ncoo, np.int32) if ncoo else np.empty(0, np.int32),
            "coo_values": get(r.coo_values, ncoo, vt) if ncoo else np.empty(0, vt),
            "coo_rows_sorted": get(r.coo_rows_sorted, ncoo, np.int32),
            "coo_cols_sorted": get(r.coo_cols_sorted, ncoo, np.int32),
            "coo_values_sorted": get(r.coo_values_sorted, ncoo, vt),
            "num_pairs": pairs, "num_coo": ncoo, "num_blocks": nb, "number_diagonals": r.number_diagonals,
            "runs3_count": r.runs3_count, "coo_sorted": bool(r.coo_sorted),
        }
    finally:
        check_hip(hip.hbdia_convert_release(ctypes.byref(r)))


class DeviceLayout:
    """Host HBDIA arrays uploaded for the raw kernels (no C++ objects).

    values/offsets/block_start describe the blocked-DIA part; coo_* the
    row-sorted fallback.  The group pointer is built on the device.
    """

    def __init__(self, values: np.ndarray, offsets: np.ndarray, block_start: np.ndarray, num_rows: int,
                 block_width: int, coo_rows=None, coo_cols=None, coo_values=None, stream: Stream | None = None,
                 pad_offsets: bool = True):
        self.num_rows = int(num_rows)
        self.block_width = int(block_width)
        self.num_blocks = int(len(block_start) - 1)
        bits = runs3_bits(np.asarray(offsets), np.asarray(block_start, np.int32), self.block_width)
        self.runs3 = DeviceArray.from_host(bits) if bits is not None else None
        self.dtype = np.dtype(values.dtype)
        self.values = DeviceArray.from_host(values if values.size else np.zeros(1, self.dtype))
        # the multi-RHS and reduced-precision kernels read HBDIA_OFFSETS_PAD entries past the last
        # offset; pad_offsets=False uploads the array at its exact size (single-vector kernels only)
        offs = np.asarray(offsets, np.int32)
        if pad_offsets:
            offs = np.concatenate([offs, np.zeros(OFFSETS_PAD, np.int32)])
        self.offsets = DeviceArray.from_host(offs if offs.size else np.zeros(1, np.int32))
        self.block_start = DeviceArray.from_host(np.asarray(block_start, dtype=np.int32))
        self.coo_nnz = 0 if coo_rows is None else int(len(coo_rows))
        if self.coo_nnz:
            rows = np.asarray(coo_rows, dtype=np.int32)
            if np.any(np.diff(rows) < 0):
                raise ValueError("COO rows must be sorted")
            self.coo_rows = DeviceArray.from_host(rows)
            self.coo_cols = DeviceArray.from_host(np.asarray(coo_cols, dtype=np.int32))
            self.coo_values = DeviceArray.from_host(np.asarray(coo_values, dtype=self.dtype))
            self.group_ptr = DeviceArray(num_groups(self.num_rows) + 1, np.int32)
            coo_group_ptr(self.coo_rows, self.coo_nnz, self.num_rows, self.group_ptr, stream)
            if stream is not None:
                stream.synchronize()
            else:
                from .device import synchronize
                synchronize()

    def spmm(self, X: DeviceArray, Y: DeviceArray, nrhs: int, stream: Stream | None = None,
             hints: bool = False) -> None:
        coo = (self.coo_rows, self.coo_cols, self.coo_values, self.group_ptr) if self.coo_nnz else None
        hbdia_spmm(self.values, self.offsets, self.block_start, self.num_blocks, self.block_width, coo, X,
                   self.num_rows, Y, self.num_rows, nrhs, self.num_rows, stream, self.runs3 if hints else None)

    def spmm_rm(self, X: DeviceArray, Y: DeviceArray, nrhs: int, stream: Stream | None = None) -> None:
        coo = (self.coo_rows, self.coo_cols, self.coo_values, self.group_ptr) if self.coo_nnz else None
        hbdia_spmm_rm(self.values, self.offsets, self.block_start, self.num_blocks, self.block_width, coo, X, Y, nrhs,
                      self.num_rows, stream)

    def spmv(self, x: DeviceArray, y: DeviceArray, stream: Stream | None = None, hints: bool = False) -> None:
        """y = A x with the fused kernel (blocked-DIA only when there is no fallback); `hints` passes
        the runs-of-three block bitset (computed from the host offsets)."""
        runs3 = self.runs3 if hints else None
        if self.coo_nnz:
            hbdia_spmv_ex(self.values, self.offsets, self.block_start, self.num_blocks, self.block_width,
                          self.coo_rows, self.coo_cols, self.coo_values, self.group_ptr, x, y, self.num_rows, runs3,
                          stream)
        else:
            bdia_spmv_ex(self.values, self.offsets, self.block_start, x, y, self.num_rows, self.num_blocks,
                         self.block_width, runs3, stream)
