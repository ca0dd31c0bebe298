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


_BASELINE_PATH = _DIR / "build" / "libhbdia_cpu_baseline.so"


def _load(path: Path = _LIB_PATH):
    if not path.exists():
        raise RuntimeError(f"{path} missing; run `make -C oracle`")
    lib = ctypes.CDLL(str(path), mode=os.RTLD_NOW)
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
        "orc_convert_g_f64": (P, [P, P, P, c_int64, c_int, c_int, c_int, c_int, c_int, P]),
        "orc_convert_g_f32": (P, [P, P, P, c_int64, c_int, c_int, c_int, c_int, c_int, P]),
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


class CpuBaseline:
    """The restatement built with the reference's host flags (oracle/Makefile `baseline`: -O3
    -march=native -ffp-contract=fast -fopenmp), compiled on the machine that times it.  Only
    bench.py's cpu_baseline leg uses it, as the timed CPU path; the checker stays the oracle above."""

    FLAGS = "-O3 -march=native -ffp-contract=fast -fopenmp"

    def __init__(self):
        import shutil
        import subprocess
        make = shutil.which("make")
        proc = subprocess.run([make or "make", "-s", "-C", str(_DIR), "baseline"], stdout=subprocess.PIPE,
                              stderr=subprocess.STDOUT, text=True) if make else None
        if proc is None or proc.returncode != 0:
            raise RuntimeError("building the CPU baseline failed: " + (proc.stdout if proc else "make not found"))
        self._lib = _load(_BASELINE_PATH)

    def max_threads(self) -> int:
        return int(self._lib.orc_max_threads())

    def hbdia_spmv(self, hb: dict, x: np.ndarray, y: np.ndarray, num_rows: int, coo_buf: np.ndarray) -> np.ndarray:
        """hbdia_spmv below, through this build."""
        getattr(self._lib, "orc_hbdia_spmv_" + _sfx(hb["values"].dtype))(
            _ptr(hb["values"]), _ptr(hb["offsets"]), _ptr(hb["block_start"]), int(hb["num_blocks"]),
            int(hb["block_width"]), _ptr(hb["coo_rows"]), _ptr(hb["coo_cols"]), _ptr(hb["coo_values"]),
            len(hb["coo_values"]), x.ctypes.data, y.ctypes.data, coo_buf.ctypes.data, int(num_rows))
        return y


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
            max_coo_entries: int = 1_000_000_000, global_rows=None) -> dict:
    """Reference convertToHBDIAFormat + prepareForGPU flattening; `global_rows` is a partial
    matrix's global row mapping (offsets col - global row, blocks and lanes local)."""
    v = np.ascontiguousarray(values)
    s = _sfx(v.dtype)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    c = np.ascontiguousarray(cols, dtype=np.int32)
    if global_rows is None:
        h = getattr(_lib, "orc_convert_" + s)(_ptr(r), _ptr(c), _ptr(v), len(v), int(num_rows), int(num_cols),
                                             int(block_width), int(threshold), int(max_coo_entries))
    else:
        g = np.ascontiguousarray(global_rows, dtype=np.int32)
        if len(g) != num_rows:
            raise ValueError("global_rows needs one entry per local row")
        h = getattr(_lib, "orc_convert_g_" + s)(_ptr(r), _ptr(c), _ptr(v), len(v), int(num_rows), int(num_cols),
                                               int(block_width), int(threshold), int(max_coo_entries), _ptr(g))
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


# ---- one rank's matrix on one GPU ------------------------------------------------------------

def partition_row_wise(rows, cols, values, num_rows: int, size: int) -> list[dict]:
    """Restates BasicExtractor<T>::partitionRowWise (src/DataExchange/BasicExtractor.cpp:125-234):
    partition p owns rows [p*q, (p+1)*q - 1] (q = num_rows // size; the last one also the remainder,
    :136-165), found by the first range that holds the row (:196-205); entries keep their input order
    with local rows and global columns (:212-217)."""
    rows = np.asarray(rows, dtype=np.int64)
    q = num_rows // size
    rem = num_rows % size
    if q == 0:
        raise ValueError("fewer rows than partitions")
    starts = [p * q for p in range(size)]
    ends = [starts[p] + q + (rem if p == size - 1 else 0) - 1 for p in range(size)]  # inclusive
    owner = np.full(len(rows), -1, np.int64)
    for p in range(size - 1, -1, -1):
        owner[(rows >= starts[p]) & (rows <= ends[p])] = p
    out = []
    for p in range(size):
        m = owner == p
        out.append({"rows": (rows[m] - starts[p]).astype(np.int32), "cols": np.asarray(cols, np.int32)[m],
                    "values": np.asarray(values)[m], "first_row": starts[p], "num_rows": ends[p] - starts[p] + 1})
    return out


def data_ranges(hb: dict, cols, global_rows, num_rows: int, num_global_cols: int) -> list[tuple[int, int]]:
    """Restates HBDIA<T>::analyzeDataRanges (src/Format/HBDIA.cpp:1333-1412): every COO column
    (:1357-1360) and, for every stored (block, offset), the global column of each lane (:1366-1384),
    kept when inside [0, numGlobalCols); sorted and merged into [start, end) runs (:1392-1411).  Lanes
    past the last row are skipped: the reference reads past its row mapping there."""
    g = np.asarray(global_rows, np.int64)
    bw = int(hb["block_width"])
    bs = hb["block_start"].astype(np.int64)
    pair_block = np.repeat(np.arange(len(bs) - 1), np.diff(bs))
    lane_rows = pair_block[:, None] * bw + np.arange(bw)[None, :]
    ok = lane_rows < num_rows
    c = g[np.minimum(lane_rows, num_rows - 1)] + hb["offsets"].astype(np.int64)[:, None]
    ok &= (c >= 0) & (c < num_global_cols)
    u = np.unique(np.concatenate([np.asarray(cols, np.int64), c[ok]]))
    if len(u) == 0:
        return []
    brk = np.nonzero(np.diff(u) != 1)[0]
    first = np.concatenate([[0], brk + 1])
    last = np.concatenate([brk, [len(u) - 1]])
    return [(int(u[a]), int(u[b]) + 1) for a, b in zip(first, last)]


def process_data_ranges(ranges, num_global_rows: int, size: int) -> tuple[list[list[tuple[int, int]]], int]:
    """Restates BasicExtractor<T>::extractPartialMatrixMetadataRowWise (BasicExtractor.cpp:269-333):
    each range is cut at the ownership boundaries (p+1)*q, the last process taking the rest; returns
    the per-process lists and totalBufferSize."""
    q = num_global_rows // size
    pdr = [[] for _ in range(size)]
    total = 0
    for rs, re in ranges:
        total += re - rs
        p0 = max(0, min(rs // q, size - 1))
        p1 = max(0, min((re - 1) // q, size - 1))
        at = rs
        p = p0
        while p <= p1 and at < re:
            boundary = re if p == size - 1 else (p + 1) * q
            end = min(re, boundary)
            if at < end:
                pdr[p].append((at, end))
                at = end
            p += 1
    return pdr, total


def vector_offsets(hb: dict, pdr, rank: int, size: int, num_global_rows: int, num_rows: int) -> np.ndarray:
    """Restates HBDIA<T>::calculateVectorOffsets (src/Format/HBDIA.cpp:1034-1122) with
    findMemoryOffsetForGlobalIndex (:1124-1178): per (block, offset), where lane 0's x entry
    (rank*q + block*bw + offset) sits in [left | local | right]; INT_MIN when in no range."""
    left = sum(e - s for p in range(rank) for s, e in pdr[p])
    q = num_global_rows // size
    g_start = rank * q
    g_end = g_start + q + (num_global_rows % q if rank == size - 1 else 0)
    bw = int(hb["block_width"])
    bs = hb["block_start"]
    out = np.empty(len(hb["offsets"]), np.int64)

    def where(g):
        if g_start <= g < g_end:
            return left + (g - g_start)
        if g < 0:
            return g
        off = 0
        if g < g_start:
            for p in range(rank):
                for s, e in pdr[p]:
                    if s <= g < e:
                        return off + (g - s)
                    off += e - s
        else:
            off = left + num_rows
            for p in range(rank + 1, len(pdr)):
                for s, e in pdr[p]:
                    if s <= g < e:
                        return off + (g - s)
                    off += e - s
        return -(2 ** 31)

    for b in range(len(bs) - 1):
        for k in range(bs[b], bs[b + 1]):
            out[k] = where(g_start + b * bw + int(hb["offsets"][k]))
    return out.astype(np.int32)


def rank_local(rows, cols, values, num_rows: int, num_cols: int, rank: int, size: int, block_width: int,
               threshold: int, max_coo_entries: int = 1_000_000_000) -> dict:
    """One rank of a row-wise partition set up as the reference's distributor does
    (src/DataExchange/BasicDistributor.cpp:57-79) and laid out for one GPU: the rank's conversion with
    global-row offsets, its data ranges and request lists, the halo columns (requests to lower ranks,
    then to higher ranks) and the rank-local COO part.  The last has no reference counterpart (the
    reference's rank kernel would read halo entries through vector offsets): the fallback entries in
    their order, then every stored non-zero whose column is outside the local rows in
    (block, offset, lane) order, stably sorted by row, columns relative to the local segment."""
    part = partition_row_wise(rows, cols, values, num_rows, size)[rank]
    n = part["num_rows"]
    r0 = part["first_row"]
    grm = np.arange(r0, r0 + n, dtype=np.int32)
    hb = convert(part["rows"], part["cols"], part["values"], n, num_cols, block_width, threshold,
                 max_coo_entries, global_rows=grm)
    ranges = data_ranges(hb, part["cols"], grm, n, num_cols)
    pdr, total = process_data_ranges(ranges, num_rows, size)
    seg = lambda ps: np.concatenate([np.arange(s, e) for p in ps for s, e in pdr[p]] or [np.zeros(0, np.int64)])
    left = seg(range(rank)).astype(np.int64)
    right = seg(range(rank + 1, size)).astype(np.int64)
    L = len(left)

    def local_col(g):
        g = np.asarray(g, np.int64)
        out = np.empty(len(g), np.int64)
        loc = (g >= r0) & (g < r0 + n)
        out[loc] = g[loc] - r0
        lo = g < r0
        il = np.searchsorted(left, g[lo])
        if np.any(il >= len(left)) or np.any(left[np.minimum(il, len(left) - 1)] != g[lo]):
            raise ValueError("column in no halo range")
        out[lo] = il - L
        hi = g >= r0 + n
        ir = np.searchsorted(right, g[hi])
        if np.any(ir >= len(right)) or np.any(right[np.minimum(ir, len(right) - 1)] != g[hi]):
            raise ValueError("column in no halo range")
        out[hi] = n + ir
        return out

    bw = int(block_width)
    bs = hb["block_start"].astype(np.int64)
    pair_block = np.repeat(np.arange(len(bs) - 1), np.diff(bs))
    lane_rows = pair_block[:, None] * bw + np.arange(bw)[None, :]
    c_local = lane_rows + hb["offsets"].astype(np.int64)[:, None]
    v = hb["values"].reshape(-1, bw) if len(hb["values"]) else np.zeros((0, bw), hb["values"].dtype)
    fold = (lane_rows < n) & (v != 0) & ((c_local < 0) | (c_local >= n))
    fold &= (c_local + r0 >= 0) & (c_local + r0 < num_cols)
    f_rows = lane_rows[fold]
    f_cols = local_col(c_local[fold] + r0)
    f_vals = v[fold]
    rows_all = np.concatenate([hb["coo_rows"].astype(np.int64), f_rows])
    cols_all = np.concatenate([local_col(hb["coo_cols"]), f_cols])
    vals_all = np.concatenate([hb["coo_values"], f_vals])
    order = np.argsort(rows_all, kind="stable")
    return {"hb": hb, "first_row": r0, "num_rows": n, "ranges": ranges, "process_data_ranges": pdr,
            "total_buffer_size": total, "halo_cols": np.concatenate([left, right]).astype(np.int32),
            "halo_left": L, "halo_right": len(right), "coo_rows": rows_all[order].astype(np.int32),
            "coo_cols": cols_all[order].astype(np.int32), "coo_values": vals_all[order],
            "folded": int(fold.sum()), "part": part}


def rank_local_spmv(loc: dict, x_ext: np.ndarray, y: np.ndarray) -> np.ndarray:
    """y (local rows, in/out) for x_ext = [left halo | local | right halo]: the blocked part over the
    local segment with the kernel's bound check, the rank-local COO part into a zeroed buffer, then
    y += buffer when the part is non-empty -- the order of hbdia_spmv above."""
    n = loc["num_rows"]
    L = loc["halo_left"]
    x_local = np.ascontiguousarray(x_ext[L:L + n])
    bdia_spmv(loc["hb"], x_local, y, n)
    if len(loc["coo_values"]):
        buf = np.zeros(n, y.dtype)
        cpu_coo_spmv(loc["coo_rows"], loc["coo_cols"].astype(np.int64) + L, loc["coo_values"], x_ext, n, buf)
        y += buf
    return y
