"""Python mirror of the reference's HBDIA API, backed by libhbdia.so.

Names, argument meaning and error behaviour follow the reference's C++ API
(include/Format/HBDIA.hpp, include/Format/HBDIAVector.hpp,
include/Operations/HBDIASpMV.hpp), so tests and scripts read like the
reference's own drivers:

    A = HBDIA(); A.create3DStencil27Point(8, 8, 8, 0.0)
    A.convertToHBDIAFormat(32, 2)
    x = HBDIAVector(np.arange(1, 513, dtype=np.float64))
    y = HBDIAVector(np.zeros(512))
    hbdiaSpMV(A, x, y, execCOOCPU=True, execCOOGPU=False)
    A.synchronize(); y.download()
"""
from __future__ import annotations

import ctypes
import enum
from ctypes import byref, c_double, c_void_p

import numpy as np

from ._native import HBDIAError, capi, check_capi

DEFAULT_BLOCK_WIDTH = 32
DEFAULT_THRESHOLD = 1
MAX_COO_ENTRIES = 1_000_000_000
RAND_SEED = 42

F64, F32 = 0, 1
_INFO = ["rows", "cols", "nnz", "has_coo", "has_hbdia", "block_width", "threshold", "num_blocks", "num_values",
         "num_offsets", "num_coo", "num_diagonals", "hist_len", "on_device", "device_coo", "stored_bytes",
         "runs3_blocks", "host_values"]


_REPORT = ["rows", "cols", "nnz", "number_diagonals", "num_blocks", "block_width", "threshold", "max_coo_entries",
           "coo_fallback_entries", "hbdia_data_size", "stored_offsets", "value_bytes", "hbdia_bytes", "coo_bytes",
           "dia_bytes", "sell_bytes", "csr_bytes", "sell_applies", "hist_rows"]
# hbdia_matrix_local_info entries (HBDIA_LOCAL_* in include/hbdia_capi.h)
_LOCAL = ["first_row", "halo_left", "halo_right", "coo", "folded", "ranges"]


class ExecutionMode(enum.IntEnum):
    CPU_COO = 0
    GPU_COO = 1


def _code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise TypeError(f"unsupported dtype {dt} (float64 or float32)")


def _map_error(e: HBDIAError, what: str):
    msg = str(e)
    if "invalid_argument" in msg:
        return ValueError(msg)
    if "out_of_range" in msg:
        return IndexError(msg)
    return e


class HBDIA:
    """Hybrid blocked-diagonal matrix (reference HBDIA<T>)."""

    def __init__(self, rowIndices=None, colIndices=None, values=None, numRows: int | None = None,
                 numCols: int | None = None, dtype=np.float64, _handle=None):
        self.dtype = np.dtype(dtype if values is None else np.asarray(values).dtype)
        if _handle is not None:
            self._h = _handle
            self.dtype = np.dtype(dtype)
            return
        h = c_void_p(None)
        if rowIndices is None:
            check_capi(capi.hbdia_matrix_create(byref(h), _code(self.dtype)))
        else:
            r = np.ascontiguousarray(rowIndices, dtype=np.int32)
            c = np.ascontiguousarray(colIndices, dtype=np.int32)
            v = np.ascontiguousarray(values, dtype=self.dtype)
            if not (len(r) == len(c) == len(v)):
                raise ValueError("Row indices, column indices, and values vectors must have the same size")
            nr = -1 if numRows is None else int(numRows)
            nc = -1 if numCols is None else int(numCols)
            if (numRows is None) != (numCols is None):
                raise ValueError("give both numRows and numCols or neither")
            if numRows is not None and (nr <= 0 or nc <= 0):
                raise ValueError("Matrix dimensions must be positive")
            try:
                check_capi(capi.hbdia_matrix_create_coo(byref(h), _code(self.dtype), r.ctypes.data, c.ctypes.data,
                                                        v.ctypes.data, len(v), nr, nc))
            except HBDIAError as e:
                raise _map_error(e, "HBDIA") from None
        self._h = h.value

    # ---- generators -------------------------------------------------------
    @classmethod
    def stencil(cls, n: int, points: int = 27, scatter: float = 0.0, seed: int = 0, dtype=np.float64) -> "HBDIA":
        """Seeded stencil + scatter benchmark matrix (see include/Workloads.hpp)."""
        h = c_void_p(None)
        check_capi(capi.hbdia_gen_stencil(byref(h), _code(dtype), int(n), int(points), float(scatter), int(seed)))
        return cls(dtype=dtype, _handle=h.value)

    @classmethod
    def banded(cls, n: int, noffsets: int, max_offset: int = 64, offband: float = 0.01, seed: int = 0,
               dtype=np.float64) -> "HBDIA":
        h = c_void_p(None)
        check_capi(capi.hbdia_gen_banded(byref(h), _code(dtype), int(n), int(noffsets), int(max_offset),
                                         float(offband), int(seed)))
        return cls(dtype=dtype, _handle=h.value)

    @classmethod
    def banded_batch(cls, count: int = 2048, min_rows: int = 1000, max_rows: int = 100000, min_offsets: int = 3,
                     max_offsets: int = 9, max_offset: int = 64, offband: float = 0.01, seed: int = 0xC4B,
                     dtype=np.float64) -> list["HBDIA"]:
        """The seeded set of independent banded matrices (BandedBatchSpec in include/Workloads.hpp)."""
        hs = (c_void_p * int(count))()
        check_capi(capi.hbdia_gen_banded_batch(hs, _code(dtype), int(count), int(min_rows), int(max_rows),
                                               int(min_offsets), int(max_offsets), int(max_offset), float(offband),
                                               int(seed)))
        return [cls(dtype=dtype, _handle=h) for h in hs]

    @staticmethod
    def banded_batch_shape(i: int, min_rows: int = 1000, max_rows: int = 100000, min_offsets: int = 3,
                           max_offsets: int = 9, seed: int = 0xC4B) -> tuple[int, int, int]:
        """(rows, offsets, seed) of matrix i of banded_batch."""
        r, k, sd = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_uint64(0)
        check_capi(capi.hbdia_banded_batch_shape(int(i), int(min_rows), int(max_rows), int(min_offsets),
                                                 int(max_offsets), int(seed), byref(r), byref(k), byref(sd)))
        return int(r.value), int(k.value), int(sd.value)

    def create3DStencil27Point(self, nx: int, ny: int, nz: int, noise: float = 0.0, iteration: int = 0) -> None:
        check_capi(capi.hbdia_matrix_stencil27(self._h, nx, ny, nz, float(noise), iteration))

    # ---- conversion -------------------------------------------------------
    def convertToHBDIAFormat(self, blockWidth: int = DEFAULT_BLOCK_WIDTH, threshold: int = DEFAULT_THRESHOLD,
                             max_coo_entries: int = MAX_COO_ENTRIES, COOisUnique: bool = False,
                             execMode: ExecutionMode = ExecutionMode.GPU_COO) -> None:
        try:
            check_capi(capi.hbdia_matrix_convert(self._h, int(blockWidth), int(threshold), int(max_coo_entries),
                                                 int(bool(COOisUnique)), int(execMode == ExecutionMode.GPU_COO)))
        except HBDIAError as e:
            raise _map_error(e, "convert") from None

    def convertToHBDIAFormatOnHost(self, blockWidth: int = DEFAULT_BLOCK_WIDTH,
                                   threshold: int = DEFAULT_THRESHOLD, max_coo_entries: int = MAX_COO_ENTRIES,
                                   COOisUnique: bool = False,
                                   execMode: ExecutionMode = ExecutionMode.GPU_COO) -> None:
        """The conversion without the device upload (prepareForGPU uploads later)."""
        try:
            check_capi(capi.hbdia_matrix_convert_host(self._h, int(blockWidth), int(threshold),
                                                      int(max_coo_entries), int(bool(COOisUnique)),
                                                      int(execMode == ExecutionMode.GPU_COO)))
        except HBDIAError as e:
            raise _map_error(e, "convert") from None

    def convertToHBDIAFormatOnDevice(self, blockWidth: int = DEFAULT_BLOCK_WIDTH,
                                     threshold: int = DEFAULT_THRESHOLD, max_coo_entries: int = MAX_COO_ENTRIES,
                                     COOisUnique: bool = False,
                                     execMode: ExecutionMode = ExecutionMode.GPU_COO, hostValues: bool = True) -> None:
        """The conversion computed on the GPU: the host state of convertToHBDIAFormatOnHost, bit for bit,
        and the device arrays of prepareForGPU built in place.  hostValues=False leaves the blocked values
        on the device until a host access reads them back."""
        try:
            check_capi(capi.hbdia_matrix_convert_device_ex(self._h, int(blockWidth), int(threshold),
                                                           int(max_coo_entries), int(bool(COOisUnique)),
                                                           int(execMode == ExecutionMode.GPU_COO), int(bool(hostValues))))
        except HBDIAError as e:
            raise _map_error(e, "convert") from None

    def loadMTX(self, filename: str, numRows: int | None = None, numCols: int | None = None) -> bool:
        """Matrix Market input (reference HBDIA<T>::loadMTX): without dimensions the file's own, entries
        sorted by (row, col); with them, indices checked (IndexError) and file order kept."""
        ok = ctypes.c_int(0)
        try:
            check_capi(capi.hbdia_matrix_load_mtx(self._h, str(filename).encode(), int(numRows or 0),
                                                  int(numCols or 0), byref(ok)))
        except HBDIAError as e:
            raise _map_error(e, "loadMTX") from None
        return bool(ok.value)

    def deleteHBDIAFormat(self) -> None:
        """Drops the blocked format, host and device; the COO stays for another conversion."""
        check_capi(capi.hbdia_matrix_delete_hbdia(self._h))

    def removeCOODuplicates(self) -> None:
        check_capi(capi.hbdia_matrix_remove_duplicates(self._h))

    def prepareForGPU(self) -> None:
        check_capi(capi.hbdia_matrix_prepare_gpu(self._h))

    def synchronize(self) -> None:
        check_capi(capi.hbdia_matrix_synchronize(self._h))

    # ---- getters ----------------------------------------------------------
    def info(self) -> dict:
        arr = np.zeros(len(_INFO), dtype=np.int64)
        check_capi(capi.hbdia_matrix_info(self._h, arr.ctypes.data))
        return dict(zip(_INFO, (int(v) for v in arr)))

    def storage_report(self) -> dict:
        """Statistics and storage costs (reference printMatrixInfo, benchmarking/benchmark.cpp:144-230):
        byte counts as HBDIA / COO / DIA / SELL / CSR, their ratios to COO, and the per-block histogram
        rows (diagonals, blocks, entries)."""
        counts = np.zeros(len(_REPORT), np.int64)
        ratios = np.zeros(4, np.float64)
        check_capi(capi.hbdia_matrix_storage_report(self._h, counts.ctypes.data, ratios.ctypes.data, None, 0))
        rows = int(counts[_REPORT.index("hist_rows")])
        hist = np.zeros((max(rows, 1), 3), np.int64)
        check_capi(capi.hbdia_matrix_storage_report(self._h, counts.ctypes.data, ratios.ctypes.data,
                                                    hist.ctypes.data, rows))
        out = dict(zip(_REPORT, (int(v) for v in counts)))
        out["sell_applies"] = bool(out["sell_applies"])
        out.update(zip(("hbdia_vs_coo", "dia_vs_coo", "csr_vs_coo", "sell_vs_coo"), (float(r) for r in ratios)))
        out["histogram"] = [tuple(int(v) for v in row) for row in hist[:rows]]
        return out

    def storage_report_csv(self, name: str = "") -> str:
        """The storage report as the reference's matrix_info.csv and histogram.csv text."""
        n = ctypes.c_size_t(0)
        check_capi(capi.hbdia_matrix_storage_report_csv(self._h, name.encode(), None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        check_capi(capi.hbdia_matrix_storage_report_csv(self._h, name.encode(), buf, n.value + 1, ctypes.byref(n)))
        return buf.value.decode()

    def getNumRows(self) -> int: return self.info()["rows"]
    def getNumCols(self) -> int: return self.info()["cols"]
    def getNumNonZeros(self) -> int: return self.info()["nnz"]
    def getBlockWidth(self) -> int: return self.info()["block_width"]
    def getThreshold(self) -> int: return self.info()["threshold"]
    def getNumBlocks(self) -> int: return self.info()["num_blocks"]
    def getNumberDiagonals(self) -> int: return self.info()["num_diagonals"]
    def isHBDIAFormat(self) -> bool: return bool(self.info()["has_hbdia"])
    def isCOOFormat(self) -> bool: return bool(self.info()["has_coo"])

    def getCOO(self):
        nnz = self.info()["nnz"]
        r = np.empty(nnz, np.int32)
        c = np.empty(nnz, np.int32)
        v = np.empty(nnz, self.dtype)
        check_capi(capi.hbdia_matrix_get_coo(self._h, r.ctypes.data, c.ctypes.data, v.ctypes.data))
        return r, c, v

    def getRowIndices(self): return self.getCOO()[0]
    def getColIndices(self): return self.getCOO()[1]
    def getValues(self): return self.getCOO()[2]

    def getHBDIA(self) -> dict:
        """All blocked-DIA host arrays: values, offsets (flattened), block_start,
        coo_rows/cols/values, row_ptr, hist_blocks, hist_nnz."""
        i = self.info()
        out = {
            "values": np.empty(i["num_values"], self.dtype),
            "offsets": np.empty(i["num_offsets"], np.int32),
            "block_start": np.empty(i["num_blocks"] + 1, np.int32),
            "coo_rows": np.empty(i["num_coo"], np.int32),
            "coo_cols": np.empty(i["num_coo"], np.int32),
            "coo_values": np.empty(i["num_coo"], self.dtype),
            "row_ptr": np.empty(i["rows"] + 1 if i["has_hbdia"] else 0, np.int32),
            "hist_blocks": np.empty(i["hist_len"], np.int32),
            "hist_nnz": np.empty(i["hist_len"], np.int32),
        }
        check_capi(capi.hbdia_matrix_get_hbdia(self._h, *(a.ctypes.data for a in out.values())))
        return out

    # ---- rank-local (partial) matrices -----------------------------------
    @classmethod
    def createRankLocalMatrix(cls, globalMatrix: "HBDIA", rank: int, size: int, blockWidth: int = DEFAULT_BLOCK_WIDTH,
                              threshold: int = DEFAULT_THRESHOLD, max_coo_entries: int = MAX_COO_ENTRIES,
                              execMode: ExecutionMode = ExecutionMode.GPU_COO) -> "HBDIA":
        """Rows [rank*q, rank*q + q) of `globalMatrix` (q = rows // size; the last rank also takes the
        remainder) as the reference's distributor sets up each rank (src/DataExchange/BasicDistributor.cpp:57-79,
        partitionRowWise at BasicExtractor.cpp:125-234): partial matrix, host conversion with global-row
        offsets, data ranges, per-process request lists, vector offsets.  prepareForGPU() then builds the
        rank-local COO part and uploads."""
        h = c_void_p(None)
        try:
            check_capi(capi.hbdia_matrix_rank_local(byref(h), globalMatrix._h, int(rank), int(size), int(blockWidth),
                                                    int(threshold), int(max_coo_entries),
                                                    int(execMode == ExecutionMode.GPU_COO)))
        except HBDIAError as e:
            raise _map_error(e, "createRankLocalMatrix") from None
        return cls(dtype=globalMatrix.dtype, _handle=h.value)

    def buildLocalCOO(self) -> None:
        try:
            check_capi(capi.hbdia_matrix_build_local_coo(self._h))
        except HBDIAError as e:
            raise _map_error(e, "buildLocalCOO") from None

    def local_info(self) -> dict:
        arr = np.zeros(len(_LOCAL), np.int64)
        check_capi(capi.hbdia_matrix_local_info(self._h, arr.ctypes.data))
        return dict(zip(_LOCAL, (int(v) for v in arr)))

    def getLocalLayout(self) -> dict:
        """Halo columns (left then right, global ids), the rank-local COO part (columns relative to the
        local segment), the data ranges as a (k, 2) array, and local_info()'s counts."""
        i = self.local_info()
        out = {
            "halo_cols": np.empty(i["halo_left"] + i["halo_right"], np.int32),
            "coo_rows": np.empty(i["coo"], np.int32),
            "coo_cols": np.empty(i["coo"], np.int32),
            "coo_values": np.empty(i["coo"], self.dtype),
            "data_ranges": np.empty((i["ranges"], 2), np.int32),
        }
        check_capi(capi.hbdia_matrix_get_local(self._h, *(a.ctypes.data for a in out.values())))
        out.update(i)
        return out

    def getVectorOffsets(self) -> np.ndarray:
        out = np.empty(self.info()["num_offsets"], np.int32)
        check_capi(capi.hbdia_matrix_get_vector_offsets(self._h, out.ctypes.data))
        return out

    def getOffsetsPerBlock(self) -> list[np.ndarray]:
        hb = self.getHBDIA()
        bs = hb["block_start"]
        return [hb["offsets"][bs[b]:bs[b + 1]] for b in range(len(bs) - 1)]

    def device_ptrs(self) -> dict:
        arr = (ctypes.c_void_p * 10)()
        check_capi(capi.hbdia_matrix_device_ptrs(self._h, arr))
        keys = ["values", "offsets", "block_start", "block_sizes", "coo_rows", "coo_cols", "coo_values",
                "coo_row_ptr", "coo_group_ptr", "bdia_stream"]
        return dict(zip(keys, (arr[k] for k in range(10))))

    def close(self) -> None:
        if getattr(self, "_h", None):
            check_capi(capi.hbdia_matrix_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                capi.hbdia_matrix_destroy(self._h)
                self._h = None
        except Exception:
            pass


class HBDIAVector:
    """Device-resident vector (reference HBDIAVector<T>)."""

    def __init__(self, localData, unifiedMemory: bool = False, unifiedMemoryMalloc: bool = False,
                 unifiedMemoryManagedMalloc: bool = False, unifiedNumaAllocOnNode: bool = False, dtype=None):
        a = np.ascontiguousarray(localData, dtype=dtype or np.asarray(localData).dtype)
        self.dtype = a.dtype
        self.size = int(a.size)
        self.unified = bool(unifiedMemory)
        h = c_void_p(None)
        check_capi(capi.hbdia_vector_create(byref(h), _code(a.dtype), a.ctypes.data, a.size, int(unifiedMemory),
                                            int(unifiedMemoryMalloc), int(unifiedMemoryManagedMalloc),
                                            int(unifiedNumaAllocOnNode)))
        self._h = h.value

    @classmethod
    def withHalo(cls, localData, matrix: HBDIA, dtype=None) -> "HBDIAVector":
        """HBDIAVector(localData, matrix, rank, size) of the reference: [left halo | local | right halo],
        halo sizes from the rank-local matrix's request lists."""
        a = np.ascontiguousarray(localData, dtype=dtype or matrix.dtype)
        self = cls.__new__(cls)
        self.dtype = a.dtype
        self.size = int(a.size)
        self.unified = False
        h = c_void_p(None)
        check_capi(capi.hbdia_vector_create_halo(byref(h), _code(a.dtype), a.ctypes.data, a.size, matrix._h))
        self._h = h.value
        return self

    def fillHalo(self, matrix: HBDIA, globalVector) -> None:
        """Copies the halo entries from the global vector (the one-GPU stand-in for the halo exchange)."""
        g = np.ascontiguousarray(globalVector, dtype=self.dtype)
        check_capi(capi.hbdia_vector_fill_halo(self._h, matrix._h, g.ctypes.data, g.size))

    def verifyHalo(self, matrix: HBDIA) -> bool:
        """The reference's verifyDataExchange for a vector filled from x[g] = g."""
        ok = ctypes.c_int(0)
        check_capi(capi.hbdia_vector_verify_halo(self._h, matrix._h, ctypes.byref(ok)))
        return bool(ok.value)

    def getLocalSize(self) -> int:
        return self.size

    def upload(self, data) -> None:
        a = np.ascontiguousarray(data, dtype=self.dtype)
        check_capi(capi.hbdia_vector_upload(self._h, a.ctypes.data, a.size))

    def download(self) -> np.ndarray:
        out = np.empty(self.size, self.dtype)
        check_capi(capi.hbdia_vector_download(self._h, out.ctypes.data, self.size))
        return out

    def download_coo(self) -> np.ndarray:
        out = np.empty(self.size, self.dtype)
        check_capi(capi.hbdia_vector_download_coo(self._h, out.ctypes.data, self.size))
        return out

    def device_ptrs(self) -> tuple[int, int]:
        a, b = c_void_p(None), c_void_p(None)
        check_capi(capi.hbdia_vector_device_ptrs(self._h, byref(a), byref(b)))
        return a.value, b.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            check_capi(capi.hbdia_vector_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                capi.hbdia_vector_destroy(self._h)
                self._h = None
        except Exception:
            pass


def hbdiaSpMV(matrix: HBDIA, inputVector: HBDIAVector, outputVector: HBDIAVector, execCOOCPU: bool,
              execCOOGPU: bool) -> bool:
    """y = A x (reference hbdiaSpMV, src/Operations/HBDIASpMV.cpp:186-251).

    Asynchronous: call matrix.synchronize() before reading the output.  Raises
    RuntimeError("Invalid COO execution flags") when the matrix has fallback
    entries and both flags are false."""
    try:
        check_capi(capi.hbdia_spmv(matrix._h, inputVector._h, outputVector._h, int(bool(execCOOCPU)),
                                   int(bool(execCOOGPU))))
    except HBDIAError as e:
        if "Invalid COO execution flags" in str(e):
            raise RuntimeError("Invalid COO execution flags") from None
        raise _map_error(e, "hbdiaSpMV") from None
    return True


def spmv_time_synced(matrix: HBDIA, inputVector: HBDIAVector, outputVector: HBDIAVector, execCOOCPU: bool,
                     execCOOGPU: bool, iterations: int = 100) -> np.ndarray:
    """The reference's timing loop (benchmarking/benchHBDIA.cpp:30-50), run natively: per round, y
    cleared, device synchronized, then hbdiaSpMV + device synchronize timed on the host clock (ms)."""
    out = np.zeros(int(iterations), np.float64)
    check_capi(capi.hbdia_spmv_time_synced(matrix._h, inputVector._h, outputVector._h, int(bool(execCOOCPU)),
                                           int(bool(execCOOGPU)), int(iterations),
                                           out.ctypes.data_as(ctypes.POINTER(c_double))))
    return out


def multi_ld(n: int, dtype=np.float64) -> int:
    """A leading dimension for column-major multi-RHS X and Y of n rows (hbdiaSpMVMulti).

    n rounded up to whole 128-byte lines, plus one line when the column stride would be a multiple
    of 4 KiB: with a power-of-two stride the same rows of every column fall into the same L2 sets.
    128^3 rows, 32 columns: 0.520 ms at ld = n, 0.480 ms at ld = n + 16."""
    size = np.dtype(dtype).itemsize
    per = 128 // size
    ld = -(-int(n) // per) * per
    if (ld * size) % 4096 == 0:
        ld += per
    return ld


def hbdiaSpMVMulti(matrix: HBDIA, X, Y, nrhs: int, ldx: int | None = None, ldy: int | None = None,
                   execCOOGPU: bool = True, layout: str = "col") -> bool:
    """Y[:, j] = A X[:, j] for j < nrhs in one launch (X, Y: DeviceArray).

    layout "col": column-major, column j at X + j * ldx; "row": row-major (interleaved), row r's
    nrhs values adjacent (hbdiaSpMVMultiRowMajor; ldx / ldy unused).  Asynchronous like
    hbdiaSpMV; every column equals hbdiaSpMV on that column bit for bit."""
    n = matrix.getNumRows()
    if layout == "row":
        nx = max(n, matrix.getNumCols()) if matrix.info()["num_coo"] else n
        if X.size < nx * nrhs or Y.size < n * nrhs:
            raise ValueError("X or Y too small for nrhs row-major columns")
        try:
            check_capi(capi.hbdia_spmv_multi_rm(matrix._h, X.ptr, Y.ptr, int(nrhs), int(bool(execCOOGPU))))
        except HBDIAError as e:
            if "Invalid COO execution flags" in str(e):
                raise RuntimeError("Invalid COO execution flags") from None
            raise _map_error(e, "hbdiaSpMVMulti") from None
        return True
    if layout != "col":
        raise ValueError("layout must be 'col' or 'row'")
    # fallback columns reach numCols - 1 (blocked columns stay below numRows)
    nx = max(n, matrix.getNumCols()) if matrix.info()["num_coo"] else n
    ldx = nx if ldx is None else int(ldx)
    ldy = n if ldy is None else int(ldy)
    if X.size < ldx * (nrhs - 1) + nx or Y.size < ldy * (nrhs - 1) + n:
        raise ValueError("X or Y too small for nrhs columns")
    try:
        check_capi(capi.hbdia_spmv_multi(matrix._h, X.ptr, ldx, Y.ptr, ldy, int(nrhs), int(bool(execCOOGPU))))
    except HBDIAError as e:
        if "Invalid COO execution flags" in str(e):
            raise RuntimeError("Invalid COO execution flags") from None
        raise
    return True


class HBDIAMixed:
    """FP32 / BF16 value storage for an FP64 matrix (include/Operations/HBDIAMixed.hpp).

    spmv(x, y) computes y = A x in FP64 vectors with FP32 accumulation; rows whose
    cancellation ratio exceeds `cancellation_ratio` are recomputed in FP64.  Without a
    ratio, the one that guarantees DEFAULT_REL_TOLERANCE relative error per row is used;
    rel_error_bound states the guarantee of the ratio in use."""

    PRECISIONS = {"fp32": 1, "bf16": 2}
    DEFAULT_REL_TOLERANCE = 1e-2

    def __init__(self, matrix: HBDIA, precision: str = "bf16", cancellation_ratio: float | None = None):
        if matrix.dtype != np.float64:
            raise TypeError("reduced-precision storage needs an FP64 matrix")
        if cancellation_ratio is not None and not cancellation_ratio > 0:
            raise ValueError("cancellation ratio must be positive")
        self.matrix = matrix  # keeps the FP64 matrix alive
        self.precision = precision
        h = c_void_p(None)
        ratio = 0.0 if cancellation_ratio is None else float(cancellation_ratio)
        check_capi(capi.hbdia_mixed_create(byref(h), matrix._h, self.PRECISIONS[precision], ratio))
        self._h = h.value
        r, b = c_double(0.0), c_double(0.0)
        check_capi(capi.hbdia_mixed_error_bound(self._h, byref(r), byref(b)))
        self.cancellation_ratio = float(r.value)
        self.rel_error_bound = float(b.value)

    def spmv(self, x: HBDIAVector, y: HBDIAVector) -> None:
        check_capi(capi.hbdia_mixed_spmv(self._h, x._h, y._h))

    def spmv_raw(self, x_ptr, y_ptr) -> None:
        check_capi(capi.hbdia_mixed_spmv_raw(self._h, x_ptr, y_ptr))

    def fp64_rows(self) -> int:
        n = ctypes.c_int64(0)
        check_capi(capi.hbdia_mixed_fp64_rows(self._h, byref(n)))
        return int(n.value)

    def fp64_mask(self) -> np.ndarray:
        """Boolean per row: recomputed in FP64 by the last spmv."""
        n = self.matrix.getNumRows()
        words = np.zeros(-(-n // 64), np.uint64)
        check_capi(capi.hbdia_mixed_fp64_masks(self._h, words.ctypes.data, words.size))
        return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)

    def close(self) -> None:
        if getattr(self, "_h", None):
            check_capi(capi.hbdia_mixed_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                capi.hbdia_mixed_destroy(self._h)
                self._h = None
        except Exception:
            pass


class HBDIABatch:
    """Many independent matrices multiplied in one persistent launch
    (include/Operations/HBDIABatch.hpp).

    x and y are device pools (DeviceArray or raw pointers) of `rows` values;
    matrix m owns rows [row_offsets[m], row_offsets[m+1]).  Each matrix's part of
    y equals hbdiaSpMV(matrix, x_m, y_m, False, True) bit for bit."""

    INTERLEAVED, CONTIGUOUS = 0, 1
    _INFO = ["matrices", "rows", "tiles", "bdia_nnz", "coo_nnz", "algorithmic_bytes", "stored_bytes", "flops",
             "kind", "tiles_per_step", "workgroups", "dtype", "long_lists"]

    def __init__(self, matrices: list[HBDIA] | None = None, _handle=None, table=None):
        if _handle is None:
            if not matrices:
                raise ValueError("need at least one matrix")
            arr = (c_void_p * len(matrices))(*[m._h for m in matrices])
            h = c_void_p(None)
            check_capi(capi.hbdia_batch_create(byref(h), arr, len(matrices)))
            _handle = h.value
        self._h = _handle
        self.table = table or []
        info = self.info()
        self.dtype = np.dtype(np.float64 if info["dtype"] == F64 else np.float32)
        offs = np.zeros(info["matrices"] + 1, np.int64)
        check_capi(capi.hbdia_batch_row_offsets(self._h, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self.row_offsets = offs

    @classmethod
    def autotune(cls, matrices: list[HBDIA], block_widths=(64, 128, 256), threshold: int = 2,
                 reps: int = 10) -> "HBDIABatch":
        """Times every block width and tile schedule; the matrices are left converted at the chosen width.
        The timed candidates are in .table as dicts."""
        arr = (c_void_p * len(matrices))(*[m._h for m in matrices])
        bws = (ctypes.c_int * len(block_widths))(*[int(b) for b in block_widths])
        cap = 64 * len(block_widths)
        tab = (c_double * (5 * cap))()
        nrows = ctypes.c_int(0)
        h = c_void_p(None)
        check_capi(capi.hbdia_batch_autotune(byref(h), arr, len(matrices), bws, len(block_widths), int(threshold),
                                             int(reps), tab, cap, byref(nrows)))
        rows = [dict(block_width=int(tab[5 * i]), kind=int(tab[5 * i + 1]), tiles_per_step=int(tab[5 * i + 2]),
                     workgroups=int(tab[5 * i + 3]), ms=float(tab[5 * i + 4])) for i in range(min(cap, nrows.value))]
        return cls(_handle=h.value, table=rows)

    def info(self) -> dict:
        a = np.zeros(len(self._INFO), np.int64)
        check_capi(capi.hbdia_batch_info(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return dict(zip(self._INFO, (int(v) for v in a)))

    @property
    def rows(self) -> int:
        return int(self.row_offsets[-1])

    def set_schedule(self, kind: int = 0, tiles_per_step: int = 1, workgroups: int = 0) -> None:
        check_capi(capi.hbdia_batch_set_schedule(self._h, int(kind), int(tiles_per_step), int(workgroups)))

    def resident_workgroups(self, kind: int = 0) -> int:
        n = ctypes.c_int(0)
        check_capi(capi.hbdia_batch_resident_workgroups(self._h, int(kind), byref(n)))
        return int(n.value)

    @staticmethod
    def _ptr(a):
        return a.ptr if hasattr(a, "ptr") else a

    def spmv(self, x, y) -> None:
        check_capi(capi.hbdia_batch_spmv(self._h, self._ptr(x), self._ptr(y)))

    def spmv_per_matrix(self, x, y) -> None:
        check_capi(capi.hbdia_batch_spmv_per_matrix(self._h, self._ptr(x), self._ptr(y)))

    def capture_graph(self, x, y) -> None:
        check_capi(capi.hbdia_batch_capture_graph(self._h, self._ptr(x), self._ptr(y)))

    def launch_graph(self) -> None:
        check_capi(capi.hbdia_batch_launch_graph(self._h))

    def synchronize(self) -> None:
        check_capi(capi.hbdia_batch_synchronize(self._h))

    @property
    def stream_handle(self) -> int:
        s = c_void_p(None)
        check_capi(capi.hbdia_batch_stream(self._h, byref(s)))
        return s.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            check_capi(capi.hbdia_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                capi.hbdia_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass


class HBDIAServer:
    """Continuous-batching SpMV server over resident matrices (include/Operations/HBDIAServe.hpp).

    submit(m, x, y) queues y = A_m x (y starts at zero); x and y are host arrays (kept
    alive until the request completes) or, with device=True, DeviceArrays.  batching=False
    gives the plain form: one request at a time through the single-matrix kernel."""

    _INFO = ["matrices", "resident_bytes", "arena_bytes", "launches", "completed", "largest_launch", "dtype"]

    def __init__(self, matrices: list[HBDIA], hbm_budget: int = 0, max_jobs: int = 256, max_in_flight: int = 2,
                 workgroups: int = 0, tiles_per_step: int = 4, batching: bool = True):
        if not matrices:
            raise ValueError("need at least one matrix")
        arr = (c_void_p * len(matrices))(*[m._h for m in matrices])
        h = c_void_p(None)
        check_capi(capi.hbdia_server_create(byref(h), arr, len(matrices), int(hbm_budget), int(max_jobs),
                                            int(max_in_flight), int(workgroups), int(tiles_per_step),
                                            int(bool(batching))))
        self._h = h.value
        self.dtype = np.dtype(np.float64 if self.info()["dtype"] == F64 else np.float32)
        self._keep = {}

    def info(self) -> dict:
        a = np.zeros(len(self._INFO), np.int64)
        check_capi(capi.hbdia_server_info(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return dict(zip(self._INFO, (int(v) for v in a)))

    def num_rows(self, m: int) -> int:
        n = ctypes.c_int64(0)
        check_capi(capi.hbdia_server_num_rows(self._h, int(m), byref(n)))
        return int(n.value)

    def submit(self, m: int, x, y, device: bool = False) -> int:
        # The server copies num_rows(m) elements from x and into y asynchronously:
        # shorter arrays would be read and written past their end.
        n = self.num_rows(m)
        if x.dtype != self.dtype or y.dtype != self.dtype:
            raise TypeError(f"x and y must be {self.dtype} arrays")
        if x.size < n or y.size < n:
            raise ValueError(f"x and y need at least {n} elements for matrix {m} (got {x.size}, {y.size})")
        if device:
            xp, yp = x.ptr, y.ptr
        else:
            if not x.flags.c_contiguous or not y.flags.c_contiguous:
                raise TypeError(f"x and y must be contiguous {self.dtype} arrays")
            xp, yp = x.ctypes.data, y.ctypes.data
        rid = ctypes.c_int64(0)
        check_capi(capi.hbdia_server_submit(self._h, int(m), xp, yp, int(bool(device)), byref(rid)))
        self._keep[rid.value] = (x, y)
        return int(rid.value)

    def wait(self, rid: int) -> None:
        check_capi(capi.hbdia_server_wait(self._h, int(rid)))
        self._keep.pop(rid, None)

    def drain(self) -> None:
        check_capi(capi.hbdia_server_drain(self._h))
        self._keep.clear()

    def times(self, rid: int) -> tuple[int, int, int]:
        a = (ctypes.c_int64 * 3)()
        check_capi(capi.hbdia_server_times(self._h, int(rid), a))
        return int(a[0]), int(a[1]), int(a[2])

    def open_loop(self, rate: float, requests: int, seed: int, xs: list, variants: int, ys: list,
                  device: bool = False) -> tuple[np.ndarray, float]:
        """Poisson arrivals at `rate` per second (0: all at once); returns latencies (us) and wall seconds."""
        xp = (c_void_p * len(xs))(*[(a.ptr if device else a.ctypes.data) for a in xs])
        yp = (c_void_p * len(ys))(*[(a.ptr if device else a.ctypes.data) for a in ys])
        lat = np.zeros(int(requests), np.float64)
        wall = c_double(0.0)
        check_capi(capi.hbdia_server_open_loop(self._h, float(rate), int(requests), int(seed), xp, int(variants), yp,
                                               int(bool(device)),
                                               lat.ctypes.data_as(ctypes.POINTER(c_double)), byref(wall)))
        return lat, float(wall.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            check_capi(capi.hbdia_server_destroy(self._h))
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                capi.hbdia_server_destroy(self._h)
                self._h = None
        except Exception:
            pass


class VectorArena:
    """The server's first-fit allocator on its own (host only), for tests."""

    def __init__(self, capacity: int, align: int = 64):
        h = c_void_p(None)
        check_capi(capi.hbdia_arena_create(byref(h), int(capacity), int(align)))
        self._h = h.value

    def allocate(self, n: int) -> int:
        off = ctypes.c_int64(0)
        check_capi(capi.hbdia_arena_allocate(self._h, int(n), byref(off)))
        return int(off.value)

    def release(self, offset: int) -> None:
        check_capi(capi.hbdia_arena_release(self._h, int(offset)))

    def info(self) -> dict:
        a = (ctypes.c_int64 * 4)()
        check_capi(capi.hbdia_arena_info(self._h, a))
        return dict(capacity=int(a[0]), used=int(a[1]), largest_free=int(a[2]), free_blocks=int(a[3]))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                capi.hbdia_arena_destroy(self._h)
                self._h = None
        except Exception:
            pass


def gen_vector(n: int, seed: int, lo: float = -1.0, hi: float = 1.0, dtype=np.float64) -> np.ndarray:
    out = np.empty(int(n), np.dtype(dtype))
    check_capi(capi.hbdia_gen_vector(_code(dtype), int(n), int(seed), float(lo), float(hi), out.ctypes.data))
    return out


def set_num_threads(n: int) -> None:
    check_capi(capi.hbdia_set_num_threads(int(n)))
