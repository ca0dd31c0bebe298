/* hbdia_capi.h -- C handle API over the C++ HBDIA API (libhbdia.so), for
 * callers that cannot use C++ templates (Python ctypes, C, FFI).
 *
 * Every call returns an int status (0 = success) and never throws across the
 * boundary; hbdia_capi_last_error() describes the last failure on the calling
 * thread.  dtype is HBDIA_DTYPE_F64 (double) or HBDIA_DTYPE_F32 (float); value
 * buffers are of that type.
 *
 * Reference interfaces each call wraps (paths relative to the reference
 * checkout):
 *   hbdia_matrix_create / _create_coo  -> HBDIA<T> constructors, include/Format/HBDIA.hpp:31-39
 *   hbdia_matrix_stencil27             -> HBDIA<T>::create3DStencil27Point, HBDIA.hpp:45
 *   hbdia_matrix_convert[_host|_device] -> HBDIA<T>::convertToHBDIAFormat, HBDIA.hpp:52-54
 *   hbdia_matrix_remove_duplicates     -> HBDIA<T>::removeCOODuplicates, HBDIA.hpp:71
 *   hbdia_matrix_delete_hbdia          -> HBDIA<T>::deleteHBDIAFormat, HBDIA.hpp:67
 *   hbdia_matrix_load_mtx              -> HBDIA<T>::loadMTX, HBDIA.hpp:43-44
 *   hbdia_matrix_info / _get_*         -> the getters of HBDIA.hpp:74-110
 *   hbdia_matrix_storage_report        -> printMatrixInfo's figures, benchmarking/benchmark.cpp:144-230
 *   hbdia_vector_create                -> HBDIAVector<T>(localData[, four memory flags]),
 *                                         include/Format/HBDIAVector.hpp:31-32
 *   hbdia_spmv                         -> hbdiaSpMV<T>, include/Operations/HBDIASpMV.hpp:8-10
 *   hbdia_matrix_rank_local            -> one rank's matrix as the reference's distributor builds it:
 *                                         BasicExtractor<T>::partitionRowWise (src/DataExchange/
 *                                         BasicExtractor.cpp:125-234) and BasicDistributor.cpp:57-79
 *   hbdia_matrix_get_vector_offsets    -> HBDIA<T>::calculateVectorOffsets, src/Format/HBDIA.cpp:1034-1178
 *   hbdia_vector_create_halo           -> HBDIAVector<T>(localData, matrix, rank, size), HBDIAVector.hpp:33
 *   hbdia_vector_fill_halo             -> stands in for MPICommunicator<T>::exchangeData on one GPU
 *   hbdia_vector_verify_halo           -> MPICommunicator<T>::verifyDataExchange, MPICommunicator.cpp:415-498
 *   hbdia_matrix_build_local_coo, _local_info, _get_local
 *                                      -> no reference counterpart: the rank-local COO part
 *   hbdia_spmv_multi                   -> no reference counterpart: many right-hand sides in one launch
 *   hbdia_mixed_*                      -> no reference counterpart: FP32 / BF16 value storage
 *   hbdia_batch_*                      -> no reference counterpart: many matrices in one launch
 *   hbdia_server_*, hbdia_arena_*      -> no reference counterpart: continuous-batching request loop
 *   hbdia_gen_*                        -> no reference counterpart: seeded benchmark inputs
 */
#ifndef HBDIA_CAPI_H
#define HBDIA_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { HBDIA_DTYPE_F64 = 0, HBDIA_DTYPE_F32 = 1 };

/* indices into the info array of hbdia_matrix_info */
enum {
  HBDIA_INFO_ROWS = 0,
  HBDIA_INFO_COLS = 1,
  HBDIA_INFO_NNZ = 2,
  HBDIA_INFO_HAS_COO = 3,
  HBDIA_INFO_HAS_HBDIA = 4,
  HBDIA_INFO_BLOCK_WIDTH = 5,
  HBDIA_INFO_THRESHOLD = 6,
  HBDIA_INFO_NUM_BLOCKS = 7,
  HBDIA_INFO_NUM_VALUES = 8,  /* stored blocked-DIA values incl. padding */
  HBDIA_INFO_NUM_OFFSETS = 9, /* (block, offset) pairs */
  HBDIA_INFO_NUM_COO = 10,    /* fallback entries */
  HBDIA_INFO_NUM_DIAGONALS = 11,
  HBDIA_INFO_HIST_LEN = 12,
  HBDIA_INFO_ON_DEVICE = 13,
  HBDIA_INFO_DEVICE_COO = 14,
  HBDIA_INFO_STORED_BYTES = 15,
  HBDIA_INFO_RUNS3_BLOCKS = 16, /* blocks given the runs-of-three hint (0 before prepareForGPU) */
  HBDIA_INFO_HOST_VALUES = 17,  /* 0 between a device-only conversion and the first host access */
  HBDIA_INFO_COUNT = 18
};

typedef struct hbdia_matrix_opaque* hbdia_matrix_t;
typedef struct hbdia_vector_opaque* hbdia_vector_t;

const char* hbdia_capi_last_error(void);
int hbdia_set_num_threads(int n);
int hbdia_get_num_threads(void);

int hbdia_matrix_create(hbdia_matrix_t* m, int dtype);
/* num_rows <= 0 and num_cols <= 0 select the 3-argument constructor that
 * infers the shape from the largest indices. */
int hbdia_matrix_create_coo(hbdia_matrix_t* m, int dtype, const int32_t* rows, const int32_t* cols,
                            const void* values, int64_t nnz, int32_t num_rows, int32_t num_cols);
int hbdia_matrix_destroy(hbdia_matrix_t m);
int hbdia_matrix_stencil27(hbdia_matrix_t m, int nx, int ny, int nz, double noise, int iteration);
int hbdia_matrix_convert(hbdia_matrix_t m, int block_width, int threshold, int max_coo_entries,
                         int coo_is_unique, int gpu_coo);
/* convertToHBDIAFormatOnHost: the same conversion without the device upload
 * (no GPU needed); hbdia_matrix_prepare_gpu uploads afterwards. */
int hbdia_matrix_convert_host(hbdia_matrix_t m, int block_width, int threshold, int max_coo_entries,
                              int coo_is_unique, int gpu_coo);
/* convertToHBDIAFormatOnDevice: the same conversion computed on the GPU
 * (hbdia_convert_coo), leaving the host state of the host conversion and the
 * device arrays of hbdia_matrix_prepare_gpu. */
int hbdia_matrix_convert_device(hbdia_matrix_t m, int block_width, int threshold, int max_coo_entries,
                                int coo_is_unique, int gpu_coo);
/* The same with host_values = 0: the blocked values stay on the device until a host access
 * (hbdia_matrix_get_hbdia, prepare_gpu) reads them back -- the admission path of a matrix that
 * only runs on the GPU.  host_values = 1 is hbdia_matrix_convert_device. */
int hbdia_matrix_convert_device_ex(hbdia_matrix_t m, int block_width, int threshold, int max_coo_entries,
                                   int coo_is_unique, int gpu_coo, int host_values);
/* deleteHBDIAFormat: drops the blocked format (host and device); the COO stays. */
int hbdia_matrix_delete_hbdia(hbdia_matrix_t m);
int hbdia_matrix_remove_duplicates(hbdia_matrix_t m);
/* loadMTX: num_rows, num_cols <= 0 take the file's dimensions (and sort by
 * (row, col)); otherwise the given ones, with index checks.  *loaded = the
 * reference's return value (any entry read). */
int hbdia_matrix_load_mtx(hbdia_matrix_t m, const char* filename, int num_rows, int num_cols, int* loaded);
int hbdia_matrix_prepare_gpu(hbdia_matrix_t m);
int hbdia_matrix_synchronize(hbdia_matrix_t m);
int hbdia_matrix_info(hbdia_matrix_t m, int64_t* info /* [HBDIA_INFO_COUNT] */);

/* Storage report of a converted matrix (hbdiaStorageReport, include/Operations/HBDIAStats.hpp;
 * the figures of the reference's printMatrixInfo, benchmarking/benchmark.cpp:144-230). */
enum {
  HBDIA_REPORT_ROWS = 0,
  HBDIA_REPORT_COLS = 1,
  HBDIA_REPORT_NNZ = 2,
  HBDIA_REPORT_NUMBER_DIAGONALS = 3,
  HBDIA_REPORT_NUM_BLOCKS = 4,
  HBDIA_REPORT_BLOCK_WIDTH = 5,
  HBDIA_REPORT_THRESHOLD = 6,
  HBDIA_REPORT_MAX_COO_ENTRIES = 7,
  HBDIA_REPORT_COO_FALLBACK_ENTRIES = 8,
  HBDIA_REPORT_HBDIA_DATA_SIZE = 9,
  HBDIA_REPORT_STORED_OFFSETS = 10,
  HBDIA_REPORT_VALUE_BYTES = 11,
  HBDIA_REPORT_HBDIA_BYTES = 12,
  HBDIA_REPORT_COO_BYTES = 13,
  HBDIA_REPORT_DIA_BYTES = 14,
  HBDIA_REPORT_SELL_BYTES = 15,
  HBDIA_REPORT_CSR_BYTES = 16,
  HBDIA_REPORT_SELL_APPLIES = 17,
  HBDIA_REPORT_HIST_ROWS = 18,
  HBDIA_REPORT_COUNT = 19
};
/* counts: [HBDIA_REPORT_COUNT]; ratios: [4] = HBDIA, DIA, CSR, SELL bytes over COO bytes;
 * hist: up to hist_capacity rows of (diagonals, blocks, entries), counts[HBDIA_REPORT_HIST_ROWS]
 * rows in all (hist may be NULL with capacity 0). */
int hbdia_matrix_storage_report(hbdia_matrix_t m, int64_t* counts, double* ratios, int64_t* hist,
                                int64_t hist_capacity);
/* The report as CSV text (hbdiaStorageReportCSV): writes at most capacity bytes
 * (NUL-terminated when they fit) and the full length, without the NUL, to *length. */
int hbdia_matrix_storage_report_csv(hbdia_matrix_t m, const char* name, char* buf, size_t capacity,
                                    size_t* length);
int hbdia_matrix_get_coo(hbdia_matrix_t m, int32_t* rows, int32_t* cols, void* values);
/* Any output pointer may be NULL to skip it.  Sizes from hbdia_matrix_info:
 * values NUM_VALUES, offsets NUM_OFFSETS, block_start NUM_BLOCKS+1, coo_* NUM_COO,
 * row_ptr ROWS+1, hist_* HIST_LEN. */
int hbdia_matrix_get_hbdia(hbdia_matrix_t m, void* values, int32_t* offsets, int32_t* block_start,
                           int32_t* coo_rows, int32_t* coo_cols, void* coo_values, int32_t* row_ptr,
                           int32_t* hist_blocks, int32_t* hist_nnz);
/* Device arrays: [0] values [1] offsets [2] block_start [3] block_sizes
 * [4] coo_rows [5] coo_cols [6] coo_values [7] coo_row_ptr [8] coo_group_ptr
 * and the BDIA stream at [9]. */
int hbdia_matrix_device_ptrs(hbdia_matrix_t m, void** ptrs /* [10] */);

int hbdia_vector_create(hbdia_vector_t* v, int dtype, const void* data, int64_t n, int unified,
                        int unified_malloc, int unified_managed, int unified_numa);
int hbdia_vector_destroy(hbdia_vector_t v);
int hbdia_vector_upload(hbdia_vector_t v, const void* data, int64_t n);
int hbdia_vector_download(hbdia_vector_t v, void* out, int64_t n);
int hbdia_vector_download_coo(hbdia_vector_t v, void* out, int64_t n);
int hbdia_vector_device_ptrs(hbdia_vector_t v, void** local, void** coo);

int hbdia_spmv(hbdia_matrix_t m, hbdia_vector_t x, hbdia_vector_t y, int exec_coo_cpu, int exec_coo_gpu);
/* The reference's timing loop (benchmarking/benchHBDIA.cpp:30-50): each of
 * `iterations` iterations clears y and its COO-result buffer, synchronizes the
 * device, then times hbdiaSpMV plus a device synchronize on the host clock
 * into times_ms[i].  The reference's plots drop the first iteration as warm-up. */
int hbdia_spmv_time_synced(hbdia_matrix_t m, hbdia_vector_t x, hbdia_vector_t y, int exec_coo_cpu, int exec_coo_gpu,
                           int iterations, double* times_ms);

/* Rank-local form on one GPU (createRankLocalMatrix, include/DataExchange/BasicExtractor.hpp):
 * rows [rank*q, rank*q + q) of the global matrix (q = rows / size, the last rank also the
 * remainder), converted on the host with the global-row offsets, its data ranges and
 * per-process request lists computed.  hbdia_matrix_prepare_gpu then builds the rank-local
 * COO part (halo columns folded in) and uploads; hbdia_matrix_build_local_coo builds it
 * without the upload. */
int hbdia_matrix_rank_local(hbdia_matrix_t* local, hbdia_matrix_t global, int rank, int size, int block_width,
                            int threshold, int max_coo_entries, int gpu_coo);
int hbdia_matrix_build_local_coo(hbdia_matrix_t m);
enum {
  HBDIA_LOCAL_FIRST_ROW = 0,  /* global row of local row 0 */
  HBDIA_LOCAL_HALO_LEFT = 1,  /* entries of x before the local segment */
  HBDIA_LOCAL_HALO_RIGHT = 2, /* entries of x after it */
  HBDIA_LOCAL_COO = 3,        /* entries of the rank-local COO part */
  HBDIA_LOCAL_FOLDED = 4,     /* of which come from the blocked part */
  HBDIA_LOCAL_RANGES = 5,     /* data ranges (analyzeDataRanges) */
  HBDIA_LOCAL_COUNT = 6
};
int hbdia_matrix_local_info(hbdia_matrix_t m, int64_t* info /* [HBDIA_LOCAL_COUNT] */);
/* Any output may be NULL.  halo_cols: LEFT + RIGHT global columns (left halo first);
 * coo_*: LOCAL_COO entries, columns relative to the local segment; ranges: 2 * RANGES
 * (start, end) pairs. */
int hbdia_matrix_get_local(hbdia_matrix_t m, int32_t* halo_cols, int32_t* coo_rows, int32_t* coo_cols,
                           void* coo_values, int32_t* ranges);
/* calculateVectorOffsets' table, flattened like the offsets (NUM_OFFSETS entries). */
int hbdia_matrix_get_vector_offsets(hbdia_matrix_t m, int32_t* out);
/* HBDIAVector(localData, matrix, matrix rank, matrix size): halo segments sized from the
 * matrix's partial metadata. */
int hbdia_vector_create_halo(hbdia_vector_t* v, int dtype, const void* data, int64_t n, hbdia_matrix_t m);
/* Fills the halo segments from the global vector (n entries) on the host and copies them to
 * the device. */
int hbdia_vector_fill_halo(hbdia_vector_t v, hbdia_matrix_t m, const void* global, int64_t n);
/* verifyHaloExchange: *ok = 1 when the vector, filled from x[g] = g, holds every halo entry's
 * global index and every local entry's global row (MPICommunicator::verifyDataExchange,
 * src/DataExchange/MPICommunicator.cpp:415-498). */
int hbdia_vector_verify_halo(hbdia_vector_t v, hbdia_matrix_t m, int* ok);
/* hbdiaSpMVMulti: X, Y device arrays (column-major, leading dimensions ldx, ldy). */
int hbdia_spmv_multi(hbdia_matrix_t m, const void* X, int64_t ldx, void* Y, int64_t ldy, int nrhs, int exec_coo_gpu);
/* hbdiaSpMVMultiRowMajor: X, Y device arrays, row-major (row r's nrhs values adjacent). */
int hbdia_spmv_multi_rm(hbdia_matrix_t m, const void* X, void* Y, int nrhs, int exec_coo_gpu);

/* Reduced-precision storage (HBDIAMixed, include/Operations/HBDIAMixed.hpp):
 * precision 1 = FP32, 2 = BF16; an FP64 matrix only.  x, y: FP64 vectors.
 * cancellation_ratio <= 0 selects the threshold that guarantees
 * HBDIAMixed::kDefaultRelTolerance (1e-2) relative error per row. */
typedef struct hbdia_mixed_opaque* hbdia_mixed_t;
int hbdia_mixed_create(hbdia_mixed_t* h, hbdia_matrix_t m, int precision, double cancellation_ratio);
int hbdia_mixed_destroy(hbdia_mixed_t h);
int hbdia_mixed_spmv(hbdia_mixed_t h, hbdia_vector_t x, hbdia_vector_t y);
int hbdia_mixed_spmv_raw(hbdia_mixed_t h, const void* x, void* y);
int hbdia_mixed_fp64_rows(hbdia_mixed_t h, int64_t* rows);
/* The rows the last spmv recomputed in FP64: bit r of masks[g] is row 64 g + r;
 * words >= ceil(rows / 64). */
int hbdia_mixed_fp64_masks(hbdia_mixed_t h, uint64_t* masks, int64_t words);
/* The cancellation ratio in use and the relative error guaranteed for every
 * row kept in FP32 (HBDIAMixed::relErrorBound). */
int hbdia_mixed_error_bound(hbdia_mixed_t h, double* ratio, double* rel_error_bound);

/* Many matrices in one persistent launch (HBDIABatch and autotuneBatch,
 * include/Operations/HBDIABatch.hpp).  The matrices share one dtype, are
 * square and converted with block widths that are multiples of 64; the batch
 * copies their arrays, so they may be destroyed afterwards.  x and y are
 * device pools of info[HBDIA_BATCH_INFO_ROWS] values; matrix m's rows start
 * at row offset m.  Launches are asynchronous on the batch's stream. */
typedef struct hbdia_batch_opaque* hbdia_batch_handle_t;
enum {
  HBDIA_BATCH_INFO_MATRICES = 0,
  HBDIA_BATCH_INFO_ROWS = 1,
  HBDIA_BATCH_INFO_TILES = 2,
  HBDIA_BATCH_INFO_BDIA_NNZ = 3,
  HBDIA_BATCH_INFO_COO_NNZ = 4,
  HBDIA_BATCH_INFO_ALG_BYTES = 5,
  HBDIA_BATCH_INFO_STORED_BYTES = 6,
  HBDIA_BATCH_INFO_FLOPS = 7,
  HBDIA_BATCH_INFO_KIND = 8,
  HBDIA_BATCH_INFO_TILES_PER_STEP = 9,
  HBDIA_BATCH_INFO_WORKGROUPS = 10,
  HBDIA_BATCH_INFO_DTYPE = 11,
  HBDIA_BATCH_INFO_LONG_LISTS = 12, /* 1: the kernel that pairs tiles of long offset lists */
  HBDIA_BATCH_INFO_COUNT = 13
};
int hbdia_batch_create(hbdia_batch_handle_t* b, const hbdia_matrix_t* matrices, int count);
/* Tries every block width in block_widths (host conversion with `threshold`,
 * left at the chosen width) and every tile schedule, `reps` timed launches
 * each; returns the fastest batch.  table (may be NULL) receives up to
 * table_capacity rows of 5 doubles (block_width, kind, tiles_per_step,
 * workgroups, ms per launch); *table_rows the number of candidates timed. */
int hbdia_batch_autotune(hbdia_batch_handle_t* b, const hbdia_matrix_t* matrices, int count,
                         const int* block_widths, int num_widths, int threshold, int reps, double* table,
                         int table_capacity, int* table_rows);
int hbdia_batch_destroy(hbdia_batch_handle_t b);
int hbdia_batch_info(hbdia_batch_handle_t b, int64_t* info); /* HBDIA_BATCH_INFO_COUNT entries */
int hbdia_batch_row_offsets(hbdia_batch_handle_t b, int64_t* offsets); /* matrices + 1 entries */
int hbdia_batch_set_schedule(hbdia_batch_handle_t b, int kind, int tiles_per_step, int workgroups);
int hbdia_batch_resident_workgroups(hbdia_batch_handle_t b, int kind, int* workgroups);
int hbdia_batch_spmv(hbdia_batch_handle_t b, const void* x, void* y);
int hbdia_batch_spmv_per_matrix(hbdia_batch_handle_t b, const void* x, void* y);
int hbdia_batch_capture_graph(hbdia_batch_handle_t b, const void* x, void* y);
int hbdia_batch_launch_graph(hbdia_batch_handle_t b);
int hbdia_batch_synchronize(hbdia_batch_handle_t b);
int hbdia_batch_stream(hbdia_batch_handle_t b, void** stream); /* its hbdia_stream_t */

/* Continuous-batching server over resident matrices (HBDIAServer,
 * include/Operations/HBDIAServe.hpp).  Matrices: one dtype, square, converted
 * with block widths that are multiples of 64 (copied; may be destroyed after
 * creation).  hbm_budget_bytes covers the resident matrices plus the request
 * arena (0: resident bytes + 1 GiB); batching 0 selects the plain form (one
 * request at a time through the single-matrix kernel).  x, y of a request:
 * host memory, or device memory with device_vectors; valid until it completes.
 * y = A x with y starting at zero. */
typedef struct hbdia_server_opaque* hbdia_server_t;
enum {
  HBDIA_SERVER_INFO_MATRICES = 0,
  HBDIA_SERVER_INFO_RESIDENT_BYTES = 1,
  HBDIA_SERVER_INFO_ARENA_BYTES = 2,
  HBDIA_SERVER_INFO_LAUNCHES = 3,
  HBDIA_SERVER_INFO_COMPLETED = 4,
  HBDIA_SERVER_INFO_LARGEST_LAUNCH = 5,
  HBDIA_SERVER_INFO_DTYPE = 6,
  HBDIA_SERVER_INFO_COUNT = 7
};
int hbdia_server_create(hbdia_server_t* s, const hbdia_matrix_t* matrices, int count, int64_t hbm_budget_bytes,
                        int max_jobs_per_launch, int max_launches_in_flight, int workgroups, int tiles_per_step,
                        int batching);
int hbdia_server_destroy(hbdia_server_t s); /* completes every submitted request first */
int hbdia_server_submit(hbdia_server_t s, int matrix, const void* x, void* y, int device_vectors, int64_t* id);
int hbdia_server_wait(hbdia_server_t s, int64_t id);
int hbdia_server_drain(hbdia_server_t s);
int hbdia_server_times(hbdia_server_t s, int64_t id, int64_t* times); /* submit, admit, complete (ns) */
int hbdia_server_info(hbdia_server_t s, int64_t* info);             /* HBDIA_SERVER_INFO_COUNT entries */
int hbdia_server_num_rows(hbdia_server_t s, int matrix, int64_t* rows);
/* Open-loop load (runOpenLoop): xs holds matrices * variants vectors (matrix
 * m, variant v at m * variants + v), ys one per matrix; latencies_us receives
 * `requests` latencies. */
int hbdia_server_open_loop(hbdia_server_t s, double rate, int64_t requests, uint64_t seed, const void* const* xs,
                           int variants, void* const* ys, int device_vectors, double* latencies_us,
                           double* wall_seconds);
/* The server's first-fit arena allocator on its own (host only). */
typedef struct hbdia_arena_opaque* hbdia_arena_t;
int hbdia_arena_create(hbdia_arena_t* a, int64_t capacity, int64_t align);
int hbdia_arena_destroy(hbdia_arena_t a);
int hbdia_arena_allocate(hbdia_arena_t a, int64_t n, int64_t* offset); /* *offset = -1 when it does not fit */
int hbdia_arena_release(hbdia_arena_t a, int64_t offset);
int hbdia_arena_info(hbdia_arena_t a, int64_t* info); /* capacity, used, largest free block, free blocks */
/* Elements of each request arena (x and y) that a server with this HBM budget,
 * resident bytes and element size gets (host only): half the remainder per
 * vector, a multiple of 64, capped at 2^31 - 64 because the serve kernel
 * addresses rows as int32. */
int hbdia_serve_arena_elems(int64_t budget_bytes, int64_t resident_bytes, int64_t elem_bytes, int64_t* elems);

/* Seeded benchmark inputs (generators in Workloads.hpp). */
int hbdia_gen_stencil(hbdia_matrix_t* m, int dtype, int n, int points, double scatter, uint64_t seed);
int hbdia_gen_banded(hbdia_matrix_t* m, int dtype, int n, int noffsets, int max_offset, double offband,
                     uint64_t seed);
int hbdia_gen_vector(int dtype, int64_t n, uint64_t seed, double lo, double hi, void* out);
/* The banded matrix set of BandedBatchSpec: out receives `count` new handles;
 * matrix i is hbdia_gen_banded(rows_i, noffsets_i, max_offset, offband,
 * seed_i) with the shape hbdia_banded_batch_shape reports. */
int hbdia_gen_banded_batch(hbdia_matrix_t* out, int dtype, int count, int min_rows, int max_rows,
                           int min_offsets, int max_offsets, int max_offset, double offband, uint64_t seed);
int hbdia_banded_batch_shape(int i, int min_rows, int max_rows, int min_offsets, int max_offsets, uint64_t seed,
                             int32_t* rows, int32_t* noffsets, uint64_t* matrix_seed);

#ifdef __cplusplus
}
#endif

#endif /* HBDIA_CAPI_H */
