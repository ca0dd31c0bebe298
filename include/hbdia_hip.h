/* hbdia_hip.h -- C ABI of libhbdia_hip.so, the device side of the HBDIA SpMV
 * path on AMD Instinct MI355X (gfx950).
 *
 * Every entry point returns an int status (HBDIA_SUCCESS == 0) and never exits
 * or throws; hbdia_last_error() returns a description of the last failure on
 * the calling thread.  Plain pointers and sizes only.  Device pointers come
 * from hbdia_malloc (or any HIP device allocation).  A NULL stream means the
 * default (null) stream.
 *
 * Reference interfaces each group replaces (paths relative to the reference
 * checkout):
 *   memory / copies  -> cudaMalloc / cudaMemcpy / cudaMemset / cudaMallocHost
 *                       calls in src/Format/HBDIA.cpp:768-814 and
 *                       src/Format/HBDIAVector.cpp:84-153
 *   streams / events -> HBDIA<T>::initializeStreams, src/Format/HBDIA.cpp:858-876,
 *                       and the event waits in src/Operations/HBDIASpMV.cpp:221,236-239
 *   BDIA kernel      -> launch_bdia_spmv_kernel<T>, src/Operations/HBDIASpMV.cu:91-117
 *   vector add       -> launch_vector_add_kernel<T>, src/Operations/HBDIASpMV.cu:120-135
 *   COO fallback     -> executeCOOOnGPU<T> (cusparseSpMV), src/Operations/HBDIASpMV.cpp:133-183
 */
#ifndef HBDIA_HIP_H
#define HBDIA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Rows handled by one 64-lane wavefront; COO group pointers are per group. */
#define HBDIA_ROW_GROUP 64
/* The single-vector entry points hbdia_bdia_spmv_* and hbdia_hbdia_spmv_*,
 * which replace the reference's launch_bdia_spmv_kernel, read nothing past
 * block_start[num_blocks] in `offsets`: an exactly-sized array is enough (so
 * is the _ex form with runs3_blocks == NULL).  Every other blocked entry point
 * -- the _ex forms given a runs-of-three bitset, multi-RHS (hbdia_hbdia_spmm_*)
 * and reduced precision (hbdia_hbdia_spmv_mixed*) -- reads a chunk's offsets
 * as one wide load and needs the array readable for HBDIA_OFFSETS_PAD int32
 * past its last offset; what lies there is ignored.  HBDIA<T>::prepareForGPU
 * and hbdia_convert_coo allocate this padding. */
#define HBDIA_OFFSETS_PAD 32

typedef struct hbdia_stream_opaque* hbdia_stream_t;
typedef struct hbdia_event_opaque* hbdia_event_t;
typedef struct hbdia_graph_opaque* hbdia_graph_t;

enum {
  HBDIA_SUCCESS = 0,
  HBDIA_ERROR_INVALID_ARGUMENT = 1,
  HBDIA_ERROR_DEVICE = 2,
  HBDIA_ERROR_OUT_OF_MEMORY = 3,
  HBDIA_ERROR_NOT_SUPPORTED = 4
};

/* ---- errors, devices ---------------------------------------------------- */
const char* hbdia_last_error(void);
int hbdia_get_device_count(int* count);
int hbdia_set_device(int device);
int hbdia_get_device(int* device);
int hbdia_device_synchronize(void);
int hbdia_mem_info(size_t* free_bytes, size_t* total_bytes);
int hbdia_device_name(char* buf, size_t len);
int hbdia_device_compute_units(int* count); /* compute units of the current device */

/* ---- memory (HBDIA.cpp:768-814, HBDIAVector.cpp:84-153) ------------------ */
int hbdia_malloc(void** ptr, size_t bytes);
int hbdia_free(void* ptr);
int hbdia_malloc_host(void** ptr, size_t bytes);    /* pinned host memory */
int hbdia_free_host(void* ptr);
int hbdia_malloc_managed(void** ptr, size_t bytes); /* hipMallocManaged */
int hbdia_memcpy_h2d(void* dst, const void* src, size_t bytes, hbdia_stream_t stream);
int hbdia_memcpy_d2h(void* dst, const void* src, size_t bytes, hbdia_stream_t stream);
int hbdia_memcpy_d2d(void* dst, const void* src, size_t bytes, hbdia_stream_t stream);
int hbdia_memcpy(void* dst, const void* src, size_t bytes); /* synchronous, any direction */
int hbdia_memset(void* dst, int value, size_t bytes, hbdia_stream_t stream);

/* ---- streams, events, graphs (HBDIA.cpp:858-876, HBDIASpMV.cpp:221,236-239) */
int hbdia_stream_create(hbdia_stream_t* stream); /* non-blocking, default priority */
int hbdia_stream_destroy(hbdia_stream_t stream);
int hbdia_stream_synchronize(hbdia_stream_t stream);
int hbdia_event_create(hbdia_event_t* event);
int hbdia_event_destroy(hbdia_event_t event);
int hbdia_event_record(hbdia_event_t event, hbdia_stream_t stream);
int hbdia_event_synchronize(hbdia_event_t event);
int hbdia_event_query(hbdia_event_t event, int* done); /* *done = 1 once the event has completed */
int hbdia_stream_wait_event(hbdia_stream_t stream, hbdia_event_t event);
int hbdia_event_elapsed_ms(float* ms, hbdia_event_t start, hbdia_event_t stop);
int hbdia_graph_begin_capture(hbdia_stream_t stream);
int hbdia_graph_end_capture(hbdia_stream_t stream, hbdia_graph_t* graph);
int hbdia_graph_launch(hbdia_graph_t graph, hbdia_stream_t stream);
int hbdia_graph_destroy(hbdia_graph_t graph);

/* ---- kernels ------------------------------------------------------------- */
/* Blocked-DIA part only; replaces launch_bdia_spmv_kernel<T>
 * (HBDIASpMV.cu:91-117) with the semantics of bdia_spmv_kernel (:7-74):
 *   for each row r < num_rows with block b = r / block_width < num_blocks and a
 *   non-empty offset list, y[r] = sum over the block's offsets in ascending
 *   order of values[(block_start[b] + o) * block_width + r % block_width] *
 *   x[r + offsets[block_start[b] + o]], skipping zero values and columns
 *   outside [0, num_rows), accumulated as FMAs.  Other rows keep y.
 * values: sum_b (block_start[b+1]-block_start[b]) * block_width elements.
 * block_start: num_blocks + 1 prefix offsets into `offsets`. */
int hbdia_bdia_spmv_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                        const double* x, double* y, int64_t num_rows, int32_t num_blocks,
                        int32_t block_width, hbdia_stream_t stream);
int hbdia_bdia_spmv_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                        const float* x, float* y, int64_t num_rows, int32_t num_blocks,
                        int32_t block_width, hbdia_stream_t stream);

/* Blocked-DIA plus the COO fallback in one launch; replaces the BDIA launch,
 * executeCOOOnGPU and the vector add of hbdiaSpMV (HBDIASpMV.cpp:186-251) when
 * the fallback is non-empty.  Result: y[r] = bdia(r) + coo(r) for every row,
 * where bdia(r) is as above (or the old y[r] for rows the BDIA part does not
 * write) and coo(r) is the sequential FMA sum of the row's COO entries in entry
 * order (+0 for rows without entries).  coo_rows must be sorted ascending;
 * coo_group_ptr (ceil(num_rows/64)+1 entries) comes from hbdia_coo_group_ptr.
 * x must hold every column the COO entries reference. */
int hbdia_hbdia_spmv_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                         int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                         const int32_t* coo_cols, const double* coo_values,
                         const int32_t* coo_group_ptr, const double* x, double* y,
                         int64_t num_rows, hbdia_stream_t stream);
int hbdia_hbdia_spmv_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                         int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                         const int32_t* coo_cols, const float* coo_values,
                         const int32_t* coo_group_ptr, const float* x, float* y,
                         int64_t num_rows, hbdia_stream_t stream);

/* The same two products with a layout hint: runs3_blocks is a device bitset
 * (bit b & 31 of word b >> 5) whose bit b says that block b's offsets are runs
 * of three consecutive values (o-1, o, o+1) -- a multiple of three offsets,
 * each triple consecutive -- as a 27-point stencil's are.  Such a block loads
 * each run's x once and shifts it across lanes for the run's neighbours.  A
 * set bit must be true of its block; clear bits are always safe; NULL means no
 * hints.  Used when block_width % 64 == 0 or block_width == 32.  Results are bit-identical to the
 * entry points above. */
int hbdia_bdia_spmv_ex_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                           const double* x, double* y, int64_t num_rows, int32_t num_blocks, int32_t block_width,
                           const uint32_t* runs3_blocks, hbdia_stream_t stream);
int hbdia_bdia_spmv_ex_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                           const float* x, float* y, int64_t num_rows, int32_t num_blocks, int32_t block_width,
                           const uint32_t* runs3_blocks, hbdia_stream_t stream);
int hbdia_hbdia_spmv_ex_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const double* coo_values, const int32_t* coo_group_ptr,
                            const double* x, double* y, int64_t num_rows, const uint32_t* runs3_blocks,
                            hbdia_stream_t stream);
int hbdia_hbdia_spmv_ex_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const float* coo_values, const int32_t* coo_group_ptr,
                            const float* x, float* y, int64_t num_rows, const uint32_t* runs3_blocks,
                            hbdia_stream_t stream);

/* Multiple right-hand sides (no reference counterpart: the reference's
 * hbdiaSpMV takes one vector; this is the batched form of the same product).
 * For j < nrhs: Y[:, j] = A X[:, j], column j at X + j*ldx and Y + j*ldy
 * (ldx, ldy >= num_rows).  With coo_group_ptr == NULL the result follows
 * hbdia_bdia_spmv_* (no fallback); otherwise hbdia_hbdia_spmv_*.  Every column
 * equals the single-vector result bit for bit.  block_width % 64 == 0. */
int hbdia_hbdia_spmm_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                         int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                         const int32_t* coo_cols, const double* coo_values, const int32_t* coo_group_ptr,
                         const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t nrhs,
                         int64_t num_rows, hbdia_stream_t stream);
int hbdia_hbdia_spmm_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                         int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                         const int32_t* coo_cols, const float* coo_values, const int32_t* coo_group_ptr,
                         const float* X, int64_t ldx, float* Y, int64_t ldy, int32_t nrhs,
                         int64_t num_rows, hbdia_stream_t stream);
/* The same with the runs-of-three block bitset of hbdia_hbdia_spmv_ex_*
 * (NULL: no hints); results are bit-identical. */
int hbdia_hbdia_spmm_ex_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const double* coo_values, const int32_t* coo_group_ptr,
                            const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t nrhs, int64_t num_rows,
                            const uint32_t* runs3_blocks, hbdia_stream_t stream);
int hbdia_hbdia_spmm_ex_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const float* coo_values, const int32_t* coo_group_ptr,
                            const float* X, int64_t ldx, float* Y, int64_t ldy, int32_t nrhs, int64_t num_rows,
                            const uint32_t* runs3_blocks, hbdia_stream_t stream);

/* The same product with X and Y row-major (interleaved): row r's values of
 * the nrhs columns are adjacent, X[r * nrhs + j], Y[r * nrhs + j].  Every
 * column equals the single-vector result bit for bit, as above.  A row's nrhs
 * values must be a multiple of 16 bytes and at most
 * HBDIA_SPMM_RM_MAX_CHUNKS x 16 bytes (FP64: nrhs even, <= 64; FP32: nrhs a
 * multiple of 4, <= 128); X and Y 16-byte aligned; block_width % 64 == 0;
 * offsets padded as for the other multi-RHS entry points.  The x windows of
 * each run of consecutive offsets are staged into LDS with direct
 * global-to-LDS loads. */
#define HBDIA_SPMM_RM_MAX_CHUNKS 32
int hbdia_hbdia_spmm_rm_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const double* coo_values, const int32_t* coo_group_ptr,
                            const double* X, double* Y, int32_t nrhs, int64_t num_rows, hbdia_stream_t stream);
int hbdia_hbdia_spmm_rm_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const float* coo_values, const int32_t* coo_group_ptr,
                            const float* X, float* Y, int32_t nrhs, int64_t num_rows, hbdia_stream_t stream);

/* Reduced-precision storage (no reference counterpart: the reference stores
 * and computes in one type).  values_lp / coo_values_lp are FP32 or BF16 copies
 * (vdtype) of the FP64 arrays, made with hbdia_convert_values (FP64 -> FP32
 * round to nearest even, then for BF16 FP32 -> BF16 round to nearest even).
 * x and y are FP64.  Each row is accumulated in FP32 (products of the stored
 * value and x rounded to FP32) together with m = sum |a x| and the bound
 * e = (u_store + (n+2) 2^-24) m on its error (u_store = 2^-24 for FP32,
 * 2^-8 + 2^-24 for BF16; n terms).  A row is recomputed from the FP64 arrays,
 * with exactly the arithmetic of hbdia_hbdia_spmv_f64 / hbdia_bdia_spmv_f64,
 * when its cancellation ratio against the smallest consistent |y| exceeds
 * `ratio`: m > ratio (|sum| - e), or the sum is not finite.  Kept rows then
 * have a relative error of at most e / (|sum| - e) < u_eff * ratio.  coo_group_ptr == NULL: no
 * fallback.  workspace: hbdia_mixed_workspace_bytes(num_rows) bytes of device
 * memory (8-byte aligned); after the call (stream-ordered) it holds one 64-bit
 * mask per 64-row group whose bit r is set when row 64 g + r was recomputed in
 * FP64.  Every call rewrites every mask, so nothing needs clearing; calls that
 * share a workspace must not overlap.  block_width % 64 == 0. */
#define HBDIA_VALUES_F32 1
#define HBDIA_VALUES_BF16 2
int hbdia_convert_values(const double* src, void* dst, int64_t n, int vdtype, hbdia_stream_t stream);
size_t hbdia_mixed_workspace_bytes(int64_t num_rows);
int hbdia_hbdia_spmv_mixed(int vdtype, const void* values_lp, const double* values_f64, const int32_t* offsets,
                           const int32_t* block_start, int32_t num_blocks, int32_t block_width,
                           const int32_t* coo_rows, const int32_t* coo_cols, const void* coo_values_lp,
                           const double* coo_values_f64, const int32_t* coo_group_ptr, const double* x,
                           double* y, int64_t num_rows, double ratio, void* workspace, hbdia_stream_t stream);
/* The same with the runs-of-three block bitset (NULL: no hints); results and
 * masks are bit-identical. */
int hbdia_hbdia_spmv_mixed_ex(int vdtype, const void* values_lp, const double* values_f64, const int32_t* offsets,
                              const int32_t* block_start, int32_t num_blocks, int32_t block_width,
                              const int32_t* coo_rows, const int32_t* coo_cols, const void* coo_values_lp,
                              const double* coo_values_f64, const int32_t* coo_group_ptr, const double* x,
                              double* y, int64_t num_rows, double ratio, void* workspace,
                              const uint32_t* runs3_blocks, hbdia_stream_t stream);

/* Allocation-free form of hbdia_hbdia_spmv_mixed_ex (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same product and results, with the per-group FP64 masks kept in a
 * library-owned device array (2 MiB per device, held until the process ends)
 * instead of a workspace argument; the masks are not returned.  It calls no
 * allocating routine; the first call on a device resolves that device's copy
 * of the array.  num_rows <= HBDIA_NOALLOC_MAX_ROWS, else
 * HBDIA_ERROR_INVALID_ARGUMENT.  All calls on one device share the array:
 * calls that overlap on one device (other streams, other threads) race. */
#define HBDIA_NOALLOC_MAX_ROWS (1LL << 24)
int hbdia_hbdia_spmv_mixed_noalloc(int vdtype, const void* values_lp, const double* values_f64,
                                   const int32_t* offsets, const int32_t* block_start, int32_t num_blocks,
                                   int32_t block_width, const int32_t* coo_rows, const int32_t* coo_cols,
                                   const void* coo_values_lp, const double* coo_values_f64,
                                   const int32_t* coo_group_ptr, const double* x, double* y, int64_t num_rows,
                                   double ratio, const uint32_t* runs3_blocks, hbdia_stream_t stream);

/* Many independent matrices in one persistent launch (no reference
 * counterpart: the reference multiplies one matrix per call).  The matrices'
 * HBDIA arrays are concatenated into pools (values, offsets, fallback rows /
 * columns / values); each matrix keeps the single-matrix layout inside its part
 * (fallback rows and columns relative to the matrix), and its rows occupy a
 * contiguous range of the x and y pools.  The work is cut into tiles of 64
 * consecutive rows of one matrix, each described by a tile record (built on
 * the host by hbdia_batch_make_tiles), and every matrix's rows come out bit for
 * bit as hbdia_hbdia_spmv_* (matrix with a fallback) or hbdia_bdia_spmv_*
 * (without) would give them for that matrix alone.  The grid is persistent:
 * num_workgroups workgroups of 4 waves, each wave taking tiles
 *   HBDIA_BATCH_INTERLEAVED: runs of tiles_per_step consecutive tiles, wave w
 *     taking runs w, w + W, w + 2W, ... (W = 4 num_workgroups);
 *   HBDIA_BATCH_CONTIGUOUS: tiles [wave_tiles[w], wave_tiles[w+1]), a split
 *     that the host balances by work (wave_tiles has W + 1 entries);
 *   HBDIA_BATCH_XCD: the tiles cut into eight contiguous ranges balanced by
 *     work, [wave_tiles[k], wave_tiles[k+1]) (wave_tiles has 9 entries), range
 *     k taken by the workgroups that run on XCD k (workgroup b runs on XCD
 *     b % 8; num_workgroups a multiple of 8) with the interleaved order inside
 *     it.  A matrix then runs on one XCD, whose L2 keeps its x lines for the
 *     band loads of neighbouring tiles and the fallback's gathers.
 * A tile record carries its block's first offsets, so a wave needs one
 * scalar load per tile, issued for the next tile while the current tile's
 * loads are in flight.  Pools need padding past their last entry: offsets
 * HBDIA_OFFSETS_PAD entries, fallback rows / columns / values
 * HBDIA_BATCH_COO_PAD entries (any values; columns must be 0), values 64. */
#define HBDIA_BATCH_INTERLEAVED 0
#define HBDIA_BATCH_CONTIGUOUS 1
#define HBDIA_BATCH_XCD 2
#define HBDIA_BATCH_COO_PAD 64
#define HBDIA_BATCH_CHUNK 9 /* offsets the batch kernel takes per step */

/* One matrix of a batch, as placed in the pools (host-side description). */
typedef struct hbdia_batch_desc {
  int64_t values;      /* first value of the matrix in the values pool */
  int64_t rows;        /* first row of the matrix in the x and y pools */
  int64_t first_tile;  /* index of its first tile */
  int32_t offsets;     /* first of its flattened offsets in the offsets pool */
  int32_t block_start; /* first of its num_blocks + 1 block starts (relative to its offsets) */
  int32_t coo;         /* first of its fallback entries in the fallback pools */
  int32_t coo_nnz;     /* its fallback entries (0: blocked part only) */
  int32_t group_ptr;   /* first of its ceil(num_rows/64) + 1 group pointers (relative to its fallback) */
  int32_t num_rows;
  int32_t num_blocks;
  int32_t block_width; /* multiple of 64 */
} hbdia_batch_desc_t;

/* One tile: 64 rows of one matrix (80 bytes, one scalar load). */
#define HBDIA_BATCH_INLINE_OFFSETS 10
typedef struct hbdia_batch_tile {
  int32_t value_line; /* values-pool index / 64 of the tile's first row in its block's first offset line */
  int32_t rows;       /* first row of the tile's matrix in the x and y pools */
  int32_t row0;       /* first row of the tile within its matrix (a multiple of 64) */
  int32_t num_rows;   /* rows of the matrix */
  int32_t stride;     /* block width: distance between offset lines, in values */
  int32_t count;      /* offsets of the tile's block (0: empty block) */
  int32_t offsets;    /* offsets-pool index of the block's offset list, or of the previous tile's
                         when the two lists are equal (so equal lists of consecutive tiles share
                         one index) */
  int32_t coo_begin;  /* fallback-pool range [coo_begin, coo_end) of the tile's rows */
  int32_t coo_end;
  int32_t fused;      /* 1: every row gets (blocked sum or old y) + fallback; 0: blocked part only */
  int32_t first_offsets[HBDIA_BATCH_INLINE_OFFSETS]; /* the block's first offsets; 0 past count */
} hbdia_batch_tile_t;

typedef struct hbdia_batch {
  const hbdia_batch_tile_t* tiles; /* device, num_tiles records */
  const int32_t* wave_tiles;       /* device, contiguous schedule only (else NULL) */
  const void* values;              /* device pools (padded as above) */
  const int32_t* offsets;
  const int32_t* coo_rows;
  const int32_t* coo_cols;
  const void* coo_values;
  int64_t num_tiles;
  int32_t has_coo; /* any tile with fused == 1 */
  int32_t x32;     /* every matrix: num_rows * sizeof(value) < 2^32 */
  int32_t long_lists; /* 1: run the kernel that pairs tiles with more than HBDIA_BATCH_CHUNK offsets
                         (a hint: either value is correct; HBDIABatch sets it when such tiles hold
                         most of the values, since that kernel holds more registers) */
} hbdia_batch_t;

typedef struct hbdia_batch_schedule {
  int32_t kind;           /* HBDIA_BATCH_INTERLEAVED, HBDIA_BATCH_CONTIGUOUS or HBDIA_BATCH_XCD */
  int32_t tiles_per_step; /* interleaved: tiles per run, >= 1 */
  int32_t num_workgroups; /* persistent grid size, >= 1 */
} hbdia_batch_schedule_t;

/* Host function: fills tiles[desc[m].first_tile + g] for every matrix m and
 * 64-row group g from the host copies of the offsets, block-start and
 * group-pointer pools.  value_size: 4 or 8.  Fails when a block's values
 * exceed 2^32 bytes or a pool index exceeds 2^31. */
int hbdia_batch_make_tiles(const hbdia_batch_desc_t* desc, int32_t num_matrices, const int32_t* offsets,
                           const int32_t* block_start, const int32_t* group_ptr, int32_t value_size,
                           hbdia_batch_tile_t* tiles, int64_t num_tiles);
int hbdia_batch_spmv_f64(const hbdia_batch_t* batch, const hbdia_batch_schedule_t* schedule, const double* x,
                         double* y, hbdia_stream_t stream);
int hbdia_batch_spmv_f32(const hbdia_batch_t* batch, const hbdia_batch_schedule_t* schedule, const float* x,
                         float* y, hbdia_stream_t stream);
/* Serving: one persistent launch over a list of jobs, each job one matrix of a
 * resident pool (tiles [tile_begin, tile_begin + its tile count) of the pool's
 * tile records) applied to its own x and y: device addresses of the job's
 * vectors, element 0 at the matrix's row 0 (a client's own device vectors, or
 * slots of a staging arena).  The launch's work is the concatenation of the
 * jobs' tiles: job k owns work indices [work_begin_k, work_begin_{k+1})
 * (work_begin ascending, starting at 0; total_tiles = the sum of all jobs'
 * tiles).  y is written in full, as hbdiaSpMV on the matrix with y starting at
 * zero: nothing needs clearing.  Interleaved schedule (runs of tiles_per_step
 * tiles, num_workgroups workgroups); every job's rows come out bit for bit as
 * in the batch and single-matrix kernels. */
typedef struct hbdia_serve_job {
  int64_t work_begin;
  int64_t tile_begin;
  uint64_t x; /* const T* */
  uint64_t y; /* T* */
} hbdia_serve_job_t;
int hbdia_serve_spmv_f64(const hbdia_batch_t* pool, const hbdia_serve_job_t* jobs, int32_t num_jobs,
                         int64_t total_tiles, int32_t num_workgroups, int32_t tiles_per_step, hbdia_stream_t stream);
int hbdia_serve_spmv_f32(const hbdia_batch_t* pool, const hbdia_serve_job_t* jobs, int32_t num_jobs,
                         int64_t total_tiles, int32_t num_workgroups, int32_t tiles_per_step, hbdia_stream_t stream);

/* Workgroups of the batch kernel that one compute unit holds at once (the
 * kernel variant that `batch`, `kind` and value_size (4 or 8) select); a grid
 * of at most this times the compute units is resident as a whole. */
int hbdia_batch_occupancy(const hbdia_batch_t* batch, int32_t kind, int32_t value_size, int* workgroups_per_cu);

/* COO fallback alone, into a result buffer: res[r] = sequential FMA sum of the
 * row's entries for rows that have entries; other rows keep res.  The GPU
 * counterpart of cpu_coo_spmv_kernel (HBDIASpMV.cpp:35-99). */
int hbdia_coo_spmv_f64(const int32_t* coo_rows, const int32_t* coo_cols, const double* coo_values,
                       const int32_t* coo_group_ptr, const double* x, double* res,
                       int64_t num_rows, hbdia_stream_t stream);
int hbdia_coo_spmv_f32(const int32_t* coo_rows, const int32_t* coo_cols, const float* coo_values,
                       const int32_t* coo_group_ptr, const float* x, float* res,
                       int64_t num_rows, hbdia_stream_t stream);

/* group_ptr[g] = first COO index whose row is >= 64 g, for g in [0, ceil(num_rows/64)];
 * coo_rows sorted ascending, all in [0, num_rows). */
int hbdia_coo_group_ptr(const int32_t* coo_rows, int64_t coo_nnz, int64_t num_rows,
                        int32_t* group_ptr, hbdia_stream_t stream);

/* dst[i] += src[i] for i < n; replaces launch_vector_add_kernel<T> (HBDIASpMV.cu:120-135). */
int hbdia_vector_add_f64(double* dst, const double* src, int64_t n, hbdia_stream_t stream);
int hbdia_vector_add_f32(float* dst, const float* src, int64_t n, hbdia_stream_t stream);

/* ---- format conversion on the device ----
 * The COO -> HBDIA conversion of HBDIA<T>::convertToHBDIAFormat
 * (src/Format/HBDIA.cpp:507-718) computed on the GPU, with the same results:
 * nb = ceil(min(num_rows, num_cols) / block_width) blocks of block_width rows;
 * per block the distinct offsets col - row (col - global_rows[row] when
 * global_rows is given) with their entry counts; a (block, offset) pair stays
 * in the block storage iff count >= threshold or the entries already sent to
 * the fallback (pairs visited blocks ascending, offsets ascending) number
 * max_coo_entries or more; kept offsets ascending per block; each entry of a
 * kept pair is assigned to values[pair * block_width + row % block_width], the
 * last one in input order winning; empty slots hold 0; every other entry goes
 * to the fallback COO in input order.  Entries must satisfy
 * 0 <= row < num_rows and 0 <= col < num_cols. */
typedef struct hbdia_convert_args {
  const int32_t* rows;        /* device, nnz entries */
  const int32_t* cols;        /* device, nnz entries */
  const void* values;         /* device, nnz elements of elem_size bytes */
  const int32_t* global_rows; /* device, num_rows entries, or NULL */
  int64_t nnz;                /* 1 .. 2^31 - 1 */
  int32_t num_rows, num_cols;
  int32_t block_width;        /* > 0 */
  int32_t threshold;
  int32_t max_coo_entries;
  int32_t elem_size;          /* 4 (int32, float) or 8 (double); values are copied bit for bit */
  int32_t want_row_sorted_coo; /* also return a row-sorted copy of the fallback when it is not */
  int32_t want_runs3;         /* also return the runs-of-three block bitset (block_width % 64 == 0 or 32) */
} hbdia_convert_args_t;

/* Device arrays in the result are allocated by hbdia_convert_coo and owned by the
 * caller: release each with hbdia_free, or all of them with
 * hbdia_convert_release.  NULL when empty. */
typedef struct hbdia_convert_result {
  void* values;               /* num_pairs * block_width elements: the block storage */
  int32_t* offsets;           /* num_pairs + HBDIA_OFFSETS_PAD; kept offsets, blocks ascending; zero tail */
  int32_t* block_start;       /* num_blocks + 1: first pair of each block */
  int32_t* block_sizes;       /* num_blocks + 1: kept pairs per block (last entry 0) */
  int32_t* block_placed;      /* num_blocks: entries placed in each block's storage */
  uint32_t* runs3_blocks;     /* ceil(num_blocks / 32) words: bit b set when block b's offsets are
                                 runs of three consecutive values; NULL if no block is */
  int32_t* coo_rows;          /* num_coo: fallback in input order */
  int32_t* coo_cols;
  void* coo_values;
  int32_t* coo_rows_sorted;   /* num_coo: the fallback stably sorted by row; NULL when coo_sorted */
  int32_t* coo_cols_sorted;
  void* coo_values_sorted;
  int64_t num_pairs;          /* kept (block, offset) pairs */
  int64_t num_coo;            /* fallback entries */
  int32_t num_blocks;
  int32_t number_diagonals;   /* distinct offsets over all pairs, kept or not */
  int32_t runs3_count;        /* blocks with the runs-of-three bit */
  int32_t coo_sorted;         /* 1 when the input-order fallback is already sorted by row */
} hbdia_convert_result_t;

/* Runs on `stream` and returns after the stream has finished the conversion.
 * On failure every array it allocated is released and *out is zeroed.
 * Temporaries (28 to 60 bytes per entry) come from a cache of device blocks of
 * the current device that keeps them for the next conversion until
 * hbdia_convert_trim.  Calls on different threads may overlap. */
int hbdia_convert_coo(const hbdia_convert_args_t* args, hbdia_convert_result_t* out, hbdia_stream_t stream);
/* Frees the current device's cached conversion temporaries.  Must not overlap a
 * conversion on the same device. */
int hbdia_convert_trim(void);
/* hbdia_free on every non-NULL array of *res, then zero it. */
int hbdia_convert_release(hbdia_convert_result_t* res);

#ifdef __cplusplus
}
#endif

#endif /* HBDIA_HIP_H */
