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
 *   hbdia_matrix_convert[_host]        -> HBDIA<T>::convertToHBDIAFormat, HBDIA.hpp:52-54
 *   hbdia_matrix_remove_duplicates     -> HBDIA<T>::removeCOODuplicates, HBDIA.hpp:71
 *   hbdia_matrix_info / _get_*         -> the getters of HBDIA.hpp:74-110
 *   hbdia_vector_create                -> HBDIAVector<T>(localData[, four memory flags]),
 *                                         include/Format/HBDIAVector.hpp:31-32
 *   hbdia_spmv                         -> hbdiaSpMV<T>, include/Operations/HBDIASpMV.hpp:8-10
 *   hbdia_spmv_multi                   -> no reference counterpart: many right-hand sides in one launch
 *   hbdia_mixed_*                      -> no reference counterpart: FP32 / BF16 value storage
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
  HBDIA_INFO_COUNT = 16
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
int hbdia_matrix_remove_duplicates(hbdia_matrix_t m);
int hbdia_matrix_prepare_gpu(hbdia_matrix_t m);
int hbdia_matrix_synchronize(hbdia_matrix_t m);
int hbdia_matrix_info(hbdia_matrix_t m, int64_t* info /* [HBDIA_INFO_COUNT] */);
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
/* hbdiaSpMVMulti: X, Y device arrays (column-major, leading dimensions ldx, ldy). */
int hbdia_spmv_multi(hbdia_matrix_t m, const void* X, int64_t ldx, void* Y, int64_t ldy, int nrhs, int exec_coo_gpu);

/* Reduced-precision storage (HBDIAMixed, include/Operations/HBDIAMixed.hpp):
 * precision 1 = FP32, 2 = BF16; an FP64 matrix only.  x, y: FP64 vectors. */
typedef struct hbdia_mixed_opaque* hbdia_mixed_t;
int hbdia_mixed_create(hbdia_mixed_t* h, hbdia_matrix_t m, int precision, double cancellation_ratio);
int hbdia_mixed_destroy(hbdia_mixed_t h);
int hbdia_mixed_spmv(hbdia_mixed_t h, hbdia_vector_t x, hbdia_vector_t y);
int hbdia_mixed_spmv_raw(hbdia_mixed_t h, const void* x, void* y);
int hbdia_mixed_fp64_rows(hbdia_mixed_t h, int64_t* rows);

/* Seeded benchmark inputs (generators in Workloads.hpp). */
int hbdia_gen_stencil(hbdia_matrix_t* m, int dtype, int n, int points, double scatter, uint64_t seed);
int hbdia_gen_banded(hbdia_matrix_t* m, int dtype, int n, int noffsets, int max_offset, double offband,
                     uint64_t seed);
int hbdia_gen_vector(int dtype, int64_t n, uint64_t seed, double lo, double hi, void* out);

#ifdef __cplusplus
}
#endif

#endif /* HBDIA_CAPI_H */
