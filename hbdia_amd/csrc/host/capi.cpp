// capi.cpp -- C handle API over the C++ HBDIA classes (see hbdia_capi.h).
#include "hbdia_capi.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "DataExchange/BasicExtractor.hpp"
#include "Format/HBDIA.hpp"
#include "Format/HBDIAVector.hpp"
#include "Operations/HBDIABatch.hpp"
#include "Operations/HBDIAMixed.hpp"
#include "Operations/HBDIAServe.hpp"
#include "Operations/HBDIASpMV.hpp"
#include "Operations/HBDIAStats.hpp"
#include "Workloads.hpp"

struct hbdia_matrix_opaque {
  int dtype = HBDIA_DTYPE_F64;
  std::unique_ptr<HBDIA<double>> d;
  std::unique_ptr<HBDIA<float>> f;
};

struct hbdia_mixed_opaque {
  std::unique_ptr<HBDIAMixed> impl;
};

struct hbdia_batch_opaque {
  int dtype = HBDIA_DTYPE_F64;
  std::unique_ptr<HBDIABatch<double>> d;
  std::unique_ptr<HBDIABatch<float>> f;
};

struct hbdia_server_opaque {
  int dtype = HBDIA_DTYPE_F64;
  std::unique_ptr<HBDIAServer<double>> d;
  std::unique_ptr<HBDIAServer<float>> f;
};

struct hbdia_arena_opaque {
  hbdia::VectorArena arena{1};
};

struct hbdia_vector_opaque {
  int dtype = HBDIA_DTYPE_F64;
  std::unique_ptr<HBDIAVector<double>> d;
  std::unique_ptr<HBDIAVector<float>> f;
};

namespace {

thread_local std::string g_err;

template <typename F>
int guard(const char* what, F&& fn) {
  try {
    fn();
    return 0;
  } catch (const hbdia::DeviceError& e) {
    g_err = std::string(what) + ": " + e.what();
    return e.code() ? e.code() : 2;
  } catch (const std::invalid_argument& e) {
    g_err = std::string(what) + ": invalid_argument: " + e.what();
    return 1;
  } catch (const std::out_of_range& e) {
    g_err = std::string(what) + ": out_of_range: " + e.what();
    return 1;
  } catch (const std::exception& e) {
    g_err = std::string(what) + ": " + e.what();
    return 5;
  } catch (...) {
    g_err = std::string(what) + ": unknown exception";
    return 5;
  }
}

void need(bool ok, const char* msg) {
  if (!ok) throw std::invalid_argument(msg);
}

template <typename T>
void info_of(const HBDIA<T>& m, int64_t* info) {
  info[HBDIA_INFO_ROWS] = m.getNumRows();
  info[HBDIA_INFO_COLS] = m.getNumCols();
  info[HBDIA_INFO_NNZ] = m.getNumNonZeros();
  info[HBDIA_INFO_HAS_COO] = m.getHasCOO();
  info[HBDIA_INFO_HAS_HBDIA] = m.getHasHBDIA();
  info[HBDIA_INFO_BLOCK_WIDTH] = m.getBlockWidth();
  info[HBDIA_INFO_THRESHOLD] = m.getThreshold();
  info[HBDIA_INFO_NUM_BLOCKS] = m.getNumBlocks();
  info[HBDIA_INFO_NUM_VALUES] = static_cast<int64_t>(m.getNumStoredValues());
  int64_t noff = 0;
  for (const auto& o : m.getOffsetsPerBlock()) noff += static_cast<int64_t>(o.size());
  info[HBDIA_INFO_NUM_OFFSETS] = noff;
  info[HBDIA_INFO_NUM_COO] = static_cast<int64_t>(m.getCpuValues().size());
  info[HBDIA_INFO_NUM_DIAGONALS] = m.getNumberDiagonals();
  info[HBDIA_INFO_HIST_LEN] = static_cast<int64_t>(m.getHistogramBlocks().size());
  info[HBDIA_INFO_ON_DEVICE] = m.areStreamsInitialized();
  info[HBDIA_INFO_DEVICE_COO] = m.hasDeviceCOO();
  info[HBDIA_INFO_STORED_BYTES] = static_cast<int64_t>(m.storedBytes());
  info[HBDIA_INFO_RUNS3_BLOCKS] = m.getRuns3BlockCount();
  info[HBDIA_INFO_HOST_VALUES] = m.hasHostValues() ? 1 : 0;
}

template <typename T>
void get_hbdia(const HBDIA<T>& m, void* values, int32_t* offsets, int32_t* blockStart, int32_t* cr, int32_t* cc,
               void* cv, int32_t* rp, int32_t* hb, int32_t* hn) {
  if (values) std::copy(m.getHBDIAData().begin(), m.getHBDIAData().end(), static_cast<T*>(values));
  const auto& ob = m.getOffsetsPerBlock();
  int64_t pos = 0;
  for (size_t b = 0; b < ob.size(); ++b) {
    if (blockStart) blockStart[b] = static_cast<int32_t>(pos);
    if (offsets) std::copy(ob[b].begin(), ob[b].end(), offsets + pos);
    pos += static_cast<int64_t>(ob[b].size());
  }
  if (blockStart) blockStart[ob.size()] = static_cast<int32_t>(pos);
  if (cr) std::copy(m.getCpuRowIndices().begin(), m.getCpuRowIndices().end(), cr);
  if (cc) std::copy(m.getCpuColIndices().begin(), m.getCpuColIndices().end(), cc);
  if (cv) std::copy(m.getCpuValues().begin(), m.getCpuValues().end(), static_cast<T*>(cv));
  if (rp) std::copy(m.getCpuRowPtr().begin(), m.getCpuRowPtr().end(), rp);
  if (hb) std::copy(m.getHistogramBlocks().begin(), m.getHistogramBlocks().end(), hb);
  if (hn) std::copy(m.getHistogramNnz().begin(), m.getHistogramNnz().end(), hn);
}

template <typename T>
void dev_ptrs(const HBDIA<T>& m, void** p) {
  p[0] = const_cast<T*>(m.getHBDIADataDevice());
  p[1] = const_cast<int*>(m.getFlattenedOffsetsDevice());
  p[2] = const_cast<int*>(m.getBlockStartIndicesDevice());
  p[3] = const_cast<int*>(m.getBlockSizesDevice());
  p[4] = const_cast<int*>(m.getCpuRowIndicesDevice());
  p[5] = const_cast<int*>(m.getCpuColIndicesDevice());
  p[6] = const_cast<T*>(m.getCpuValuesDevice());
  p[7] = const_cast<int*>(m.getCpuRowPtrDevice());
  p[8] = const_cast<int*>(m.getCooGroupPtrDevice());
  p[9] = m.getBDIAStream();
}

template <typename T>
std::unique_ptr<HBDIA<T>> from_coo(const int32_t* r, const int32_t* c, const void* v, int64_t nnz, int32_t nr,
                                   int32_t nc) {
  need(nnz >= 0, "nnz must be non-negative");
  need(nnz == 0 || (r && c && v), "null COO pointer");
  std::vector<int> rr(r, r + nnz), cc(c, c + nnz);
  std::vector<T> vv(static_cast<const T*>(v), static_cast<const T*>(v) + nnz);
  if (nr <= 0 && nc <= 0) return std::make_unique<HBDIA<T>>(rr, cc, vv);
  return std::make_unique<HBDIA<T>>(std::move(rr), std::move(cc), std::move(vv), nr, nc);
}

template <typename T>
std::unique_ptr<HBDIA<T>> from_gen(hbdia::CooMatrix<T>& g) {
  const int nr = g.numRows, nc = g.numCols;
  return std::make_unique<HBDIA<T>>(std::move(g.rows), std::move(g.cols), std::move(g.vals), nr, nc);
}

}  // namespace

namespace {

template <typename T>
void local_info(const HBDIA<T>& m, int64_t* info) {
  info[HBDIA_LOCAL_FIRST_ROW] = m.getLocalFirstRow();
  info[HBDIA_LOCAL_HALO_LEFT] = m.getHaloLeftSize();
  info[HBDIA_LOCAL_HALO_RIGHT] = m.getHaloRightSize();
  info[HBDIA_LOCAL_COO] = static_cast<int64_t>(m.getLocalCOOValues().size());
  info[HBDIA_LOCAL_FOLDED] = m.getLocalCOOFolded();
  info[HBDIA_LOCAL_RANGES] = static_cast<int64_t>(m.getDataRanges().size());
}

template <typename T>
void get_local(const HBDIA<T>& m, int32_t* halo, int32_t* cr, int32_t* cc, void* cv, int32_t* ranges) {
  if (halo) std::copy(m.getHaloColumns().begin(), m.getHaloColumns().end(), halo);
  if (cr) std::copy(m.getLocalCOORowIndices().begin(), m.getLocalCOORowIndices().end(), cr);
  if (cc) std::copy(m.getLocalCOOColIndices().begin(), m.getLocalCOOColIndices().end(), cc);
  if (cv) std::copy(m.getLocalCOOValues().begin(), m.getLocalCOOValues().end(), static_cast<T*>(cv));
  if (ranges) {
    const auto& dr = m.getDataRanges();
    for (size_t i = 0; i < dr.size(); ++i) {
      ranges[2 * i] = std::get<0>(dr[i]);
      ranges[2 * i + 1] = std::get<1>(dr[i]);
    }
  }
}

template <typename T>
void get_vector_offsets(const HBDIA<T>& m, int32_t* out) {
  int64_t pos = 0;
  for (const auto& vo : m.getVectorOffsets()) {
    std::copy(vo.begin(), vo.end(), out + pos);
    pos += static_cast<int64_t>(vo.size());
  }
}

}  // namespace

extern "C" {

const char* hbdia_capi_last_error(void) { return g_err.c_str(); }

int hbdia_set_num_threads(int n) {
  return guard("hbdia_set_num_threads", [&] {
    need(n > 0, "thread count must be positive");
    omp_set_num_threads(n);
  });
}
int hbdia_get_num_threads(void) { return omp_get_max_threads(); }

int hbdia_matrix_create(hbdia_matrix_t* m, int dtype) {
  return guard("hbdia_matrix_create", [&] {
    need(m != nullptr, "null handle pointer");
    need(dtype == HBDIA_DTYPE_F64 || dtype == HBDIA_DTYPE_F32, "dtype must be F64 or F32");
    auto h = std::make_unique<hbdia_matrix_opaque>();
    h->dtype = dtype;
    if (dtype == HBDIA_DTYPE_F64) h->d = std::make_unique<HBDIA<double>>();
    else h->f = std::make_unique<HBDIA<float>>();
    *m = h.release();
  });
}

int hbdia_matrix_create_coo(hbdia_matrix_t* m, int dtype, const int32_t* rows, const int32_t* cols,
                            const void* values, int64_t nnz, int32_t num_rows, int32_t num_cols) {
  return guard("hbdia_matrix_create_coo", [&] {
    need(m != nullptr, "null handle pointer");
    need(dtype == HBDIA_DTYPE_F64 || dtype == HBDIA_DTYPE_F32, "dtype must be F64 or F32");
    auto h = std::make_unique<hbdia_matrix_opaque>();
    h->dtype = dtype;
    if (dtype == HBDIA_DTYPE_F64) h->d = from_coo<double>(rows, cols, values, nnz, num_rows, num_cols);
    else h->f = from_coo<float>(rows, cols, values, nnz, num_rows, num_cols);
    *m = h.release();
  });
}

int hbdia_matrix_destroy(hbdia_matrix_t m) {
  return guard("hbdia_matrix_destroy", [&] { delete m; });
}

#define DISPATCH(m, expr_d, expr_f)             \
  do {                                          \
    need((m) != nullptr, "null matrix handle"); \
    if ((m)->dtype == HBDIA_DTYPE_F64) {        \
      auto& M = *(m)->d;                        \
      expr_d;                                   \
    } else {                                    \
      auto& M = *(m)->f;                        \
      expr_f;                                   \
    }                                           \
  } while (0)

int hbdia_matrix_stencil27(hbdia_matrix_t m, int nx, int ny, int nz, double noise, int iteration) {
  return guard("hbdia_matrix_stencil27", [&] {
    need(nx > 0 && ny > 0 && nz > 0, "grid dimensions must be positive");
    DISPATCH(m, M.create3DStencil27Point(nx, ny, nz, noise, iteration),
             M.create3DStencil27Point(nx, ny, nz, noise, iteration));
  });
}

int hbdia_matrix_convert(hbdia_matrix_t m, int bw, int thr, int max_coo, int coo_unique, int gpu_coo) {
  return guard("hbdia_matrix_convert", [&] {
    const ExecutionMode em = gpu_coo ? ExecutionMode::GPU_COO : ExecutionMode::CPU_COO;
    DISPATCH(m, M.convertToHBDIAFormat(bw, thr, max_coo, coo_unique != 0, em),
             M.convertToHBDIAFormat(bw, thr, max_coo, coo_unique != 0, em));
  });
}

int hbdia_matrix_convert_host(hbdia_matrix_t m, int bw, int thr, int max_coo, int coo_unique, int gpu_coo) {
  return guard("hbdia_matrix_convert_host", [&] {
    const ExecutionMode em = gpu_coo ? ExecutionMode::GPU_COO : ExecutionMode::CPU_COO;
    DISPATCH(m, M.convertToHBDIAFormatOnHost(bw, thr, max_coo, coo_unique != 0, em),
             M.convertToHBDIAFormatOnHost(bw, thr, max_coo, coo_unique != 0, em));
  });
}

int hbdia_matrix_convert_device(hbdia_matrix_t m, int bw, int thr, int max_coo, int coo_unique, int gpu_coo) {
  return hbdia_matrix_convert_device_ex(m, bw, thr, max_coo, coo_unique, gpu_coo, 1);
}

int hbdia_matrix_convert_device_ex(hbdia_matrix_t m, int bw, int thr, int max_coo, int coo_unique, int gpu_coo,
                                   int host_values) {
  return guard("hbdia_matrix_convert_device", [&] {
    const ExecutionMode em = gpu_coo ? ExecutionMode::GPU_COO : ExecutionMode::CPU_COO;
    DISPATCH(m, M.convertToHBDIAFormatOnDevice(bw, thr, max_coo, coo_unique != 0, em, host_values != 0),
             M.convertToHBDIAFormatOnDevice(bw, thr, max_coo, coo_unique != 0, em, host_values != 0));
  });
}

int hbdia_matrix_remove_duplicates(hbdia_matrix_t m) {
  return guard("hbdia_matrix_remove_duplicates", [&] { DISPATCH(m, M.removeCOODuplicates(), M.removeCOODuplicates()); });
}

int hbdia_matrix_load_mtx(hbdia_matrix_t m, const char* filename, int num_rows, int num_cols, int* loaded) {
  return guard("hbdia_matrix_load_mtx", [&] {
    need(m != nullptr && filename != nullptr, "null argument");
    bool ok;
    auto load = [&](auto& A) {
      return (num_rows <= 0 && num_cols <= 0) ? A.loadMTX(filename) : A.loadMTX(filename, num_rows, num_cols);
    };
    ok = m->dtype == HBDIA_DTYPE_F64 ? load(*m->d) : load(*m->f);
    if (loaded) *loaded = ok ? 1 : 0;
  });
}

int hbdia_matrix_delete_hbdia(hbdia_matrix_t m) {
  return guard("hbdia_matrix_delete_hbdia", [&] { DISPATCH(m, M.deleteHBDIAFormat(), M.deleteHBDIAFormat()); });
}

int hbdia_matrix_prepare_gpu(hbdia_matrix_t m) {
  return guard("hbdia_matrix_prepare_gpu", [&] { DISPATCH(m, M.prepareForGPU(), M.prepareForGPU()); });
}

int hbdia_matrix_synchronize(hbdia_matrix_t m) {
  return guard("hbdia_matrix_synchronize", [&] { DISPATCH(m, M.synchronize(), M.synchronize()); });
}

int hbdia_matrix_info(hbdia_matrix_t m, int64_t* info) {
  return guard("hbdia_matrix_info", [&] {
    need(info != nullptr, "null info");
    DISPATCH(m, info_of(M, info), info_of(M, info));
  });
}

int hbdia_matrix_storage_report(hbdia_matrix_t m, int64_t* counts, double* ratios, int64_t* hist,
                                int64_t hist_capacity) {
  return guard("hbdia_matrix_storage_report", [&] {
    need(counts != nullptr && ratios != nullptr, "null output");
    need(hist_capacity >= 0 && (hist_capacity == 0 || hist != nullptr), "bad histogram buffer");
    HBDIAStorageReport r;
    DISPATCH(m, r = hbdiaStorageReport(M), r = hbdiaStorageReport(M));
    const int64_t c[HBDIA_REPORT_COUNT] = {r.rows, r.cols, r.nnz, r.numberDiagonals, r.numBlocks, r.blockWidth,
                                           r.threshold, r.maxCooEntries, r.cooFallbackEntries, r.hbdiaDataSize,
                                           r.storedOffsets, r.valueBytes, r.hbdiaBytes, r.cooBytes, r.diaBytes,
                                           r.sellBytes, r.csrBytes, r.sellApplies ? 1 : 0,
                                           static_cast<int64_t>(r.histogram.size())};
    std::copy(c, c + HBDIA_REPORT_COUNT, counts);
    ratios[0] = r.hbdiaVsCoo;
    ratios[1] = r.diaVsCoo;
    ratios[2] = r.csrVsCoo;
    ratios[3] = r.sellVsCoo;
    const int64_t n = std::min<int64_t>(hist_capacity, static_cast<int64_t>(r.histogram.size()));
    for (int64_t k = 0; k < n; ++k) {
      hist[3 * k] = r.histogram[k].diagonals;
      hist[3 * k + 1] = r.histogram[k].blocks;
      hist[3 * k + 2] = r.histogram[k].nnz;
    }
  });
}

int hbdia_matrix_storage_report_csv(hbdia_matrix_t m, const char* name, char* buf, size_t capacity,
                                    size_t* length) {
  return guard("hbdia_matrix_storage_report_csv", [&] {
    need(length != nullptr && (capacity == 0 || buf != nullptr), "bad output buffer");
    HBDIAStorageReport r;
    DISPATCH(m, r = hbdiaStorageReport(M), r = hbdiaStorageReport(M));
    const std::string text = hbdiaStorageReportCSV(r, name ? name : "");
    *length = text.size();
    if (capacity > 0) {
      const size_t n = std::min(capacity - 1, text.size());
      std::memcpy(buf, text.data(), n);
      buf[n] = '\0';
    }
  });
}

int hbdia_matrix_get_coo(hbdia_matrix_t m, int32_t* rows, int32_t* cols, void* values) {
  return guard("hbdia_matrix_get_coo", [&] {
    DISPATCH(
        m,
        {
          if (rows) std::copy(M.getRowIndices().begin(), M.getRowIndices().end(), rows);
          if (cols) std::copy(M.getColIndices().begin(), M.getColIndices().end(), cols);
          if (values) std::copy(M.getValues().begin(), M.getValues().end(), static_cast<double*>(values));
        },
        {
          if (rows) std::copy(M.getRowIndices().begin(), M.getRowIndices().end(), rows);
          if (cols) std::copy(M.getColIndices().begin(), M.getColIndices().end(), cols);
          if (values) std::copy(M.getValues().begin(), M.getValues().end(), static_cast<float*>(values));
        });
  });
}

int hbdia_matrix_get_hbdia(hbdia_matrix_t m, void* values, int32_t* offsets, int32_t* block_start,
                           int32_t* coo_rows, int32_t* coo_cols, void* coo_values, int32_t* row_ptr,
                           int32_t* hist_blocks, int32_t* hist_nnz) {
  return guard("hbdia_matrix_get_hbdia", [&] {
    DISPATCH(m, get_hbdia(M, values, offsets, block_start, coo_rows, coo_cols, coo_values, row_ptr, hist_blocks, hist_nnz),
             get_hbdia(M, values, offsets, block_start, coo_rows, coo_cols, coo_values, row_ptr, hist_blocks, hist_nnz));
  });
}

int hbdia_matrix_device_ptrs(hbdia_matrix_t m, void** ptrs) {
  return guard("hbdia_matrix_device_ptrs", [&] {
    need(ptrs != nullptr, "null output");
    DISPATCH(m, dev_ptrs(M, ptrs), dev_ptrs(M, ptrs));
  });
}

int hbdia_vector_create(hbdia_vector_t* v, int dtype, const void* data, int64_t n, int unified, int umalloc,
                        int managed, int numa) {
  return guard("hbdia_vector_create", [&] {
    need(v != nullptr, "null handle pointer");
    need(n >= 0 && (n == 0 || data), "bad data");
    need(dtype == HBDIA_DTYPE_F64 || dtype == HBDIA_DTYPE_F32, "dtype must be F64 or F32");
    auto h = std::make_unique<hbdia_vector_opaque>();
    h->dtype = dtype;
    if (dtype == HBDIA_DTYPE_F64) {
      std::vector<double> d(static_cast<const double*>(data), static_cast<const double*>(data) + n);
      h->d = std::make_unique<HBDIAVector<double>>(d, unified != 0, umalloc != 0, managed != 0, numa != 0);
    } else {
      std::vector<float> d(static_cast<const float*>(data), static_cast<const float*>(data) + n);
      h->f = std::make_unique<HBDIAVector<float>>(d, unified != 0, umalloc != 0, managed != 0, numa != 0);
    }
    *v = h.release();
  });
}

int hbdia_vector_destroy(hbdia_vector_t v) {
  return guard("hbdia_vector_destroy", [&] { delete v; });
}

int hbdia_vector_upload(hbdia_vector_t v, const void* data, int64_t n) {
  return guard("hbdia_vector_upload", [&] {
    need(v != nullptr && (n == 0 || data), "bad argument");
    if (v->dtype == HBDIA_DTYPE_F64)
      v->d->upload(std::vector<double>(static_cast<const double*>(data), static_cast<const double*>(data) + n));
    else
      v->f->upload(std::vector<float>(static_cast<const float*>(data), static_cast<const float*>(data) + n));
  });
}

int hbdia_vector_download(hbdia_vector_t v, void* out, int64_t n) {
  return guard("hbdia_vector_download", [&] {
    need(v != nullptr && out != nullptr, "bad argument");
    const size_t es = v->dtype == HBDIA_DTYPE_F64 ? sizeof(double) : sizeof(float);
    const size_t local = v->dtype == HBDIA_DTYPE_F64 ? v->d->getLocalSize() : v->f->getLocalSize();
    need(static_cast<size_t>(n) == local, "size mismatch");
    void* src = v->dtype == HBDIA_DTYPE_F64 ? static_cast<void*>(v->d->getDeviceLocalPtr())
                                            : static_cast<void*>(v->f->getDeviceLocalPtr());
    CHECK_HBDIA(hbdia_memcpy(out, src, local * es));
  });
}

int hbdia_vector_download_coo(hbdia_vector_t v, void* out, int64_t n) {
  return guard("hbdia_vector_download_coo", [&] {
    need(v != nullptr && out != nullptr, "bad argument");
    const size_t es = v->dtype == HBDIA_DTYPE_F64 ? sizeof(double) : sizeof(float);
    const size_t local = v->dtype == HBDIA_DTYPE_F64 ? v->d->getLocalSize() : v->f->getLocalSize();
    need(static_cast<size_t>(n) == local, "size mismatch");
    void* src = v->dtype == HBDIA_DTYPE_F64 ? static_cast<void*>(v->d->getDeviceCOOResultsPtr())
                                            : static_cast<void*>(v->f->getDeviceCOOResultsPtr());
    CHECK_HBDIA(hbdia_memcpy(out, src, local * es));
  });
}

int hbdia_vector_device_ptrs(hbdia_vector_t v, void** local, void** coo) {
  return guard("hbdia_vector_device_ptrs", [&] {
    need(v != nullptr, "null vector handle");
    if (v->dtype == HBDIA_DTYPE_F64) {
      if (local) *local = v->d->getDeviceLocalPtr();
      if (coo) *coo = v->d->getDeviceCOOResultsPtr();
    } else {
      if (local) *local = v->f->getDeviceLocalPtr();
      if (coo) *coo = v->f->getDeviceCOOResultsPtr();
    }
  });
}

int hbdia_spmv(hbdia_matrix_t m, hbdia_vector_t x, hbdia_vector_t y, int exec_coo_cpu, int exec_coo_gpu) {
  return guard("hbdia_spmv", [&] {
    need(m && x && y, "null handle");
    need(m->dtype == x->dtype && m->dtype == y->dtype, "dtype mismatch between matrix and vectors");
    if (m->dtype == HBDIA_DTYPE_F64) hbdiaSpMV<double>(*m->d, *x->d, *y->d, exec_coo_cpu != 0, exec_coo_gpu != 0);
    else hbdiaSpMV<float>(*m->f, *x->f, *y->f, exec_coo_cpu != 0, exec_coo_gpu != 0);
  });
}

}  // extern "C"

namespace {
// benchHBDIA's loop (reference benchmarking/benchHBDIA.cpp:30-50) for one value type.
template <typename T>
void time_synced(HBDIA<T>& A, const HBDIAVector<T>& x, HBDIAVector<T>& y, bool cpu, bool gpu, int iterations,
                 double* times_ms) {
  const size_t bytes = static_cast<size_t>(A.getNumRows()) * sizeof(T);
  for (int i = 0; i < iterations; ++i) {
    CHECK_HBDIA(hbdia_memset(y.getDeviceLocalPtr(), 0, bytes, nullptr));
    CHECK_HBDIA(hbdia_memset(y.getDeviceCOOResultsPtr(), 0, bytes, nullptr));
    CHECK_HBDIA(hbdia_device_synchronize());
    const auto t0 = std::chrono::steady_clock::now();
    hbdiaSpMV<T>(A, x, y, cpu, gpu);
    CHECK_HBDIA(hbdia_device_synchronize());
    const auto t1 = std::chrono::steady_clock::now();
    times_ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
}
}  // namespace

extern "C" {

int hbdia_spmv_time_synced(hbdia_matrix_t m, hbdia_vector_t x, hbdia_vector_t y, int exec_coo_cpu, int exec_coo_gpu,
                           int iterations, double* times_ms) {
  return guard("hbdia_spmv_time_synced", [&] {
    need(m && x && y && times_ms, "null argument");
    need(iterations >= 0, "negative iteration count");
    need(m->dtype == x->dtype && m->dtype == y->dtype, "dtype mismatch between matrix and vectors");
    if (m->dtype == HBDIA_DTYPE_F64)
      time_synced<double>(*m->d, *x->d, *y->d, exec_coo_cpu != 0, exec_coo_gpu != 0, iterations, times_ms);
    else
      time_synced<float>(*m->f, *x->f, *y->f, exec_coo_cpu != 0, exec_coo_gpu != 0, iterations, times_ms);
  });
}

int hbdia_matrix_rank_local(hbdia_matrix_t* local, hbdia_matrix_t global, int rank, int size, int block_width,
                            int threshold, int max_coo_entries, int gpu_coo) {
  return guard("hbdia_matrix_rank_local", [&] {
    need(local != nullptr && global != nullptr, "null handle");
    const ExecutionMode em = gpu_coo ? ExecutionMode::GPU_COO : ExecutionMode::CPU_COO;
    auto h = std::make_unique<hbdia_matrix_opaque>();
    h->dtype = global->dtype;
    if (global->dtype == HBDIA_DTYPE_F64)
      h->d = std::make_unique<HBDIA<double>>(
          createRankLocalMatrix<double>(*global->d, rank, size, block_width, threshold, max_coo_entries, em));
    else
      h->f = std::make_unique<HBDIA<float>>(
          createRankLocalMatrix<float>(*global->f, rank, size, block_width, threshold, max_coo_entries, em));
    *local = h.release();
  });
}

int hbdia_matrix_build_local_coo(hbdia_matrix_t m) {
  return guard("hbdia_matrix_build_local_coo", [&] { DISPATCH(m, M.buildLocalCOO(), M.buildLocalCOO()); });
}


int hbdia_matrix_local_info(hbdia_matrix_t m, int64_t* info) {
  return guard("hbdia_matrix_local_info", [&] {
    need(info != nullptr, "null info");
    DISPATCH(m, local_info(M, info), local_info(M, info));
  });
}

int hbdia_matrix_get_local(hbdia_matrix_t m, int32_t* halo_cols, int32_t* coo_rows, int32_t* coo_cols,
                           void* coo_values, int32_t* ranges) {
  return guard("hbdia_matrix_get_local", [&] {
    DISPATCH(m, get_local(M, halo_cols, coo_rows, coo_cols, coo_values, ranges),
             get_local(M, halo_cols, coo_rows, coo_cols, coo_values, ranges));
  });
}

int hbdia_matrix_get_vector_offsets(hbdia_matrix_t m, int32_t* out) {
  return guard("hbdia_matrix_get_vector_offsets", [&] {
    need(out != nullptr, "null output");
    DISPATCH(m, get_vector_offsets(M, out), get_vector_offsets(M, out));
  });
}

int hbdia_vector_create_halo(hbdia_vector_t* v, int dtype, const void* data, int64_t n, hbdia_matrix_t m) {
  return guard("hbdia_vector_create_halo", [&] {
    need(v != nullptr && m != nullptr, "null handle");
    need(n >= 0 && (n == 0 || data), "bad data");
    need(dtype == m->dtype, "dtype mismatch between matrix and vector");
    auto h = std::make_unique<hbdia_vector_opaque>();
    h->dtype = dtype;
    if (dtype == HBDIA_DTYPE_F64) {
      std::vector<double> d(static_cast<const double*>(data), static_cast<const double*>(data) + n);
      h->d = std::make_unique<HBDIAVector<double>>(d, *m->d, m->d->getRank(), m->d->getSize());
    } else {
      std::vector<float> d(static_cast<const float*>(data), static_cast<const float*>(data) + n);
      h->f = std::make_unique<HBDIAVector<float>>(d, *m->f, m->f->getRank(), m->f->getSize());
    }
    *v = h.release();
  });
}

int hbdia_vector_fill_halo(hbdia_vector_t v, hbdia_matrix_t m, const void* global, int64_t n) {
  return guard("hbdia_vector_fill_halo", [&] {
    need(v != nullptr && m != nullptr && n >= 0 && (n == 0 || global), "bad argument");
    need(v->dtype == m->dtype, "dtype mismatch between matrix and vector");
    bool ok;
    if (m->dtype == HBDIA_DTYPE_F64)
      ok = exchangeHaloFromGlobal<double>(
          *m->d, *v->d, std::vector<double>(static_cast<const double*>(global), static_cast<const double*>(global) + n));
    else
      ok = exchangeHaloFromGlobal<float>(
          *m->f, *v->f, std::vector<float>(static_cast<const float*>(global), static_cast<const float*>(global) + n));
    if (!ok) throw std::invalid_argument("halo layout of the vector does not match the matrix");
  });
}

int hbdia_vector_verify_halo(hbdia_vector_t v, hbdia_matrix_t m, int* ok) {
  return guard("hbdia_vector_verify_halo", [&] {
    need(v != nullptr && m != nullptr && ok != nullptr, "bad argument");
    need(v->dtype == m->dtype, "dtype mismatch between matrix and vector");
    *ok = (m->dtype == HBDIA_DTYPE_F64 ? verifyHaloExchange<double>(*m->d, *v->d)
                                       : verifyHaloExchange<float>(*m->f, *v->f))
              ? 1
              : 0;
  });
}

int hbdia_spmv_multi(hbdia_matrix_t m, const void* X, int64_t ldx, void* Y, int64_t ldy, int nrhs, int exec_coo_gpu) {
  return guard("hbdia_spmv_multi", [&] {
    need(m != nullptr && X != nullptr && Y != nullptr, "null argument");
    if (m->dtype == HBDIA_DTYPE_F64)
      hbdiaSpMVMulti<double>(*m->d, static_cast<const double*>(X), ldx, static_cast<double*>(Y), ldy, nrhs,
                             exec_coo_gpu != 0);
    else
      hbdiaSpMVMulti<float>(*m->f, static_cast<const float*>(X), ldx, static_cast<float*>(Y), ldy, nrhs,
                            exec_coo_gpu != 0);
  });
}

int hbdia_spmv_multi_rm(hbdia_matrix_t m, const void* X, void* Y, int nrhs, int exec_coo_gpu) {
  return guard("hbdia_spmv_multi_rm", [&] {
    need(m != nullptr && X != nullptr && Y != nullptr, "null argument");
    if (m->dtype == HBDIA_DTYPE_F64)
      hbdiaSpMVMultiRowMajor<double>(*m->d, static_cast<const double*>(X), static_cast<double*>(Y), nrhs,
                                     exec_coo_gpu != 0);
    else
      hbdiaSpMVMultiRowMajor<float>(*m->f, static_cast<const float*>(X), static_cast<float*>(Y), nrhs,
                                    exec_coo_gpu != 0);
  });
}

int hbdia_mixed_create(hbdia_mixed_t* h, hbdia_matrix_t m, int precision, double ratio) {
  return guard("hbdia_mixed_create", [&] {
    need(h != nullptr && m != nullptr, "null argument");
    need(m->dtype == HBDIA_DTYPE_F64, "reduced-precision storage needs an FP64 matrix");
    need(precision == 1 || precision == 2, "precision must be 1 (FP32) or 2 (BF16)");
    auto o = std::make_unique<hbdia_mixed_opaque>();
    const auto p = static_cast<StoragePrecision>(precision);
    o->impl = ratio > 0.0 ? std::make_unique<HBDIAMixed>(*m->d, p, ratio) : std::make_unique<HBDIAMixed>(*m->d, p);
    *h = o.release();
  });
}
int hbdia_mixed_fp64_masks(hbdia_mixed_t h, uint64_t* masks, int64_t words) {
  return guard("hbdia_mixed_fp64_masks", [&] {
    need(h && masks, "null argument");
    const std::vector<uint64_t> m = h->impl->lastFp64Masks();
    need(words >= static_cast<int64_t>(m.size()), "mask buffer shorter than ceil(rows / 64) words");
    std::copy(m.begin(), m.end(), masks);
  });
}
int hbdia_mixed_error_bound(hbdia_mixed_t h, double* ratio, double* rel_error_bound) {
  return guard("hbdia_mixed_error_bound", [&] {
    need(h && ratio && rel_error_bound, "null argument");
    *ratio = h->impl->cancellationRatio();
    *rel_error_bound = h->impl->relErrorBound();
  });
}
int hbdia_mixed_destroy(hbdia_mixed_t h) {
  return guard("hbdia_mixed_destroy", [&] { delete h; });
}
int hbdia_mixed_spmv(hbdia_mixed_t h, hbdia_vector_t x, hbdia_vector_t y) {
  return guard("hbdia_mixed_spmv", [&] {
    need(h && x && y, "null argument");
    need(x->dtype == HBDIA_DTYPE_F64 && y->dtype == HBDIA_DTYPE_F64, "vectors must be FP64");
    h->impl->spmv(*x->d, *y->d);
  });
}
int hbdia_mixed_spmv_raw(hbdia_mixed_t h, const void* x, void* y) {
  return guard("hbdia_mixed_spmv_raw", [&] {
    need(h && x && y, "null argument");
    h->impl->spmv(static_cast<const double*>(x), static_cast<double*>(y));
  });
}
int hbdia_mixed_fp64_rows(hbdia_mixed_t h, int64_t* rows) {
  return guard("hbdia_mixed_fp64_rows", [&] {
    need(h && rows, "null argument");
    *rows = h->impl->lastFp64Rows();
  });
}

int hbdia_gen_stencil(hbdia_matrix_t* m, int dtype, int n, int points, double scatter, uint64_t seed) {
  return guard("hbdia_gen_stencil", [&] {
    need(m != nullptr, "null handle pointer");
    need(dtype == HBDIA_DTYPE_F64 || dtype == HBDIA_DTYPE_F32, "dtype must be F64 or F32");
    auto h = std::make_unique<hbdia_matrix_opaque>();
    h->dtype = dtype;
    if (dtype == HBDIA_DTYPE_F64) {
      hbdia::CooMatrix<double> g;
      hbdia::generateStencilScatter<double>(n, points, scatter, seed, g);
      h->d = from_gen(g);
    } else {
      hbdia::CooMatrix<float> g;
      hbdia::generateStencilScatter<float>(n, points, scatter, seed, g);
      h->f = from_gen(g);
    }
    *m = h.release();
  });
}

int hbdia_gen_banded(hbdia_matrix_t* m, int dtype, int n, int noffsets, int max_offset, double offband,
                     uint64_t seed) {
  return guard("hbdia_gen_banded", [&] {
    need(m != nullptr, "null handle pointer");
    need(dtype == HBDIA_DTYPE_F64 || dtype == HBDIA_DTYPE_F32, "dtype must be F64 or F32");
    auto h = std::make_unique<hbdia_matrix_opaque>();
    h->dtype = dtype;
    if (dtype == HBDIA_DTYPE_F64) {
      hbdia::CooMatrix<double> g;
      hbdia::generateBanded<double>(n, noffsets, max_offset, offband, seed, g);
      h->d = from_gen(g);
    } else {
      hbdia::CooMatrix<float> g;
      hbdia::generateBanded<float>(n, noffsets, max_offset, offband, seed, g);
      h->f = from_gen(g);
    }
    *m = h.release();
  });
}

int hbdia_gen_vector(int dtype, int64_t n, uint64_t seed, double lo, double hi, void* out) {
  return guard("hbdia_gen_vector", [&] {
    need(n >= 0 && (n == 0 || out), "bad argument");
    if (dtype == HBDIA_DTYPE_F64) hbdia::generateUniformVector<double>(n, seed, lo, hi, static_cast<double*>(out));
    else if (dtype == HBDIA_DTYPE_F32) hbdia::generateUniformVector<float>(n, seed, lo, hi, static_cast<float*>(out));
    else throw std::invalid_argument("dtype must be F64 or F32");
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// many-matrix batches

namespace {

template <typename T>
std::unique_ptr<HBDIA<T>>& member(hbdia_matrix_opaque* m);
template <>
std::unique_ptr<HBDIA<double>>& member<double>(hbdia_matrix_opaque* m) { return m->d; }
template <>
std::unique_ptr<HBDIA<float>>& member<float>(hbdia_matrix_opaque* m) { return m->f; }

int batch_dtype(const hbdia_matrix_t* ms, int count) {
  need(ms != nullptr && count > 0, "need at least one matrix");
  for (int i = 0; i < count; ++i) need(ms[i] != nullptr, "null matrix handle");
  const int dt = ms[0]->dtype;
  for (int i = 1; i < count; ++i) need(ms[i]->dtype == dt, "matrices of one dtype only");
  return dt;
}

template <typename F>
auto with_batch(hbdia_batch_handle_t b, F&& fn) {
  need(b != nullptr, "null batch handle");
  if (b->dtype == HBDIA_DTYPE_F64) return fn(*b->d);
  return fn(*b->f);
}

}  // namespace

int hbdia_batch_create(hbdia_batch_handle_t* b, const hbdia_matrix_t* matrices, int count) {
  return guard("hbdia_batch_create", [&] {
    need(b != nullptr, "null handle pointer");
    const int dt = batch_dtype(matrices, count);
    auto h = std::make_unique<hbdia_batch_opaque>();
    h->dtype = dt;
    if (dt == HBDIA_DTYPE_F64) {
      std::vector<const HBDIA<double>*> v;
      for (int i = 0; i < count; ++i) v.push_back(matrices[i]->d.get());
      h->d = std::make_unique<HBDIABatch<double>>(v);
    } else {
      std::vector<const HBDIA<float>*> v;
      for (int i = 0; i < count; ++i) v.push_back(matrices[i]->f.get());
      h->f = std::make_unique<HBDIABatch<float>>(v);
    }
    *b = h.release();
  });
}

int hbdia_batch_autotune(hbdia_batch_handle_t* b, const hbdia_matrix_t* matrices, int count,
                         const int* block_widths, int num_widths, int threshold, int reps, double* table,
                         int table_capacity, int* table_rows) {
  return guard("hbdia_batch_autotune", [&] {
    need(b != nullptr, "null handle pointer");
    need(block_widths != nullptr && num_widths > 0, "need block widths");
    const int dt = batch_dtype(matrices, count);
    const std::vector<int> bws(block_widths, block_widths + num_widths);
    std::vector<BatchTuneResult> rows;
    auto h = std::make_unique<hbdia_batch_opaque>();
    h->dtype = dt;
    if (dt == HBDIA_DTYPE_F64) {
      std::vector<HBDIA<double>*> v;
      for (int i = 0; i < count; ++i) v.push_back(matrices[i]->d.get());
      h->d = autotuneBatch<double>(v, bws, threshold, reps, &rows);
    } else {
      std::vector<HBDIA<float>*> v;
      for (int i = 0; i < count; ++i) v.push_back(matrices[i]->f.get());
      h->f = autotuneBatch<float>(v, bws, threshold, reps, &rows);
    }
    if (table)
      for (int i = 0; i < std::min<int>(table_capacity, static_cast<int>(rows.size())); ++i) {
        table[5 * i + 0] = rows[i].blockWidth;
        table[5 * i + 1] = rows[i].schedule.kind;
        table[5 * i + 2] = rows[i].schedule.tilesPerStep;
        table[5 * i + 3] = rows[i].schedule.workgroups;
        table[5 * i + 4] = rows[i].ms;
      }
    if (table_rows) *table_rows = static_cast<int>(rows.size());
    *b = h.release();
  });
}

int hbdia_batch_destroy(hbdia_batch_handle_t b) {
  return guard("hbdia_batch_destroy", [&] { delete b; });
}

int hbdia_batch_info(hbdia_batch_handle_t b, int64_t* info) {
  return guard("hbdia_batch_info", [&] {
    need(info != nullptr, "null info");
    with_batch(b, [&](auto& B) {
      info[HBDIA_BATCH_INFO_MATRICES] = B.numMatrices();
      info[HBDIA_BATCH_INFO_ROWS] = B.totalRows();
      info[HBDIA_BATCH_INFO_TILES] = B.numTiles();
      info[HBDIA_BATCH_INFO_BDIA_NNZ] = B.bdiaNnz();
      info[HBDIA_BATCH_INFO_COO_NNZ] = B.cooNnz();
      info[HBDIA_BATCH_INFO_ALG_BYTES] = B.algorithmicBytes();
      info[HBDIA_BATCH_INFO_STORED_BYTES] = B.storedBytes();
      info[HBDIA_BATCH_INFO_FLOPS] = B.flops();
      info[HBDIA_BATCH_INFO_KIND] = B.schedule().kind;
      info[HBDIA_BATCH_INFO_TILES_PER_STEP] = B.schedule().tilesPerStep;
      info[HBDIA_BATCH_INFO_WORKGROUPS] = B.schedule().workgroups;
      info[HBDIA_BATCH_INFO_DTYPE] = b->dtype;
      info[HBDIA_BATCH_INFO_LONG_LISTS] = B.deviceBatch().long_lists;
      return 0;
    });
  });
}

int hbdia_batch_row_offsets(hbdia_batch_handle_t b, int64_t* offsets) {
  return guard("hbdia_batch_row_offsets", [&] {
    need(offsets != nullptr, "null output");
    with_batch(b, [&](auto& B) {
      for (int m = 0; m < B.numMatrices(); ++m) offsets[m] = B.rowOffset(m);
      offsets[B.numMatrices()] = B.totalRows();
      return 0;
    });
  });
}

int hbdia_batch_set_schedule(hbdia_batch_handle_t b, int kind, int tiles_per_step, int workgroups) {
  return guard("hbdia_batch_set_schedule", [&] {
    with_batch(b, [&](auto& B) {
      B.setSchedule(BatchSchedule{kind, tiles_per_step, workgroups});
      return 0;
    });
  });
}

int hbdia_batch_resident_workgroups(hbdia_batch_handle_t b, int kind, int* workgroups) {
  return guard("hbdia_batch_resident_workgroups", [&] {
    need(workgroups != nullptr, "null output");
    with_batch(b, [&](auto& B) {
      *workgroups = B.residentWorkgroups(kind);
      return 0;
    });
  });
}

int hbdia_batch_spmv(hbdia_batch_handle_t b, const void* x, void* y) {
  return guard("hbdia_batch_spmv", [&] {
    need(b != nullptr, "null batch handle");
    if (b->dtype == HBDIA_DTYPE_F64) b->d->spmv(static_cast<const double*>(x), static_cast<double*>(y));
    else b->f->spmv(static_cast<const float*>(x), static_cast<float*>(y));
  });
}

int hbdia_batch_spmv_per_matrix(hbdia_batch_handle_t b, const void* x, void* y) {
  return guard("hbdia_batch_spmv_per_matrix", [&] {
    need(b != nullptr, "null batch handle");
    if (b->dtype == HBDIA_DTYPE_F64) b->d->spmvPerMatrix(static_cast<const double*>(x), static_cast<double*>(y));
    else b->f->spmvPerMatrix(static_cast<const float*>(x), static_cast<float*>(y));
  });
}

int hbdia_batch_capture_graph(hbdia_batch_handle_t b, const void* x, void* y) {
  return guard("hbdia_batch_capture_graph", [&] {
    need(b != nullptr, "null batch handle");
    if (b->dtype == HBDIA_DTYPE_F64) b->d->captureGraph(static_cast<const double*>(x), static_cast<double*>(y));
    else b->f->captureGraph(static_cast<const float*>(x), static_cast<float*>(y));
  });
}

int hbdia_batch_launch_graph(hbdia_batch_handle_t b) {
  return guard("hbdia_batch_launch_graph", [&] {
    with_batch(b, [&](auto& B) {
      B.launchGraph();
      return 0;
    });
  });
}

int hbdia_batch_synchronize(hbdia_batch_handle_t b) {
  return guard("hbdia_batch_synchronize", [&] {
    with_batch(b, [&](auto& B) {
      B.synchronize();
      return 0;
    });
  });
}

int hbdia_batch_stream(hbdia_batch_handle_t b, void** stream) {
  return guard("hbdia_batch_stream", [&] {
    need(stream != nullptr, "null output");
    with_batch(b, [&](auto& B) {
      *stream = static_cast<void*>(B.stream());
      return 0;
    });
  });
}

int hbdia_gen_banded_batch(hbdia_matrix_t* out, int dtype, int count, int min_rows, int max_rows,
                           int min_offsets, int max_offsets, int max_offset, double offband, uint64_t seed) {
  return guard("hbdia_gen_banded_batch", [&] {
    need(out != nullptr && count >= 0, "bad output or count");
    need(dtype == HBDIA_DTYPE_F64 || dtype == HBDIA_DTYPE_F32, "dtype must be F64 or F32");
    hbdia::BandedBatchSpec spec;
    spec.count = count;
    spec.minRows = min_rows;
    spec.maxRows = max_rows;
    spec.minOffsets = min_offsets;
    spec.maxOffsets = max_offsets;
    spec.maxOffset = max_offset;
    spec.offband = offband;
    spec.seed = seed;
    std::vector<std::unique_ptr<hbdia_matrix_opaque>> hs(static_cast<size_t>(count));
    auto fill = [&](auto tag) {
      using T = decltype(tag);
      std::vector<hbdia::CooMatrix<T>> gens;
      hbdia::generateBandedBatch<T>(spec, gens);
      for (int i = 0; i < count; ++i) {
        hs[i] = std::make_unique<hbdia_matrix_opaque>();
        hs[i]->dtype = dtype;
        member<T>(hs[i].get()) = from_gen(gens[i]);
      }
    };
    if (dtype == HBDIA_DTYPE_F64) fill(double{});
    else fill(float{});
    for (int i = 0; i < count; ++i) out[i] = hs[i].release();
  });
}

int hbdia_banded_batch_shape(int i, int min_rows, int max_rows, int min_offsets, int max_offsets, uint64_t seed,
                             int32_t* rows, int32_t* noffsets, uint64_t* matrix_seed) {
  return guard("hbdia_banded_batch_shape", [&] {
    need(rows && noffsets && matrix_seed, "null output");
    hbdia::BandedBatchSpec spec;
    spec.minRows = min_rows;
    spec.maxRows = max_rows;
    spec.minOffsets = min_offsets;
    spec.maxOffsets = max_offsets;
    spec.seed = seed;
    const hbdia::BandedShape s = hbdia::bandedBatchShape(spec, i);
    *rows = s.rows;
    *noffsets = s.noffsets;
    *matrix_seed = s.seed;
  });
}

// ---------------------------------------------------------------------------
// serving

namespace {

template <typename F>
auto with_server(hbdia_server_t s, F&& fn) {
  need(s != nullptr, "null server handle");
  if (s->dtype == HBDIA_DTYPE_F64) return fn(*s->d);
  return fn(*s->f);
}

}  // namespace

int hbdia_server_create(hbdia_server_t* s, const hbdia_matrix_t* matrices, int count, int64_t hbm_budget_bytes,
                        int max_jobs_per_launch, int max_launches_in_flight, int workgroups, int tiles_per_step,
                        int batching) {
  return guard("hbdia_server_create", [&] {
    need(s != nullptr, "null handle pointer");
    const int dt = batch_dtype(matrices, count);
    ServeConfig cfg;
    cfg.hbmBudgetBytes = hbm_budget_bytes;
    cfg.maxJobsPerLaunch = max_jobs_per_launch;
    cfg.maxLaunchesInFlight = max_launches_in_flight;
    cfg.workgroups = workgroups;
    cfg.tilesPerStep = tiles_per_step;
    cfg.batching = batching != 0;
    auto h = std::make_unique<hbdia_server_opaque>();
    h->dtype = dt;
    if (dt == HBDIA_DTYPE_F64) {
      std::vector<const HBDIA<double>*> v;
      for (int i = 0; i < count; ++i) v.push_back(matrices[i]->d.get());
      h->d = std::make_unique<HBDIAServer<double>>(v, cfg);
    } else {
      std::vector<const HBDIA<float>*> v;
      for (int i = 0; i < count; ++i) v.push_back(matrices[i]->f.get());
      h->f = std::make_unique<HBDIAServer<float>>(v, cfg);
    }
    *s = h.release();
  });
}

int hbdia_server_destroy(hbdia_server_t s) {
  return guard("hbdia_server_destroy", [&] { delete s; });
}

int hbdia_server_submit(hbdia_server_t s, int matrix, const void* x, void* y, int device_vectors, int64_t* id) {
  return guard("hbdia_server_submit", [&] {
    need(s != nullptr && id != nullptr, "null argument");
    if (s->dtype == HBDIA_DTYPE_F64)
      *id = s->d->submit(matrix, static_cast<const double*>(x), static_cast<double*>(y), device_vectors != 0);
    else
      *id = s->f->submit(matrix, static_cast<const float*>(x), static_cast<float*>(y), device_vectors != 0);
  });
}

int hbdia_server_wait(hbdia_server_t s, int64_t id) {
  return guard("hbdia_server_wait", [&] {
    with_server(s, [&](auto& S) {
      S.wait(id);
      return 0;
    });
  });
}

int hbdia_server_drain(hbdia_server_t s) {
  return guard("hbdia_server_drain", [&] {
    with_server(s, [&](auto& S) {
      S.drain();
      return 0;
    });
  });
}

int hbdia_server_times(hbdia_server_t s, int64_t id, int64_t* times) {
  return guard("hbdia_server_times", [&] {
    need(times != nullptr, "null output");
    with_server(s, [&](auto& S) {
      const RequestTimes t = S.times(id);
      times[0] = t.submitNs;
      times[1] = t.admitNs;
      times[2] = t.completeNs;
      return 0;
    });
  });
}

int hbdia_server_info(hbdia_server_t s, int64_t* info) {
  return guard("hbdia_server_info", [&] {
    need(info != nullptr, "null output");
    with_server(s, [&](auto& S) {
      info[HBDIA_SERVER_INFO_MATRICES] = S.numMatrices();
      info[HBDIA_SERVER_INFO_RESIDENT_BYTES] = S.residentBytes();
      info[HBDIA_SERVER_INFO_ARENA_BYTES] = S.arenaBytes();
      info[HBDIA_SERVER_INFO_LAUNCHES] = S.launches();
      info[HBDIA_SERVER_INFO_COMPLETED] = S.completed();
      info[HBDIA_SERVER_INFO_LARGEST_LAUNCH] = S.largestLaunch();
      info[HBDIA_SERVER_INFO_DTYPE] = s->dtype;
      return 0;
    });
  });
}

int hbdia_server_num_rows(hbdia_server_t s, int matrix, int64_t* rows) {
  return guard("hbdia_server_num_rows", [&] {
    need(rows != nullptr, "null output");
    with_server(s, [&](auto& S) {
      need(matrix >= 0 && matrix < S.numMatrices(), "no such matrix");
      *rows = S.numRows(matrix);
      return 0;
    });
  });
}

int hbdia_server_open_loop(hbdia_server_t s, double rate, int64_t requests, uint64_t seed, const void* const* xs,
                           int variants, void* const* ys, int device_vectors, double* latencies_us,
                           double* wall_seconds) {
  return guard("hbdia_server_open_loop", [&] {
    need(s != nullptr && xs != nullptr && ys != nullptr && variants > 0, "bad argument");
    need(requests >= 0 && (requests == 0 || latencies_us != nullptr), "bad request count or output");
    LoadReport rep;
    const int M = s->dtype == HBDIA_DTYPE_F64 ? s->d->numMatrices() : s->f->numMatrices();
    if (s->dtype == HBDIA_DTYPE_F64) {
      std::vector<const double*> x(static_cast<size_t>(M) * variants);
      std::vector<double*> y(static_cast<size_t>(M));
      for (size_t i = 0; i < x.size(); ++i) x[i] = static_cast<const double*>(xs[i]);
      for (int m = 0; m < M; ++m) y[m] = static_cast<double*>(ys[m]);
      rep = runOpenLoop<double>(*s->d, rate, requests, seed, x, variants, y, device_vectors != 0);
    } else {
      std::vector<const float*> x(static_cast<size_t>(M) * variants);
      std::vector<float*> y(static_cast<size_t>(M));
      for (size_t i = 0; i < x.size(); ++i) x[i] = static_cast<const float*>(xs[i]);
      for (int m = 0; m < M; ++m) y[m] = static_cast<float*>(ys[m]);
      rep = runOpenLoop<float>(*s->f, rate, requests, seed, x, variants, y, device_vectors != 0);
    }
    for (size_t i = 0; i < rep.latencyUs.size(); ++i) latencies_us[i] = rep.latencyUs[i];
    if (wall_seconds) *wall_seconds = rep.wallSeconds;
  });
}

int hbdia_arena_create(hbdia_arena_t* a, int64_t capacity, int64_t align) {
  return guard("hbdia_arena_create", [&] {
    need(a != nullptr, "null handle pointer");
    auto h = std::make_unique<hbdia_arena_opaque>();
    h->arena = hbdia::VectorArena(capacity, align);
    *a = h.release();
  });
}

int hbdia_arena_destroy(hbdia_arena_t a) {
  return guard("hbdia_arena_destroy", [&] { delete a; });
}

int hbdia_arena_allocate(hbdia_arena_t a, int64_t n, int64_t* offset) {
  return guard("hbdia_arena_allocate", [&] {
    need(a != nullptr && offset != nullptr, "null argument");
    *offset = a->arena.allocate(n);
  });
}

int hbdia_arena_release(hbdia_arena_t a, int64_t offset) {
  return guard("hbdia_arena_release", [&] {
    need(a != nullptr, "null argument");
    a->arena.release(offset);
  });
}

int hbdia_serve_arena_elems(int64_t budget_bytes, int64_t resident_bytes, int64_t elem_bytes, int64_t* elems) {
  return guard("hbdia_serve_arena_elems", [&] {
    need(elems != nullptr, "null argument");
    *elems = hbdia::serveArenaElems(budget_bytes, resident_bytes, elem_bytes);
  });
}

int hbdia_arena_info(hbdia_arena_t a, int64_t* info) {
  return guard("hbdia_arena_info", [&] {
    need(a != nullptr && info != nullptr, "null argument");
    info[0] = a->arena.capacity();
    info[1] = a->arena.used();
    info[2] = a->arena.largestFree();
    info[3] = a->arena.freeBlocks();
  });
}
