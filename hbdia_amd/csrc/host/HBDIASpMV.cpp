// HBDIASpMV.cpp -- orchestration of y = A x (reference src/Operations/HBDIASpMV.cpp).
#include "Operations/HBDIASpMV.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>

namespace {

template <typename T>
struct Dev;

template <>
struct Dev<double> {
  static int bdia(const HBDIA<double>& m, const double* x, double* y, hbdia_stream_t s) {
    return hbdia_bdia_spmv_ex_f64(m.getHBDIADataDevice(), m.getFlattenedOffsetsDevice(),
                                  m.getBlockStartIndicesDevice(), x, y, m.getNumRows(), m.getNumBlocks(),
                                  m.getBlockWidth(), m.getRuns3BlocksDevice(), s);
  }
  static int fused(const HBDIA<double>& m, const double* x, double* y, hbdia_stream_t s) {
    return hbdia_hbdia_spmv_ex_f64(m.getHBDIADataDevice(), m.getFlattenedOffsetsDevice(),
                                   m.getBlockStartIndicesDevice(), m.getNumBlocks(), m.getBlockWidth(),
                                   m.getCpuRowIndicesDevice(), m.getCpuColIndicesDevice(), m.getCpuValuesDevice(),
                                   m.getCooGroupPtrDevice(), x, y, m.getNumRows(), m.getRuns3BlocksDevice(), s);
  }
  static int coo(const HBDIA<double>& m, const double* x, double* res, hbdia_stream_t s) {
    return hbdia_coo_spmv_f64(m.getCpuRowIndicesDevice(), m.getCpuColIndicesDevice(), m.getCpuValuesDevice(),
                              m.getCooGroupPtrDevice(), x, res, m.getNumRows(), s);
  }
  static int add(double* d, const double* s_, int64_t n, hbdia_stream_t s) {
    return hbdia_vector_add_f64(d, s_, n, s);
  }
  static double fma(double a, double b, double c) { return std::fma(a, b, c); }
  static int multi_rm(const HBDIA<double>& m, bool coo, const double* X, double* Y, int k, hbdia_stream_t s) {
    return hbdia_hbdia_spmm_rm_f64(m.getHBDIADataDevice(), m.getFlattenedOffsetsDevice(),
                                   m.getBlockStartIndicesDevice(), m.getNumBlocks(), m.getBlockWidth(),
                                   coo ? m.getCpuRowIndicesDevice() : nullptr, coo ? m.getCpuColIndicesDevice() : nullptr,
                                   coo ? m.getCpuValuesDevice() : nullptr, coo ? m.getCooGroupPtrDevice() : nullptr, X,
                                   Y, k, m.getNumRows(), s);
  }
  static int multi(const HBDIA<double>& m, bool coo, const double* X, int64_t ldx, double* Y, int64_t ldy, int k,
                   hbdia_stream_t s) {
    return hbdia_hbdia_spmm_ex_f64(m.getHBDIADataDevice(), m.getFlattenedOffsetsDevice(),
                                   m.getBlockStartIndicesDevice(), m.getNumBlocks(), m.getBlockWidth(),
                                   coo ? m.getCpuRowIndicesDevice() : nullptr, coo ? m.getCpuColIndicesDevice() : nullptr,
                                   coo ? m.getCpuValuesDevice() : nullptr, coo ? m.getCooGroupPtrDevice() : nullptr, X,
                                   ldx, Y, ldy, k, m.getNumRows(), m.getRuns3BlocksDevice(), s);
  }
};

template <>
struct Dev<float> {
  static int bdia(const HBDIA<float>& m, const float* x, float* y, hbdia_stream_t s) {
    return hbdia_bdia_spmv_ex_f32(m.getHBDIADataDevice(), m.getFlattenedOffsetsDevice(),
                                  m.getBlockStartIndicesDevice(), x, y, m.getNumRows(), m.getNumBlocks(),
                                  m.getBlockWidth(), m.getRuns3BlocksDevice(), s);
  }
  static int fused(const HBDIA<float>& m, const float* x, float* y, hbdia_stream_t s) {
    return hbdia_hbdia_spmv_ex_f32(m.getHBDIADataDevice(), m.getFlattenedOffsetsDevice(),
                                   m.getBlockStartIndicesDevice(), m.getNumBlocks(), m.getBlockWidth(),
                                   m.getCpuRowIndicesDevice(), m.getCpuColIndicesDevice(), m.getCpuValuesDevice(),
                                   m.getCooGroupPtrDevice(), x, y, m.getNumRows(), m.getRuns3BlocksDevice(), s);
  }
  static int coo(const HBDIA<float>& m, const float* x, float* res, hbdia_stream_t s) {
    return hbdia_coo_spmv_f32(m.getCpuRowIndicesDevice(), m.getCpuColIndicesDevice(), m.getCpuValuesDevice(),
                              m.getCooGroupPtrDevice(), x, res, m.getNumRows(), s);
  }
  static int add(float* d, const float* s_, int64_t n, hbdia_stream_t s) {
    return hbdia_vector_add_f32(d, s_, n, s);
  }
  static float fma(float a, float b, float c) { return std::fmaf(a, b, c); }
  static int multi_rm(const HBDIA<float>& m, bool coo, const float* X, float* Y, int k, hbdia_stream_t s) {
    return hbdia_hbdia_spmm_rm_f32(m.getHBDIADataDevice(), m.getFlattenedOffsetsDevice(),
                                   m.getBlockStartIndicesDevice(), m.getNumBlocks(), m.getBlockWidth(),
                                   coo ? m.getCpuRowIndicesDevice() : nullptr, coo ? m.getCpuColIndicesDevice() : nullptr,
                                   coo ? m.getCpuValuesDevice() : nullptr, coo ? m.getCooGroupPtrDevice() : nullptr, X,
                                   Y, k, m.getNumRows(), s);
  }
  static int multi(const HBDIA<float>& m, bool coo, const float* X, int64_t ldx, float* Y, int64_t ldy, int k,
                   hbdia_stream_t s) {
    return hbdia_hbdia_spmm_ex_f32(m.getHBDIADataDevice(), m.getFlattenedOffsetsDevice(),
                                   m.getBlockStartIndicesDevice(), m.getNumBlocks(), m.getBlockWidth(),
                                   coo ? m.getCpuRowIndicesDevice() : nullptr, coo ? m.getCpuColIndicesDevice() : nullptr,
                                   coo ? m.getCpuValuesDevice() : nullptr, coo ? m.getCooGroupPtrDevice() : nullptr, X,
                                   ldx, Y, ldy, k, m.getNumRows(), m.getRuns3BlocksDevice(), s);
  }
};

template <typename T>
void check_ready(const HBDIA<T>& m, const HBDIAVector<T>& x, const HBDIAVector<T>& y) {
  if (!m.isHBDIAFormat())
    throw std::runtime_error("hbdiaSpMV: matrix is not in HBDIA format (call convertToHBDIAFormat)");
  if (!m.areStreamsInitialized())
    throw std::runtime_error("hbdiaSpMV: matrix is not on the device (call prepareForGPU)");
  if (y.getLocalSize() < static_cast<size_t>(m.getNumRows()))
    throw std::invalid_argument("hbdiaSpMV: output vector shorter than the number of rows");
  if (x.getLocalSize() < static_cast<size_t>(m.getNumRows()))
    throw std::invalid_argument("hbdiaSpMV: input vector shorter than the number of rows");
  // The fallback gathers x[col] for every col < numCols from the data pointer
  // (left halo included), so a rectangular matrix needs numCols entries there.
  if (!m.isPartialMatrix() && x.getTotalSize() < static_cast<size_t>(m.getNumCols()))
    throw std::invalid_argument("hbdiaSpMV: input vector shorter than the number of columns");
  if (m.isPartialMatrix()) {
    if (!m.hasLocalCOO()) throw std::runtime_error("hbdiaSpMV: partial matrix without its rank-local layout");
    if (x.getLeftSize() != static_cast<size_t>(m.getHaloLeftSize()) ||
        x.getRightSize() != static_cast<size_t>(m.getHaloRightSize()))
      throw std::invalid_argument("hbdiaSpMV: input vector halo sizes do not match the matrix's halo layout");
  }
}

// x as the kernels index it: the local segment for a rank-local matrix (its
// COO columns are relative to it, negative in the left halo), else the data
// pointer, as in the reference (HBDIASpMV.cpp:212).
template <typename T>
const T* device_x(const HBDIA<T>& m, const HBDIAVector<T>& x) {
  return m.isPartialMatrix() ? x.getDeviceLocalPtr() : x.getDeviceDataPtr();
}

}  // namespace

// Reference HBDIASpMV.cpp:35-99.
template <typename T>
void hbdia_cpu_coo_spmv(const std::vector<int>& cpuRowIndices, const std::vector<int>& cpuColIndices,
                        const std::vector<T>& cpuValues, const T* inputVector, int numRows, T* cpuResults) {
  const size_t nnz = cpuRowIndices.size();
  if (nnz == 0) return;
#pragma omp parallel
  {
    const int nt = omp_get_num_threads(), tid = omp_get_thread_num();
    const int rowsPerThread = (numRows + nt - 1) / nt;
    const int startRow = tid * rowsPerThread;
    const int endRow = std::min(startRow + rowsPerThread, numRows);
    size_t s = 0, e = nnz;
    if (startRow > 0)
      s = std::lower_bound(cpuRowIndices.begin(), cpuRowIndices.end(), startRow) - cpuRowIndices.begin();
    if (endRow < numRows)
      e = std::lower_bound(cpuRowIndices.begin(), cpuRowIndices.end(), endRow) - cpuRowIndices.begin();
    T rowSum = T(0);
    int cur = -1;
    for (size_t i = s; i < e; ++i) {
      const int row = cpuRowIndices[i];
      if (row != cur) {
        if (cur >= startRow && cur < endRow) cpuResults[cur] = rowSum;
        cur = row;
        rowSum = T(0);
      }
      rowSum = Dev<T>::fma(cpuValues[i], inputVector[cpuColIndices[i]], rowSum);
    }
    if (cur >= startRow && cur < endRow) cpuResults[cur] = rowSum;
  }
}

// Reference HBDIASpMV.cpp:101-131.
template <typename T>
void executeCOOOnCPU(HBDIA<T>& matrix, const HBDIAVector<T>& inputVector, HBDIAVector<T>& outputVector,
                     int numRows, hbdia_stream_t sD2H, hbdia_stream_t sH2D, hbdia_stream_t /*sCOO*/,
                     hbdia_event_t cooEvent) {
  const bool local = matrix.isPartialMatrix();
  const auto& rows = local ? matrix.getLocalCOORowIndices() : matrix.getCpuRowIndices();
  const auto& cols = local ? matrix.getLocalCOOColIndices() : matrix.getCpuColIndices();
  const auto& vals = local ? matrix.getLocalCOOValues() : matrix.getCpuValues();
  if (outputVector.isUnifiedMemory()) {
    hbdia_cpu_coo_spmv<T>(rows, cols, vals, inputVector.getDeviceLocalPtr(), numRows,
                          outputVector.getDeviceCOOResultsPtr());
    return;
  }
  CHECK_HBDIA(hbdia_memcpy_d2h(inputVector.getHostDataPtr(), inputVector.getDeviceDataPtr(),
                               inputVector.getTotalSize() * sizeof(T), sD2H));
  CHECK_HBDIA(hbdia_stream_synchronize(sD2H));
  hbdia_cpu_coo_spmv<T>(rows, cols, vals, inputVector.getHostLocalPtr(), numRows,
                        outputVector.getHostCOOResultsPtr());
  CHECK_HBDIA(hbdia_memcpy_h2d(outputVector.getDeviceCOOResultsPtr(), outputVector.getHostCOOResultsPtr(),
                               static_cast<size_t>(numRows) * sizeof(T), sH2D));
  CHECK_HBDIA(hbdia_event_record(cooEvent, sH2D));
}

// Reference HBDIASpMV.cpp:133-183 (cuSPARSE with beta = 0 there: every row of
// the COO buffer is written, rows without entries get 0).
template <typename T>
void executeCOOOnGPU(HBDIA<T>& matrix, const HBDIAVector<T>& inputVector, HBDIAVector<T>& outputVector,
                     int numRows, hbdia_stream_t sCOO, hbdia_event_t cooEvent) {
  if (matrix.numSpMVCOOEntries() == 0) return;
  if (!matrix.hasDeviceCOO())
    throw std::runtime_error("executeCOOOnGPU: COO fallback not on the device (convert with ExecutionMode::GPU_COO)");
  CHECK_HBDIA(hbdia_memset(outputVector.getDeviceCOOResultsPtr(), 0, static_cast<size_t>(numRows) * sizeof(T), sCOO));
  CHECK_HBDIA(Dev<T>::coo(matrix, device_x(matrix, inputVector), outputVector.getDeviceCOOResultsPtr(), sCOO));
  CHECK_HBDIA(hbdia_event_record(cooEvent, sCOO));
}

// Reference HBDIASpMV.cpp:186-251.
template <typename T>
bool hbdiaSpMV(HBDIA<T>& matrix, const HBDIAVector<T>& inputVector, HBDIAVector<T>& outputVector,
               bool execCOOCPU, bool execCOOGPU) {
  check_ready(matrix, inputVector, outputVector);
  const int numRows = matrix.getNumRows();
  hbdia_stream_t sBDIA = matrix.getBDIAStream();
  const bool hasCOOWork = matrix.numSpMVCOOEntries() > 0;
  const T* x = device_x(matrix, inputVector);
  T* y = outputVector.getDeviceLocalPtr();

  if (hasCOOWork && !execCOOCPU && execCOOGPU) {
    if (!matrix.hasDeviceCOO())
      throw std::runtime_error("hbdiaSpMV: GPU COO requested but the fallback is not on the device "
                               "(convert with ExecutionMode::GPU_COO)");
    CHECK_HBDIA(Dev<T>::fused(matrix, x, y, sBDIA));
    CHECK_HBDIA(hbdia_event_record(matrix.getBDIAEvent(), sBDIA));
    return true;
  }

  CHECK_HBDIA(Dev<T>::bdia(matrix, x, y, sBDIA));
  CHECK_HBDIA(hbdia_event_record(matrix.getBDIAEvent(), sBDIA));
  if (!hasCOOWork) return true;
  if (!execCOOCPU) {
    std::cerr << "COO data is available but execCOOCPU and execCOOGPU are both false. No COO execution "
                 "will be performed."
              << std::endl;
    throw std::runtime_error("Invalid COO execution flags");
  }
  executeCOOOnCPU<T>(matrix, inputVector, outputVector, numRows, matrix.getD2HStream(), matrix.getH2DStream(),
                     matrix.getCOOStream(), matrix.getCOOEvent());
  hbdia_stream_t sADD = matrix.getADDStream();
  CHECK_HBDIA(hbdia_stream_wait_event(sADD, matrix.getBDIAEvent()));
  if (!outputVector.isUnifiedMemory()) CHECK_HBDIA(hbdia_stream_wait_event(sADD, matrix.getCOOEvent()));
  CHECK_HBDIA(Dev<T>::add(y, outputVector.getDeviceCOOResultsPtr(), numRows, sADD));
  return true;
}

template <typename T>
bool hbdiaSpMVMulti(HBDIA<T>& matrix, const T* X, int64_t ldx, T* Y, int64_t ldy, int nrhs, bool execCOOGPU) {
  if (!matrix.isHBDIAFormat() || !matrix.areStreamsInitialized())
    throw std::runtime_error("hbdiaSpMVMulti: matrix is not converted and on the device");
  if (matrix.isPartialMatrix())
    throw std::invalid_argument("hbdiaSpMVMulti: rank-local (partial) matrices take hbdiaSpMV with halo vectors");
  const bool hasCOOWork = !matrix.getCpuValues().empty();
  if (hasCOOWork && !execCOOGPU) throw std::runtime_error("Invalid COO execution flags");
  if (nrhs < 1) throw std::invalid_argument("hbdiaSpMVMulti: nrhs must be at least 1");
  // Blocked columns are clamped into [0, numRows); fallback columns reach numCols - 1.
  const int64_t xNeed = hasCOOWork ? std::max(matrix.getNumRows(), matrix.getNumCols()) : matrix.getNumRows();
  if (ldx < xNeed || ldy < matrix.getNumRows())
    throw std::invalid_argument("hbdiaSpMVMulti: ldx must cover the columns and ldy the rows");
  if (hasCOOWork && !matrix.hasDeviceCOO())
    throw std::runtime_error("hbdiaSpMVMulti: the fallback is not on the device (convert with ExecutionMode::GPU_COO)");
  if (nrhs == 1) {
    // one column: the single-vector kernel (same arithmetic, and its runs-of-three x loads)
    if (hasCOOWork) CHECK_HBDIA(Dev<T>::fused(matrix, X, Y, matrix.getBDIAStream()));
    else CHECK_HBDIA(Dev<T>::bdia(matrix, X, Y, matrix.getBDIAStream()));
  } else {
    CHECK_HBDIA(Dev<T>::multi(matrix, hasCOOWork, X, ldx, Y, ldy, nrhs, matrix.getBDIAStream()));
  }
  CHECK_HBDIA(hbdia_event_record(matrix.getBDIAEvent(), matrix.getBDIAStream()));
  return true;
}

template <typename T>
bool hbdiaSpMVMultiRowMajor(HBDIA<T>& matrix, const T* X, T* Y, int nrhs, bool execCOOGPU) {
  if (!matrix.isHBDIAFormat() || !matrix.areStreamsInitialized())
    throw std::runtime_error("hbdiaSpMVMultiRowMajor: matrix is not converted and on the device");
  if (matrix.isPartialMatrix())
    throw std::invalid_argument("hbdiaSpMVMultiRowMajor: rank-local (partial) matrices take hbdiaSpMV with halo vectors");
  if (nrhs < 1) throw std::invalid_argument("hbdiaSpMVMultiRowMajor: nrhs must be at least 1");
  const bool hasCOOWork = !matrix.getCpuValues().empty();
  if (hasCOOWork && !execCOOGPU) throw std::runtime_error("Invalid COO execution flags");
  if (hasCOOWork && !matrix.hasDeviceCOO())
    throw std::runtime_error("hbdiaSpMVMultiRowMajor: the fallback is not on the device (convert with ExecutionMode::GPU_COO)");
  hbdia_stream_t s = matrix.getBDIAStream();
  if (nrhs == 1) {
    if (hasCOOWork) CHECK_HBDIA(Dev<T>::fused(matrix, X, Y, s));
    else CHECK_HBDIA(Dev<T>::bdia(matrix, X, Y, s));
  } else {
    CHECK_HBDIA(Dev<T>::multi_rm(matrix, hasCOOWork, X, Y, nrhs, s));
  }
  CHECK_HBDIA(hbdia_event_record(matrix.getBDIAEvent(), s));
  return true;
}

template bool hbdiaSpMVMultiRowMajor<float>(HBDIA<float>&, const float*, float*, int, bool);
template bool hbdiaSpMVMultiRowMajor<double>(HBDIA<double>&, const double*, double*, int, bool);
template bool hbdiaSpMVMulti<float>(HBDIA<float>&, const float*, int64_t, float*, int64_t, int, bool);
template bool hbdiaSpMVMulti<double>(HBDIA<double>&, const double*, int64_t, double*, int64_t, int, bool);
template bool hbdiaSpMV<float>(HBDIA<float>&, const HBDIAVector<float>&, HBDIAVector<float>&, bool, bool);
template bool hbdiaSpMV<double>(HBDIA<double>&, const HBDIAVector<double>&, HBDIAVector<double>&, bool, bool);
template void executeCOOOnCPU<float>(HBDIA<float>&, const HBDIAVector<float>&, HBDIAVector<float>&, int,
                                     hbdia_stream_t, hbdia_stream_t, hbdia_stream_t, hbdia_event_t);
template void executeCOOOnCPU<double>(HBDIA<double>&, const HBDIAVector<double>&, HBDIAVector<double>&, int,
                                      hbdia_stream_t, hbdia_stream_t, hbdia_stream_t, hbdia_event_t);
template void executeCOOOnGPU<float>(HBDIA<float>&, const HBDIAVector<float>&, HBDIAVector<float>&, int,
                                     hbdia_stream_t, hbdia_event_t);
template void executeCOOOnGPU<double>(HBDIA<double>&, const HBDIAVector<double>&, HBDIAVector<double>&, int,
                                      hbdia_stream_t, hbdia_event_t);
template void hbdia_cpu_coo_spmv<float>(const std::vector<int>&, const std::vector<int>&,
                                        const std::vector<float>&, const float*, int, float*);
template void hbdia_cpu_coo_spmv<double>(const std::vector<int>&, const std::vector<int>&,
                                         const std::vector<double>&, const double*, int, double*);
