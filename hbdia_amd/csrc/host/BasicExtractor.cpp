// BasicExtractor.cpp -- row-wise partitioning and the one-GPU halo fill
// (reference src/DataExchange/BasicExtractor.cpp).
#include "DataExchange/BasicExtractor.hpp"

#include <algorithm>
#include <iostream>
#include <stdexcept>

template <typename T>
bool BasicExtractor<T>::createMatrixPartitions(const HBDIA<T>& matrix, int numProcesses) {
  if (numProcesses <= 0) {
    std::cerr << "Invalid number of processes" << std::endl;
    return false;
  }
  switch (this->strategy_) {
    case HBDIAExtractor<T>::PartitioningStrategy::ROW_WISE:
      return partitionRowWise(matrix, numProcesses);
  }
  return false;
}

template <typename T>
bool BasicExtractor<T>::createVectorPartitions(const std::vector<T>& globalVector, int numProcesses) {
  if (numProcesses <= 0 || globalVector.empty()) {
    std::cerr << "Invalid vector partition request" << std::endl;
    return false;
  }
  return partitionVectorRowWise(globalVector, numProcesses);
}

template <typename T>
void BasicExtractor<T>::setPartitioningStrategy(typename HBDIAExtractor<T>::PartitioningStrategy strategy) {
  this->strategy_ = strategy;
}

// Reference BasicExtractor.cpp:125-234: partition p owns rows
// [p*q, p*q + q) with q = numRows / numPartitions, the last partition also
// the remainder; every entry goes, in input order, to the owner of its row
// with a local row index and its global column.
template <typename T>
bool BasicExtractor<T>::partitionRowWise(const HBDIA<T>& matrix, int numPartitions) {
  const int numRows = matrix.getNumRows();
  const int numCols = matrix.getNumCols();
  const int nnz = matrix.getNumNonZeros();
  if (numRows == 0 || numCols == 0 || nnz == 0) {
    std::cerr << "Invalid matrix dimensions" << std::endl;
    return false;
  }
  if (!matrix.getHasCOO()) {
    std::cerr << "Matrix does not have COO format available" << std::endl;
    return false;
  }
  const int q = numRows / numPartitions;
  const int rem = numRows % numPartitions;
  if (q == 0) {
    std::cerr << "Fewer rows than partitions" << std::endl;
    return false;
  }
  this->matrixData_.assign(numPartitions, MatrixData{});
  this->partitions_.assign(numPartitions, Partition<T>{});
  for (int p = 0; p < numPartitions; ++p) {
    MatrixData& md = this->matrixData_[p];
    md.numGlobalRows = numRows;
    md.numGlobalCols = numCols;
    md.numGlobalNonZeros = nnz;
    md.numLocalRows = q + (p == numPartitions - 1 ? rem : 0);
    md.numLocalCols = numCols;
    auto& g = this->partitions_[p].globalRowMapping;
    g.resize(md.numLocalRows);
    for (int i = 0; i < md.numLocalRows; ++i) g[i] = p * q + i;
  }
  const auto& rows = matrix.getRowIndices();
  const auto& cols = matrix.getColIndices();
  const auto& vals = matrix.getValues();
  std::vector<int64_t> count(numPartitions, 0);
  for (int i = 0; i < nnz; ++i) count[std::min(rows[i] / q, numPartitions - 1)]++;
  for (int p = 0; p < numPartitions; ++p) {
    auto& part = this->partitions_[p];
    part.localRowIndices.reserve(count[p]);
    part.localColIndices.reserve(count[p]);
    part.localValues.reserve(count[p]);
  }
  for (int i = 0; i < nnz; ++i) {
    const int p = std::min(rows[i] / q, numPartitions - 1);
    auto& part = this->partitions_[p];
    part.localRowIndices.push_back(rows[i] - p * q);
    part.localColIndices.push_back(cols[i]);
    part.localValues.push_back(vals[i]);
    this->matrixData_[p].numLocalNonZeros++;
  }
  return true;
}

// Reference BasicExtractor.cpp:75-122: contiguous slices of the same sizes.
template <typename T>
bool BasicExtractor<T>::partitionVectorRowWise(const std::vector<T>& globalVector, int numPartitions) {
  const int64_t n = static_cast<int64_t>(globalVector.size());
  const int64_t q = n / numPartitions, rem = n % numPartitions;
  this->vectorPartitions_.assign(numPartitions, VectorPartition<T>{});
  int64_t at = 0;
  for (int p = 0; p < numPartitions; ++p) {
    const int64_t len = q + (p == numPartitions - 1 ? rem : 0);
    this->vectorPartitions_[p].localValues.assign(globalVector.begin() + at, globalVector.begin() + at + len);
    at += len;
  }
  return true;
}

template <typename T>
void BasicExtractor<T>::extractPartialMatrixMetadata(HBDIA<T>& matrix, int numProcesses) {
  switch (this->strategy_) {
    case HBDIAExtractor<T>::PartitioningStrategy::ROW_WISE:
      extractPartialMatrixMetadataRowWise(matrix, numProcesses);
      return;
  }
}

// Reference BasicExtractor.cpp:269-333: each data range is split at the
// ownership boundaries (p + 1) * q; the last process takes the rest.
template <typename T>
void BasicExtractor<T>::extractPartialMatrixMetadataRowWise(HBDIA<T>& matrix, int numProcesses) {
  if (!matrix.isPartialMatrix()) return;
  const int q = numProcesses > 0 ? matrix.getNumGlobalRows() / numProcesses : 0;
  if (q <= 0) throw std::invalid_argument("extractPartialMatrixMetadata: fewer global rows than processes");
  PartialMatrixMetadata<T> md;
  md.dataRanges = matrix.getDataRanges();
  md.processDataRanges.resize(numProcesses);
  for (const auto& rg : md.dataRanges) {
    const int rs = std::get<0>(rg), re = std::get<1>(rg);
    md.totalBufferSize += re - rs;
    const int p0 = std::max(0, std::min(rs / q, numProcesses - 1));
    const int p1 = std::max(0, std::min((re - 1) / q, numProcesses - 1));
    int at = rs;
    for (int p = p0; p <= p1 && at < re; ++p) {
      const int boundary = p == numProcesses - 1 ? re : (p + 1) * q;
      const int end = std::min(re, boundary);
      if (at < end) {
        md.processDataRanges[p].emplace_back(at, end);
        at = end;
      }
    }
  }
  matrix.setPartialMatrixMetadata(md);
}

template <typename T>
void BasicExtractor<T>::printProcessedDataRanges(const HBDIA<T>& matrix) {
  std::cout << "=== PROCESSED DATA RANGES ===" << std::endl;
  if (!matrix.isPartialMatrix() || !matrix.hasPartialMatrixMetadata()) {
    std::cout << "No processed metadata available" << std::endl;
    return;
  }
  const auto& md = matrix.getPartialMatrixMetadata();
  std::cout << "total buffer size: " << md.totalBufferSize << std::endl;
  for (size_t p = 0; p < md.processDataRanges.size(); ++p) {
    if (md.processDataRanges[p].empty()) continue;
    std::cout << "process " << p << ": " << md.processDataRanges[p].size() << " range(s):";
    for (const auto& rg : md.processDataRanges[p]) std::cout << " [" << std::get<0>(rg) << ", " << std::get<1>(rg) << ")";
    std::cout << std::endl;
  }
}

template <typename T>
void BasicExtractor<T>::clearMatrixPartitions() {
  this->matrixData_.clear();
  this->partitions_.clear();
}

template <typename T>
void BasicExtractor<T>::clearVectorPartitions() {
  this->vectorPartitions_.clear();
}

template <typename T>
HBDIA<T> createRankLocalMatrix(const HBDIA<T>& globalMatrix, int rank, int size, int blockWidth, int threshold,
                               int maxCooEntries, ExecutionMode execMode) {
  if (size <= 0 || rank < 0 || rank >= size) throw std::invalid_argument("createRankLocalMatrix: bad rank/size");
  if (!globalMatrix.getHasCOO() || globalMatrix.getNumNonZeros() == 0)
    throw std::invalid_argument("createRankLocalMatrix: the matrix has no COO entries");
  // partitionRowWise's partition `rank`, without building the other ones
  const int numRows = globalMatrix.getNumRows();
  const int q = numRows / size;
  if (q == 0) throw std::invalid_argument("createRankLocalMatrix: fewer rows than ranks");
  const int r0 = rank * q;
  const int n = q + (rank == size - 1 ? numRows % size : 0);
  const auto& rows = globalMatrix.getRowIndices();
  const auto& cols = globalMatrix.getColIndices();
  const auto& vals = globalMatrix.getValues();
  std::vector<int> lr, lc, grm(n);
  std::vector<T> lv;
  for (size_t i = 0; i < rows.size(); ++i) {
    if (rows[i] < r0 || rows[i] >= r0 + n) continue;
    lr.push_back(rows[i] - r0);
    lc.push_back(cols[i]);
    lv.push_back(vals[i]);
  }
  for (int i = 0; i < n; ++i) grm[i] = r0 + i;
  HBDIA<T> local(lr, lc, lv, n, globalMatrix.getNumCols(), grm, true, numRows, globalMatrix.getNumCols(),
                 globalMatrix.getNumNonZeros(), rank, size);
  BasicExtractor<T> ex;
  if (!local.convertToHBDIAFormatOnHost(blockWidth, threshold, maxCooEntries, false, execMode))
    throw std::runtime_error("createRankLocalMatrix: conversion failed");
  local.analyzeDataRanges();
  ex.extractPartialMatrixMetadata(local, size);
  local.calculateVectorOffsets(rank, size);
  return local;
}

template <typename T>
bool exchangeHaloFromGlobal(const HBDIA<T>& matrix, HBDIAVector<T>& vector, const std::vector<T>& globalVector) {
  if (!matrix.hasLocalCOO()) {
    std::cerr << "exchangeHaloFromGlobal: the matrix has no rank-local layout (call prepareForGPU)" << std::endl;
    return false;
  }
  const size_t left = static_cast<size_t>(matrix.getHaloLeftSize());
  const size_t right = static_cast<size_t>(matrix.getHaloRightSize());
  if (vector.getTotalSize() - vector.getLocalSize() != left + right ||
      vector.getDeviceLocalPtr() - vector.getDeviceDataPtr() != static_cast<std::ptrdiff_t>(left)) {
    std::cerr << "exchangeHaloFromGlobal: vector halo sizes do not match the matrix" << std::endl;
    return false;
  }
  const auto& cols = matrix.getHaloColumns();
  std::vector<T> l(left), r(right);
  for (size_t i = 0; i < left; ++i) l[i] = globalVector.at(static_cast<size_t>(cols[i]));
  for (size_t i = 0; i < right; ++i) r[i] = globalVector.at(static_cast<size_t>(cols[left + i]));
  vector.uploadHalo(l, r);
  return true;
}

template <typename T>
bool verifyHaloExchange(const HBDIA<T>& matrix, const HBDIAVector<T>& vector) {
  if (!matrix.isPartialMatrix() || !matrix.hasPartialMatrixMetadata()) {
    std::cerr << "Matrix must be a partial matrix with metadata for verification" << std::endl;
    return false;
  }
  const auto& pdr = matrix.getPartialMatrixMetadata().processDataRanges;
  const auto& grm = matrix.getGlobalRowMapping();
  const std::vector<T> all = vector.downloadAll();
  int64_t at = 0, errors = 0;
  auto check = [&](int64_t pos, int64_t global) {
    if (pos >= static_cast<int64_t>(all.size()) || all[pos] != static_cast<T>(global)) {
      if (errors < 8)
        std::cerr << "Verification ERROR: rank " << matrix.getRank() << " expected " << global << " at buffer position "
                  << pos << std::endl;
      ++errors;
    }
  };
  for (int p = 0; p < matrix.getSize() && p < static_cast<int>(pdr.size()); ++p) {
    if (p == matrix.getRank()) {
      for (size_t i = 0; i < grm.size(); ++i) check(at + static_cast<int64_t>(i), grm[i]);
      at += static_cast<int64_t>(grm.size());
      continue;
    }
    for (const auto& rg : pdr[p])
      for (int g = std::get<0>(rg); g < std::get<1>(rg); ++g) check(at++, g);
  }
  if (at != static_cast<int64_t>(all.size())) ++errors;
  if (errors == 0) std::cout << "Data exchange verification PASSED for rank " << matrix.getRank() << std::endl;
  else
    std::cerr << "Data exchange verification FAILED for rank " << matrix.getRank() << " with " << errors << " errors"
              << std::endl;
  return errors == 0;
}

template class BasicExtractor<double>;
template class BasicExtractor<float>;
template HBDIA<double> createRankLocalMatrix<double>(const HBDIA<double>&, int, int, int, int, int, ExecutionMode);
template HBDIA<float> createRankLocalMatrix<float>(const HBDIA<float>&, int, int, int, int, int, ExecutionMode);
template bool exchangeHaloFromGlobal<double>(const HBDIA<double>&, HBDIAVector<double>&, const std::vector<double>&);
template bool exchangeHaloFromGlobal<float>(const HBDIA<float>&, HBDIAVector<float>&, const std::vector<float>&);
template bool verifyHaloExchange<double>(const HBDIA<double>&, const HBDIAVector<double>&);
template bool verifyHaloExchange<float>(const HBDIA<float>&, const HBDIAVector<float>&);
