// HBDIAExtractor.hpp -- row partitioning of a global matrix into rank-local
// matrices (reference include/DataExchange/HBDIAExtractor.hpp: MatrixData,
// Partition, VectorPartition and the abstract extractor, same members).
//
// On one GPU the partitions are not sent anywhere: a caller builds the one
// rank's local HBDIA from its Partition and runs it at its real shape, with
// the columns owned by other ranks folded into the COO part (see
// HBDIA<T>::prepareForGPU for partial matrices).
#ifndef HBDIA_EXTRACTOR_HPP
#define HBDIA_EXTRACTOR_HPP

#include <tuple>
#include <vector>

#include "Format/HBDIA.hpp"

struct MatrixData {
  int numGlobalRows = 0;
  int numGlobalCols = 0;
  int numGlobalNonZeros = 0;
  int numLocalRows = 0;
  int numLocalCols = 0;
  int numLocalNonZeros = 0;
};

template <typename T>
struct Partition {
  std::vector<int> localRowIndices;   // row - first row of the partition
  std::vector<int> localColIndices;   // global column indices
  std::vector<T> localValues;
  std::vector<int> globalRowMapping;  // local row -> global row
};

template <typename T>
struct VectorPartition {
  std::vector<T> localValues;
};

template <typename T>
class HBDIAExtractor {
 public:
  enum class PartitioningStrategy { ROW_WISE };

  HBDIAExtractor() = default;
  virtual ~HBDIAExtractor() = default;

  virtual bool createMatrixPartitions(const HBDIA<T>& matrix, int numProcesses) = 0;
  virtual bool createVectorPartitions(const std::vector<T>& globalVector, int numProcesses) = 0;
  virtual void extractPartialMatrixMetadata(HBDIA<T>& matrix, int numProcesses) = 0;
  virtual void printProcessedDataRanges(const HBDIA<T>& matrix) = 0;
  virtual void setPartitioningStrategy(PartitioningStrategy strategy) = 0;

  virtual const std::vector<MatrixData>& getMatrixData() const = 0;
  virtual const std::vector<Partition<T>>& getPartitions() const = 0;
  virtual const std::vector<VectorPartition<T>>& getVectorPartitions() const = 0;

  virtual void clearMatrixPartitions() = 0;
  virtual void clearVectorPartitions() = 0;

 protected:
  PartitioningStrategy strategy_ = PartitioningStrategy::ROW_WISE;
  std::vector<MatrixData> matrixData_;
  std::vector<Partition<T>> partitions_;
  std::vector<VectorPartition<T>> vectorPartitions_;
};

#endif  // HBDIA_EXTRACTOR_HPP
