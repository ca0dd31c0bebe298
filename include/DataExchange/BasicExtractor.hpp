// BasicExtractor.hpp -- row-wise partitioning (reference
// include/DataExchange/BasicExtractor.hpp, src/DataExchange/BasicExtractor.cpp).
//
// createMatrixPartitions splits the COO entries by row ownership: partition p
// owns rows [p*q, p*q + q) with q = numRows / numProcesses, the last one also
// the remainder; entries keep their input order, rows become local, columns
// stay global.  extractPartialMatrixMetadata splits a partial matrix's data
// ranges (HBDIA<T>::analyzeDataRanges) at the same ownership boundaries into
// per-process request lists.
//
// exchangeHaloFromGlobal is the one-GPU stand-in for the reference's halo
// exchange (MPICommunicator<T>::exchangeData, src/DataExchange/MPICommunicator.cpp:253-413):
// the left and right halo segments of a rank-local vector are gathered from
// the global vector on the host and copied to the device; no other process
// takes part.
#ifndef BASIC_EXTRACTOR_HPP
#define BASIC_EXTRACTOR_HPP

#include <vector>

#include "DataExchange/HBDIAExtractor.hpp"
#include "Format/HBDIAVector.hpp"

template <typename T>
class BasicExtractor : public HBDIAExtractor<T> {
 public:
  BasicExtractor() = default;
  ~BasicExtractor() override = default;

  bool createMatrixPartitions(const HBDIA<T>& matrix, int numProcesses) override;
  bool createVectorPartitions(const std::vector<T>& globalVector, int numProcesses) override;
  void setPartitioningStrategy(typename HBDIAExtractor<T>::PartitioningStrategy strategy) override;
  void extractPartialMatrixMetadata(HBDIA<T>& matrix, int numProcesses) override;
  void printProcessedDataRanges(const HBDIA<T>& matrix) override;

  const std::vector<MatrixData>& getMatrixData() const override { return this->matrixData_; }
  const std::vector<Partition<T>>& getPartitions() const override { return this->partitions_; }
  const std::vector<VectorPartition<T>>& getVectorPartitions() const override { return this->vectorPartitions_; }

  void clearMatrixPartitions() override;
  void clearVectorPartitions() override;

 private:
  bool partitionRowWise(const HBDIA<T>& matrix, int numPartitions);
  bool partitionVectorRowWise(const std::vector<T>& globalVector, int numPartitions);
  void extractPartialMatrixMetadataRowWise(HBDIA<T>& matrix, int numProcesses);
};

// The rank-local matrix of `rank` out of `size` row-wise partitions of the
// global matrix, set up as the reference's distributor sets up each rank
// (src/DataExchange/BasicDistributor.cpp:57-79): partial constructor,
// convertToHBDIAFormat(blockWidth, threshold, maxCooEntries) on the host,
// analyzeDataRanges, extractPartialMatrixMetadata, calculateVectorOffsets.
// prepareForGPU() is left to the caller.
template <typename T>
HBDIA<T> createRankLocalMatrix(const HBDIA<T>& globalMatrix, int rank, int size, int blockWidth, int threshold,
                               int maxCooEntries = MAX_COO_ENTRIES,
                               ExecutionMode execMode = ExecutionMode::GPU_COO);

// Fills the halo segments of `vector` (built with HBDIAVector(local, matrix,
// rank, size)) with the entries of `globalVector` at the matrix's halo
// columns, in the host mirror and on the device.  Returns false when the
// vector's halo sizes do not match the matrix's halo layout.
template <typename T>
bool exchangeHaloFromGlobal(const HBDIA<T>& matrix, HBDIAVector<T>& vector, const std::vector<T>& globalVector);

// The check of the reference's MPICommunicator<T>::verifyDataExchange
// (src/DataExchange/MPICommunicator.cpp:415-498) for a vector whose global
// values were x[g] = g: every halo entry, walked through the request lists of
// the other ranks in order, must hold its global index, and every local entry
// its global row.  Reads the device buffer back; prints one PASSED / FAILED
// line like the reference and returns whether all entries matched.
template <typename T>
bool verifyHaloExchange(const HBDIA<T>& matrix, const HBDIAVector<T>& vector);

#endif  // BASIC_EXTRACTOR_HPP
