// HBDIA.hpp -- hybrid blocked-diagonal sparse matrix for MI355X.
//
// Drop-in for the reference's HBDIA<T> (include/Format/HBDIA.hpp:25-291): the
// same constructors, generator, conversion, getters and GPU-preparation calls.
// Device work goes through the C ABI in hbdia_hip.h; no HIP or CUDA type
// appears here.  Streams are hbdia_stream_t handles instead of cudaStream_t,
// and there is no cuSPARSE: initializeCuSparse() only marks the GPU COO path
// ready (the project's own fused COO kernel needs no handle).
//
// Memory layout on the device (prepareForGPU):
//   values      T[sum_b nOff_b * bw]   block-major, then offset, then lane
//   offsets     int32[sum_b nOff_b]    ascending per block
//   blockStart  int32[nb + 1]          prefix of offset counts
//   blockSizes  int32[nb]
//   COO         rows/cols int32, values T, row-sorted (stable), plus
//               cpuRowPtr int32[N+1] and a per-64-row group pointer
//               int32[ceil(N/64)+1] for the fused kernel.
#ifndef HBDIA_HPP
#define HBDIA_HPP

#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

#include "types.hpp"

template <typename T>
class HBDIAVector;

template <typename T>
struct PartialMatrixMetadata {
  std::vector<std::tuple<int, int>> dataRanges;
  int totalBufferSize = 0;
  std::vector<std::vector<std::tuple<int, int>>> processDataRanges;
};

template <typename T>
class HBDIA {
  friend class HBDIAVector<T>;

 public:
  HBDIA();
  HBDIA(const std::vector<int>& rowIndices, const std::vector<int>& colIndices,
        const std::vector<T>& values);
  HBDIA(const std::vector<int>& rowIndices, const std::vector<int>& colIndices,
        const std::vector<T>& values, int numRows, int numCols);
  HBDIA(const std::vector<int>& rowIndices, const std::vector<int>& colIndices,
        const std::vector<T>& values, int numRows, int numCols,
        const std::vector<int>& globalRowMapping);
  HBDIA(const std::vector<int>& rowIndices, const std::vector<int>& colIndices,
        const std::vector<T>& values, int numRows, int numCols,
        const std::vector<int>& globalRowMapping, bool partialMatrix, int numGlobalRows,
        int numGlobalCols, int numGlobalNonZeros, int rank = -1, int size = -1);
  // Takes ownership of the COO arrays without copying them (large inputs).
  HBDIA(std::vector<int>&& rowIndices, std::vector<int>&& colIndices, std::vector<T>&& values,
        int numRows, int numCols);
  ~HBDIA();

  HBDIA(const HBDIA&) = delete;
  HBDIA& operator=(const HBDIA&) = delete;
  HBDIA(HBDIA&& other) noexcept;
  HBDIA& operator=(HBDIA&& other) noexcept;

  // format loading
  void create3DStencil27Point(int nx, int ny, int nz, double noise = 0.0, int iteration = 0);

  // format conversion
  bool isCOOFormat() const { return hasCOO; }
  bool isHBDIAFormat() const { return hasHBDIA; }
  bool isDIAFormat() const { return false; }
  void convertToHBDIAFormat(int blockWidth = DEFAULT_BLOCK_WIDTH, int threshold = DEFAULT_THRESHOLD,
                            int max_coo_entries = MAX_COO_ENTRIES, bool COOisUnique = false,
                            ExecutionMode execMode = ExecutionMode::GPU_COO);
  // The host half of convertToHBDIAFormat, without the device upload that the
  // reference always performs; call prepareForGPU() afterwards to upload.
  // Returns false (and changes nothing) in the cases where the reference
  // prints a message and returns.
  bool convertToHBDIAFormatOnHost(int blockWidth = DEFAULT_BLOCK_WIDTH, int threshold = DEFAULT_THRESHOLD,
                                  int max_coo_entries = MAX_COO_ENTRIES, bool COOisUnique = false,
                                  ExecutionMode execMode = ExecutionMode::GPU_COO);
  // convertToHBDIAFormat computed on the GPU (sort by (block, offset), run
  // counts, a scan for the threshold rule, scatter): the same host state as
  // convertToHBDIAFormatOnHost, bit for bit, and -- unless the matrix is
  // partial -- the device arrays of prepareForGPU built in place on the GPU
  // rather than uploaded.  The COO is uploaded once.  Same return value and
  // messages as the host conversion; throws hbdia::DeviceError on device failures.
  // hostValues = false leaves the blocked values on the device only: the first
  // getHBDIAData() / getPtrToBlock() (or prepareForGPU, cleanupGPUData) reads
  // them back.  Not for partial matrices, whose host state is always built.
  bool convertToHBDIAFormatOnDevice(int blockWidth = DEFAULT_BLOCK_WIDTH, int threshold = DEFAULT_THRESHOLD,
                                    int max_coo_entries = MAX_COO_ENTRIES, bool COOisUnique = false,
                                    ExecutionMode execMode = ExecutionMode::GPU_COO, bool hostValues = true);

  // print wrappers (short textual dumps)
  void print() const;
  void printCOO() const;
  void printHBDIA() const;

  // format deletion
  void deleteCOOFormat();
  void deleteHBDIAFormat();
  void deleteMatrix();

  // Sums duplicate (row, col) entries and leaves the COO sorted by (row, col).
  void removeCOODuplicates();

  // Matrix Market input (reference src/Format/HBDIA.cpp:161-370).  The first
  // form takes the dimensions from the file and sorts the entries by
  // (row, col); the second uses the given dimensions, checks every index
  // against them (std::out_of_range) and keeps file order.  "symmetric" in a
  // header comment mirrors every off-diagonal entry (right after it); pattern
  // files get value 1.  Returns whether any entry was read.
  bool loadMTX(const std::string& filename);
  bool loadMTX(const std::string& filename, int numRows, int numCols);

  // dimensions
  int getNumRows() const { return numRows; }
  int getNumCols() const { return numCols; }
  int getNumNonZeros() const { return numNonZeros; }
  int getNumGlobalRows() const { return partialMatrix ? numGlobalRows : numRows; }
  int getNumGlobalCols() const { return partialMatrix ? numGlobalCols : numCols; }
  int getNumGlobalNonZeros() const { return partialMatrix ? numGlobalNonZeros : numNonZeros; }

  // COO data
  const std::vector<T>& getValues() const { return values; }
  const std::vector<int>& getRowIndices() const { return rowIndices; }
  const std::vector<int>& getColIndices() const { return colIndices; }
  const std::vector<int>& getNnzPerRow() const { return nnzPerRow; }

  // HBDIA data
  // The blocked values on the host.  After a device-only conversion the first
  // call reads them back from the device (not safe against concurrent callers).
  const std::vector<T>& getHBDIAData() const {
    materializeHostValues();
    return hbdiaData;
  }
  const std::vector<T*>& getPtrToBlock() const {
    materializeHostValues();
    return ptrToBlock;
  }
  // Stored blocked-DIA values, padding included, without reading them back.
  size_t getNumStoredValues() const { return hostValuesPending_ ? pendingValues_ : hbdiaData.size(); }
  // False between a device-only conversion and the first host access.
  bool hasHostValues() const { return !hostValuesPending_; }
  const std::vector<std::vector<int>>& getOffsetsPerBlock() const { return offsetsPerBlock; }
  const std::vector<int>& getCpuRowIndices() const { return cpuRowIndices; }
  const std::vector<int>& getCpuColIndices() const { return cpuColIndices; }
  const std::vector<T>& getCpuValues() const { return cpuValues; }
  const std::vector<int>& getCpuRowPtr() const { return cpuRowPtr; }
  int getBlockWidth() const { return blockWidth; }
  int getThreshold() const { return threshold; }
  int getMaxCooEntries() const { return max_coo_entries; }
  int getNumBlocks() const { return static_cast<int>(offsetsPerBlock.size()); }
  ExecutionMode getExecutionMode() const { return executionMode; }

  // statistics
  int getNumberDiagonals() const { return numberDiagonals; }
  const std::vector<int>& getHistogramBlocks() const { return histogramBlocks; }
  const std::vector<int>& getHistogramNnz() const { return histogramNnz; }

  // mutable COO access
  std::vector<T>& getValuesRef() { return values; }
  std::vector<int>& getRowIndicesRef() { return rowIndices; }
  std::vector<int>& getColIndicesRef() { return colIndices; }

  // global row mapping (partial matrices)
  const std::vector<int>& getGlobalRowMapping() const { return globalRowMapping; }
  std::vector<int>& getGlobalRowMappingRef() { return globalRowMapping; }
  bool hasGlobalRowMapping() const { return !globalRowMapping.empty(); }
  int getGlobalRowIndex(int localRowIndex) const {
    return (hasGlobalRowMapping() && localRowIndex < static_cast<int>(globalRowMapping.size()))
               ? globalRowMapping[localRowIndex]
               : localRowIndex;
  }
  bool isPartialMatrix() const { return partialMatrix; }
  int getRank() const { return rank_; }
  int getSize() const { return size_; }
  const PartialMatrixMetadata<T>& getPartialMatrixMetadata() const { return partialMatrixMetadata_; }
  PartialMatrixMetadata<T>& getPartialMatrixMetadataRef() { return partialMatrixMetadata_; }
  void setPartialMatrixMetadata(const PartialMatrixMetadata<T>& metadata) { partialMatrixMetadata_ = metadata; }
  bool hasPartialMatrixMetadata() const {
    return partialMatrix && partialMatrixMetadata_.totalBufferSize > 0;
  }

  // The global columns a partial matrix reads, as sorted, merged [start, end)
  // ranges (reference src/Format/HBDIA.cpp:1333-1412): every COO column, and
  // for every stored (block, offset) the column of each lane whose row exists.
  void analyzeDataRanges();
  const std::vector<std::tuple<int, int>>& getDataRanges() const { return dataRanges; }
  // Per (block, offset), the position of lane 0's x entry in the
  // [left halo | local | right halo] vector (reference HBDIA.cpp:1034-1178);
  // INT_MIN where the column is in no range.  A host table: the kernels here
  // read the halo through the rank-local COO part instead.
  void calculateVectorOffsets(int rank, int size);
  const std::vector<std::vector<int>>& getVectorOffsets() const { return vectorOffsets; }

  // Rank-local COO part of a partial matrix (one GPU, no halo exchange).  The
  // local rows are the contiguous global rows [r0, r0 + numRows), and x is
  // [left halo | local | right halo] with the halo columns of the partial
  // metadata: processDataRanges of the ranks below this one, then above.
  // The blocked part is kept as converted and reads only local columns (its
  // bound check skips the others); every stored non-zero whose column lies
  // outside the local rows is added to the COO part instead.  The list is the
  // fallback entries in their order, then those blocked entries in
  // (block, offset, lane) order, stably sorted by row; columns are indices
  // relative to the local segment (negative in the left halo, numRows and up
  // in the right one).  Built by prepareForGPU for partial matrices; needs
  // the partial metadata (analyzeDataRanges + extractPartialMatrixMetadata).
  void buildLocalCOO();
  bool hasLocalCOO() const { return localCOOBuilt_; }
  int getLocalFirstRow() const { return localFirstRow_; }
  int getHaloLeftSize() const { return haloLeft_; }
  int getHaloRightSize() const { return haloRight_; }
  // Global column of each halo entry: left halo first, then right halo.
  const std::vector<int>& getHaloColumns() const { return haloColumns_; }
  const std::vector<int>& getLocalCOORowIndices() const { return localCooRows_; }
  const std::vector<int>& getLocalCOOColIndices() const { return localCooCols_; }
  const std::vector<T>& getLocalCOOValues() const { return localCooValues_; }
  // Entries of the local COO part that come from the blocked part.
  int64_t getLocalCOOFolded() const { return localCooFolded_; }
  // Fallback entries the SpMV runs: the rank-local list for partial
  // matrices, the conversion's fallback otherwise.
  size_t numSpMVCOOEntries() const { return partialMatrix ? localCooValues_.size() : cpuValues.size(); }

  // format flags
  bool getHasCOO() const { return hasCOO; }
  bool getHasDIA() const { return false; }
  bool getHasHBDIA() const { return hasHBDIA; }
  void setHasCOO(bool flag) { hasCOO = flag; }
  void setHasHBDIA(bool flag) { hasHBDIA = flag; }

  // device arrays
  const T* getHBDIADataDevice() const { return hbdiaData_d_; }
  const int* getFlattenedOffsetsDevice() const { return flattenedOffsets_d_; }
  const int* getBlockStartIndicesDevice() const { return blockStartIndices_d_; }
  const int* getBlockSizesDevice() const { return blockSizes_d_; }
  const int* getCpuRowIndicesDevice() const { return cpuRowIndices_d_; }
  const int* getCpuColIndicesDevice() const { return cpuColIndices_d_; }
  const T* getCpuValuesDevice() const { return cpuValues_d_; }
  const int* getCpuRowPtrDevice() const { return cpuRowPtr_d_; }
  // Per-64-row-group start into the device COO arrays (fused kernel input).
  const int* getCooGroupPtrDevice() const { return cooGroupPtr_d_; }
  bool hasDeviceCOO() const { return cooGroupPtr_d_ != nullptr; }
  // Device bitset of the blocks whose offsets are runs of three consecutive
  // values (the kernels' runs3_blocks hint), or null when no block is; built
  // by prepareForGPU for block widths that are multiples of 64 or 32.
  const uint32_t* getRuns3BlocksDevice() const { return runs3Blocks_d_; }
  int64_t getRuns3BlockCount() const { return runs3Count_; }

  void prepareForGPU();
  void cleanupGPUData();

  void initializeStreams();
  void cleanupStreams();
  bool areStreamsInitialized() const { return streamsInitialized_; }

  // No cuSPARSE on this platform; kept so reference call sites compile.
  void initializeCuSparse() { cusparseInitialized_ = true; }
  void cleanupCuSparse() { cusparseInitialized_ = false; }
  bool isCuSparseInitialized() const { return cusparseInitialized_; }

  hbdia_stream_t getBDIAStream() const { return sBDIA_; }
  hbdia_stream_t getD2HStream() const { return sD2H_; }
  hbdia_stream_t getH2DStream() const { return sH2D_; }
  hbdia_stream_t getCOOStream() const { return sCOO_; }
  hbdia_stream_t getADDStream() const { return sADD_; }
  hbdia_event_t getBDIAEvent() const { return bdiaEvent_; }
  hbdia_event_t getCOOEvent() const { return cooEvent_; }

  // Waits for every stream of this matrix.
  void synchronize() const;

  // Bytes of the stored format (values + offsets + starts + COO).
  size_t storedBytes() const;

 private:
  void releaseDevice() noexcept;
  void moveFrom(HBDIA& o) noexcept;
  void materializeHostValues() const;

  // COO
  std::vector<T> values;
  std::vector<int> rowIndices;
  std::vector<int> colIndices;

  // HBDIA (the values may be read back lazily: see materializeHostValues)
  mutable std::vector<T> hbdiaData;
  mutable std::vector<T*> ptrToBlock;
  mutable bool hostValuesPending_ = false;
  size_t pendingValues_ = 0;
  std::vector<std::vector<int>> offsetsPerBlock;
  std::vector<int> cpuRowIndices;
  std::vector<int> cpuColIndices;
  std::vector<T> cpuValues;
  std::vector<int> cpuRowPtr;
  int blockWidth = 0;
  int threshold = 0;
  int max_coo_entries = MAX_COO_ENTRIES;
  ExecutionMode executionMode = ExecutionMode::GPU_COO;

  int numberDiagonals = 0;
  std::vector<int> histogramBlocks;
  std::vector<int> histogramNnz;

  int numRows = 0;
  int numCols = 0;
  int numNonZeros = 0;
  std::vector<int> nnzPerRow;

  std::vector<int> globalRowMapping;
  bool partialMatrix = false;
  int numGlobalRows = 0;
  int numGlobalCols = 0;
  int numGlobalNonZeros = 0;
  PartialMatrixMetadata<T> partialMatrixMetadata_;
  int rank_ = -1;
  int size_ = -1;
  std::vector<std::tuple<int, int>> dataRanges;
  std::vector<std::vector<int>> vectorOffsets;
  bool localCOOBuilt_ = false;
  int localFirstRow_ = 0;
  int haloLeft_ = 0;
  int haloRight_ = 0;
  std::vector<int> haloColumns_;
  std::vector<int> localCooRows_;
  std::vector<int> localCooCols_;
  std::vector<T> localCooValues_;
  int64_t localCooFolded_ = 0;

  bool hasCOO = false;
  bool hasHBDIA = false;

  T* hbdiaData_d_ = nullptr;
  int* flattenedOffsets_d_ = nullptr;
  int* blockStartIndices_d_ = nullptr;
  int* blockSizes_d_ = nullptr;
  int* cpuRowIndices_d_ = nullptr;
  int* cpuColIndices_d_ = nullptr;
  T* cpuValues_d_ = nullptr;
  int* cpuRowPtr_d_ = nullptr;
  int* cooGroupPtr_d_ = nullptr;
  uint32_t* runs3Blocks_d_ = nullptr;
  int64_t runs3Count_ = 0;

  hbdia_stream_t sBDIA_ = nullptr;
  hbdia_stream_t sD2H_ = nullptr;
  hbdia_stream_t sH2D_ = nullptr;
  hbdia_stream_t sCOO_ = nullptr;
  hbdia_stream_t sADD_ = nullptr;
  hbdia_event_t bdiaEvent_ = nullptr;
  hbdia_event_t cooEvent_ = nullptr;
  bool streamsInitialized_ = false;
  bool cusparseInitialized_ = false;
};

#endif  // HBDIA_HPP
