// HBDIAStats.hpp -- statistics and storage-cost report of a converted matrix.
//
// The figures of the reference's printMatrixInfo (benchmarking/benchmark.cpp:
// 144-230), as data: the per-block diagonal histograms and the bytes the
// matrix takes as HBDIA, COO, DIA, SELL (slices of blockWidth rows) and CSR,
// each with its ratio to COO.  The same formulas, evaluated in 64-bit integers
// (the reference multiplies the DIA term in int), with sizeof(T) where the
// reference fixes double.  SELL needs the per-row counts that only the stencil
// generator records (getNnzPerRow); for other matrices it reduces to the slice
// pointers, as in the reference.  The reference reports SELL as meaningful
// only for threshold 1 (sellApplies).
#ifndef HBDIA_STATS_HPP
#define HBDIA_STATS_HPP

#include <cstdint>
#include <string>
#include <vector>

#include "../Format/HBDIA.hpp"

struct HBDIAHistogramRow {
  int64_t diagonals;  // offsets per block
  int64_t blocks;     // blocks with that many offsets
  int64_t nnz;        // entries placed in those blocks
};

struct HBDIAStorageReport {
  int64_t rows = 0, cols = 0, nnz = 0, numberDiagonals = 0;
  int64_t numBlocks = 0, blockWidth = 0, threshold = 0, maxCooEntries = 0;
  int64_t cooFallbackEntries = 0, hbdiaDataSize = 0, storedOffsets = 0;
  int64_t valueBytes = 0;  // sizeof(T)
  int64_t hbdiaBytes = 0, cooBytes = 0, diaBytes = 0, sellBytes = 0, csrBytes = 0;
  double hbdiaVsCoo = 0, diaVsCoo = 0, csrVsCoo = 0, sellVsCoo = 0;
  bool sellApplies = false;
  // rows of the reference's histogram.csv: counts with blocks or entries
  std::vector<HBDIAHistogramRow> histogram;
};

// Throws std::invalid_argument when the matrix is not in HBDIA format.
template <typename T>
HBDIAStorageReport hbdiaStorageReport(const HBDIA<T>& matrix);

// The report as the reference's two CSV files would hold it (matrix_info.csv
// without the run's configuration columns, then a blank line and
// histogram.csv), for logging next to benchmark results.
std::string hbdiaStorageReportCSV(const HBDIAStorageReport& r, const std::string& name);

#endif  // HBDIA_STATS_HPP
