// HBDIAPartial.cpp -- partial (rank-local) matrices: the columns they read,
// the reference's vector-offset table, and the rank-local COO part that lets
// one rank's matrix run on one GPU with its halo columns in the COO part.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>

#include "Format/HBDIA.hpp"

// Reference src/Format/HBDIA.cpp:1333-1412.  The reference walks every lane
// of every block, also lanes past the last row of a partial last block, where
// it reads past the end of the global row mapping; here only lanes whose row
// exists contribute.  Columns outside [0, numGlobalCols) are skipped, as there.
template <typename T>
void HBDIA<T>::analyzeDataRanges() {
  if (!partialMatrix) {
    std::cout << "Cannot analyze data ranges: not a partial matrix" << std::endl;
    return;
  }
  if (!hasHBDIA) {
    std::cerr << "Cannot analyze data ranges: HBDIA format not available" << std::endl;
    return;
  }
  // one flag per global column, then runs of set flags
  const int nGC = getNumGlobalCols();
  std::vector<uint8_t> used(static_cast<size_t>(std::max(nGC, 0)), 0);
  for (int c : colIndices)
    if (c >= 0 && c < nGC) used[c] = 1;
  for (size_t b = 0; b < offsetsPerBlock.size(); ++b) {
    const int64_t first = static_cast<int64_t>(b) * blockWidth;
    const int64_t last = std::min<int64_t>(first + blockWidth, numRows);
    for (int off : offsetsPerBlock[b]) {
      for (int64_t r = first; r < last; ++r) {
        const int64_t c = static_cast<int64_t>(getGlobalRowIndex(static_cast<int>(r))) + off;
        if (c >= 0 && c < nGC) used[c] = 1;
      }
    }
  }
  dataRanges.clear();
  int64_t c = 0;
  while (c < nGC) {
    if (!used[c]) {
      ++c;
      continue;
    }
    int64_t e = c + 1;
    while (e < nGC && used[e]) ++e;
    dataRanges.emplace_back(static_cast<int>(c), static_cast<int>(e));
    c = e;
  }
}

// Reference src/Format/HBDIA.cpp:1034-1178 (calculateVectorOffsets and
// findMemoryOffsetForGlobalIndex), including its local-range rule for the
// last rank (the remainder taken modulo the rows per process).  Indices in no
// range become INT_MIN; one message reports how many there were.
template <typename T>
void HBDIA<T>::calculateVectorOffsets(int rank, int size) {
  if (!partialMatrix) {
    std::cerr << "Error: Cannot calculate vector offsets - this is not a partial matrix" << std::endl;
    return;
  }
  if (!hasHBDIA) {
    std::cerr << "Error: Cannot calculate vector offsets - HBDIA format not available" << std::endl;
    return;
  }
  if (!hasPartialMatrixMetadata()) {
    std::cerr << "Error: Cannot calculate vector offsets - no partial matrix metadata available" << std::endl;
    return;
  }
  if (size <= 0 || rank < 0 || rank >= size) throw std::invalid_argument("calculateVectorOffsets: bad rank/size");
  const auto& pdr = partialMatrixMetadata_.processDataRanges;
  int64_t left = 0;
  for (int p = 0; p < static_cast<int>(pdr.size()); ++p) {
    if (p >= rank) break;
    for (const auto& rg : pdr[p]) left += std::get<1>(rg) - std::get<0>(rg);
  }
  const int rpp = numGlobalRows / size;
  if (rpp <= 0) throw std::invalid_argument("calculateVectorOffsets: fewer global rows than processes");
  const int64_t gStart = static_cast<int64_t>(rank) * rpp;
  const int64_t gEnd = gStart + rpp + (rank == size - 1 ? numGlobalRows % rpp : 0);
  const int64_t local = numRows;
  // The reference takes the first range (in request order) that holds g, at
  // its position in the buffer; for ascending, disjoint ranges (what
  // extractPartialMatrixMetadata produces) that is a binary search.
  struct Seg {
    int64_t s, e, at;
  };
  auto segments = [&](int p0, int p1, int64_t at) {
    std::vector<Seg> v;
    for (int p = p0; p < p1 && p < static_cast<int>(pdr.size()); ++p)
      for (const auto& rg : pdr[p]) {
        v.push_back({std::get<0>(rg), std::get<1>(rg), at});
        at += std::get<1>(rg) - std::get<0>(rg);
      }
    return v;
  };
  const std::vector<Seg> lsegs = segments(0, rank, 0), rsegs = segments(rank + 1, size, left + local);
  auto sorted = [](const std::vector<Seg>& v) {
    for (size_t i = 1; i < v.size(); ++i)
      if (v[i].s < v[i - 1].e) return false;
    return true;
  };
  const bool lsorted = sorted(lsegs), rsorted = sorted(rsegs);
  auto find = [](const std::vector<Seg>& v, bool bin, int64_t g) -> int64_t {
    if (bin) {
      auto it = std::upper_bound(v.begin(), v.end(), g, [](int64_t x, const Seg& q) { return x < q.s; });
      if (it != v.begin() && g < (it - 1)->e) return (it - 1)->at + (g - (it - 1)->s);
      return INT_MIN;
    }
    for (const Seg& q : v)
      if (g >= q.s && g < q.e) return q.at + (g - q.s);
    return INT_MIN;
  };
  auto memoryOffset = [&](int64_t g) -> int64_t {
    if (g >= gStart && g < gEnd) return left + (g - gStart);
    if (g < 0) return g;
    return g < gStart ? find(lsegs, lsorted, g) : find(rsegs, rsorted, g);
  };
  vectorOffsets.assign(offsetsPerBlock.size(), {});
  int64_t missing = 0;
  for (size_t b = 0; b < offsetsPerBlock.size(); ++b) {
    const int64_t blockStart = gStart + static_cast<int64_t>(b) * blockWidth;
    auto& vo = vectorOffsets[b];
    vo.reserve(offsetsPerBlock[b].size());
    for (int off : offsetsPerBlock[b]) {
      const int64_t m = memoryOffset(blockStart + off);
      if (m == INT_MIN) ++missing;
      vo.push_back(static_cast<int>(m));
    }
  }
  if (missing)
    std::cerr << rank << "\t Error: " << missing << " global indices not found in any buffer ranges" << std::endl;
}

template <typename T>
void HBDIA<T>::buildLocalCOO() {
  if (!partialMatrix) throw std::logic_error("buildLocalCOO: not a partial matrix");
  if (!hasHBDIA) throw std::logic_error("buildLocalCOO: HBDIA format not available");
  if (!hasPartialMatrixMetadata())
    throw std::runtime_error("buildLocalCOO: no partial matrix metadata (call analyzeDataRanges and "
                             "extractPartialMatrixMetadata first)");
  const auto& pdr = partialMatrixMetadata_.processDataRanges;
  if (size_ <= 0 || rank_ < 0 || rank_ >= size_ || static_cast<int>(pdr.size()) != size_)
    throw std::invalid_argument("buildLocalCOO: the matrix's rank/size do not match its process data ranges");
  if (static_cast<int>(globalRowMapping.size()) != numRows)
    throw std::invalid_argument("buildLocalCOO: a partial matrix needs one global row per local row");
  const int r0 = globalRowMapping.empty() ? 0 : globalRowMapping[0];
  for (int i = 0; i < numRows; ++i)
    if (globalRowMapping[i] != r0 + i)
      throw std::invalid_argument("buildLocalCOO: the global rows of a rank must be contiguous");
  const int64_t rEnd = static_cast<int64_t>(r0) + numRows;

  // halo columns: ranks below, then ranks above, each in range order
  std::vector<int> halo;
  int64_t left = 0;
  for (int p = 0; p < size_; ++p) {
    if (p == rank_) {
      left = static_cast<int64_t>(halo.size());
      continue;
    }
    for (const auto& rg : pdr[p])
      for (int c = std::get<0>(rg); c < std::get<1>(rg); ++c) halo.push_back(c);
  }
  const int64_t right = static_cast<int64_t>(halo.size()) - left;
  materializeHostValues();
  if (left + numRows + right > static_cast<int64_t>(INT32_MAX))
    throw std::runtime_error("buildLocalCOO: the vector with its halo exceeds 2^31-1 entries");
  for (size_t i = 0; i < halo.size(); ++i) {
    const bool ordered = i == 0 || static_cast<int64_t>(i) == left || halo[i] > halo[i - 1];
    const bool leftSide = static_cast<int64_t>(i) < left;
    if (!ordered || (leftSide && halo[i] >= r0) || (!leftSide && halo[i] < rEnd))
      throw std::invalid_argument("buildLocalCOO: halo ranges must be ascending, below the local rows for lower "
                                  "ranks and above them for higher ranks");
  }
  const auto lBegin = halo.begin(), lEnd = halo.begin() + left, rBegin = lEnd, rEndIt = halo.end();
  // global column -> index relative to the local segment
  auto localColumn = [&](int64_t g) -> int {
    if (g >= r0 && g < rEnd) return static_cast<int>(g - r0);
    if (g < r0) {
      auto it = std::lower_bound(lBegin, lEnd, static_cast<int>(g));
      if (it != lEnd && *it == g) return static_cast<int>((it - lBegin) - left);
    } else {
      auto it = std::lower_bound(rBegin, rEndIt, static_cast<int>(g));
      if (it != rEndIt && *it == g) return static_cast<int>(numRows + (it - rBegin));
    }
    throw std::runtime_error("buildLocalCOO: column " + std::to_string(g) +
                             " is neither local nor in the halo ranges");
  };

  std::vector<int> rows, cols;
  std::vector<T> vals;
  rows.reserve(cpuValues.size());
  cols.reserve(cpuValues.size());
  vals.reserve(cpuValues.size());
  for (size_t i = 0; i < cpuValues.size(); ++i) {
    rows.push_back(cpuRowIndices[i]);
    cols.push_back(localColumn(cpuColIndices[i]));
    vals.push_back(cpuValues[i]);
  }
  // blocked entries whose column is outside the local rows: only blocks whose
  // offsets reach past [0, numRows) hold any
  const int64_t nGC = getNumGlobalCols();
  int64_t folded = 0;
  for (size_t b = 0; b < offsetsPerBlock.size(); ++b) {
    const auto& ob = offsetsPerBlock[b];
    if (ob.empty() || !ptrToBlock[b]) continue;
    const int64_t first = static_cast<int64_t>(b) * blockWidth;
    const int64_t last = std::min<int64_t>(first + blockWidth, numRows);
    if (first + ob.front() >= 0 && last - 1 + ob.back() < numRows) continue;
    for (size_t k = 0; k < ob.size(); ++k) {
      for (int64_t r = first; r < last; ++r) {
        const T v = ptrToBlock[b][k * static_cast<size_t>(blockWidth) + static_cast<size_t>(r - first)];
        if (v == T(0)) continue;
        const int64_t c = r + ob[k];  // relative to the local segment
        if (c >= 0 && c < numRows) continue;
        const int64_t g = r0 + c;
        if (g < 0 || g >= nGC) continue;
        rows.push_back(static_cast<int>(r));
        cols.push_back(localColumn(g));
        vals.push_back(v);
        ++folded;
      }
    }
  }
  // stable counting sort by row
  std::vector<int64_t> start(static_cast<size_t>(numRows) + 1, 0);
  for (int r : rows) start[static_cast<size_t>(r) + 1]++;
  for (int i = 0; i < numRows; ++i) start[i + 1] += start[i];
  localCooRows_.assign(rows.size(), 0);
  localCooCols_.assign(rows.size(), 0);
  localCooValues_.assign(rows.size(), T(0));
  for (size_t i = 0; i < rows.size(); ++i) {
    const int64_t d = start[rows[i]]++;
    localCooRows_[d] = rows[i];
    localCooCols_[d] = cols[i];
    localCooValues_[d] = vals[i];
  }
  localFirstRow_ = r0;
  haloLeft_ = static_cast<int>(left);
  haloRight_ = static_cast<int>(right);
  haloColumns_ = std::move(halo);
  localCooFolded_ = folded;
  localCOOBuilt_ = true;
}

#define HBDIA_PARTIAL_INSTANTIATE(T)                     \
  template void HBDIA<T>::analyzeDataRanges();           \
  template void HBDIA<T>::calculateVectorOffsets(int, int); \
  template void HBDIA<T>::buildLocalCOO();

HBDIA_PARTIAL_INSTANTIATE(double)
HBDIA_PARTIAL_INSTANTIATE(float)
HBDIA_PARTIAL_INSTANTIATE(int)
