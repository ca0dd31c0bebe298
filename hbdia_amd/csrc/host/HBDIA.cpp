// HBDIA.cpp -- host side of the HBDIA matrix: COO construction, the 27-point
// stencil generator, the COO -> HBDIA conversion and the device upload.
//
// Semantics follow the reference (src/Format/HBDIA.cpp); the implementation
// is new: the conversion runs block-parallel with OpenMP (no std::map per
// block, no std::find per entry), the generator sorts (row, index) pairs
// instead of an index permutation, and all device work goes through the C ABI.
#include "Format/HBDIA.hpp"

#include <omp.h>

#include "staging.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <iostream>
#include <numeric>
#include <random>
#include <stdexcept>
#include <utility>

namespace {

bool verbose() {
  static const bool v = [] {
    const char* e = std::getenv("HBDIA_VERBOSE");
    return e && *e && *e != '0';
  }();
  return v;
}

// Phase timer for HBDIA_VERBOSE output.
struct Phases {
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void mark(const char* what) {
    const auto now = std::chrono::steady_clock::now();
    if (verbose())
      std::cout << "  " << what << ": " << std::chrono::duration<double, std::milli>(now - t).count() << " ms"
                << std::endl;
    t = now;
  }
};

template <typename T>
void validate_coo(const std::vector<int>& r, const std::vector<int>& c, const std::vector<T>& v,
                  int numRows, int numCols) {
  if (r.size() != c.size() || r.size() != v.size())
    throw std::invalid_argument("Row indices, column indices, and values vectors must have the same size");
  if (numRows <= 0 || numCols <= 0) throw std::invalid_argument("Matrix dimensions must be positive");
  if (r.size() > static_cast<size_t>(0x7fffffff))
    throw std::invalid_argument("More than 2^31-1 non-zeros are not supported");
  const int64_t n = static_cast<int64_t>(r.size());
  int64_t bad = -1;
#pragma omp parallel for schedule(static) reduction(max : bad)
  for (int64_t i = 0; i < n; ++i)
    if (r[i] < 0 || r[i] >= numRows || c[i] < 0 || c[i] >= numCols) bad = std::max(bad, i);
  if (bad >= 0) {
    // report the first offending entry, as the reference's sequential scan does
    for (int64_t i = 0; i < n; ++i) {
      if (r[i] < 0 || r[i] >= numRows)
        throw std::out_of_range("Row index " + std::to_string(r[i]) + " is out of range [0, " +
                                std::to_string(numRows - 1) + "]");
      if (c[i] < 0 || c[i] >= numCols)
        throw std::out_of_range("Column index " + std::to_string(c[i]) + " is out of range [0, " +
                                std::to_string(numCols - 1) + "]");
    }
  }
}

template <typename U>
void dev_upload(U** dst, const U* src, size_t count) {
  *dst = nullptr;
  if (count == 0) return;
  void* p = nullptr;
  CHECK_HBDIA(hbdia_malloc(&p, count * sizeof(U)));
  *dst = static_cast<U*>(p);
  hbdia_host::upload_staged(p, src, count * sizeof(U));
}

template <typename U>
void dev_free(U*& p) noexcept {
  if (p) (void)hbdia_free(p);
  p = nullptr;
}

// cpuRowPtr by the reference's gap-filling rule (src/Format/HBDIA.cpp:676-708):
// on each change of row r, every slot from the previous row + 1 through r gets
// the current entry index; the tail gets the entry count.  For row-sorted
// entries this is the CSR row pointer.
void fill_cpu_row_ptr(const std::vector<int>& rows, int numRows, std::vector<int>& ptr) {
  ptr.assign(static_cast<size_t>(numRows) + 1, 0);
  int currentRow = -1, nextOffset = 0;
  for (size_t q = 0; q < rows.size(); ++q) {
    const int r = rows[q];
    if (r != currentRow) {
      for (int z = currentRow + 1; z <= r; z++) ptr[z] = nextOffset;
      currentRow = r;
    }
    nextOffset++;
  }
  for (int z = currentRow + 1; z <= numRows; z++) ptr[z] = nextOffset;
}

// Owns the device arrays of one hbdia_convert_coo result until they are adopted.
struct ConvertResult {
  hbdia_convert_result_t r{};
  ~ConvertResult() { (void)hbdia_convert_release(&r); }
  template <typename P>
  P* take(P*& field) {
    P* p = field;
    field = nullptr;
    return p;
  }
};

template <typename U>
void dev_download(std::vector<U>& dst, const void* src, size_t count) {
  hbdia_host::assign_zeros(dst, count);
  if (count) hbdia_host::download_staged(dst.data(), src, count * sizeof(U));
}

}  // namespace

// ---------------------------------------------------------------------------
// construction (reference src/Format/HBDIA.cpp:20-154)

template <typename T>
HBDIA<T>::HBDIA() = default;

template <typename T>
HBDIA<T>::HBDIA(const std::vector<int>& r, const std::vector<int>& c, const std::vector<T>& v)
    : values(v), rowIndices(r), colIndices(c), hasCOO(true) {
  if (r.size() != c.size() || r.size() != v.size())
    throw std::invalid_argument("Row indices, column indices, and values vectors must have the same size");
  numNonZeros = static_cast<int>(v.size());
  numRows = 0;
  numCols = 0;
  for (int x : r) numRows = std::max(numRows, x + 1);
  for (int x : c) numCols = std::max(numCols, x + 1);
}

template <typename T>
HBDIA<T>::HBDIA(const std::vector<int>& r, const std::vector<int>& c, const std::vector<T>& v,
                int nR, int nC)
    : values(v), rowIndices(r), colIndices(c), numRows(nR), numCols(nC), hasCOO(true) {
  validate_coo(r, c, v, nR, nC);
  numNonZeros = static_cast<int>(v.size());
}

template <typename T>
HBDIA<T>::HBDIA(const std::vector<int>& r, const std::vector<int>& c, const std::vector<T>& v,
                int nR, int nC, const std::vector<int>& grm)
    : values(v), rowIndices(r), colIndices(c), numRows(nR), numCols(nC), globalRowMapping(grm),
      hasCOO(true) {
  if (r.size() != c.size() || r.size() != v.size())
    throw std::invalid_argument("Row indices, column indices, and values vectors must have the same size");
  if (nR <= 0 || nC <= 0) throw std::invalid_argument("Matrix dimensions must be positive");
  if (!grm.empty() && static_cast<int>(grm.size()) != nR)
    throw std::invalid_argument("Global row mapping size must match number of rows");
  validate_coo(r, c, v, nR, nC);
  numNonZeros = static_cast<int>(v.size());
}

template <typename T>
HBDIA<T>::HBDIA(const std::vector<int>& r, const std::vector<int>& c, const std::vector<T>& v,
                int nR, int nC, const std::vector<int>& grm, bool partial, int nGR, int nGC,
                int nGNZ, int rank, int size)
    : values(v), rowIndices(r), colIndices(c), numRows(nR), numCols(nC), globalRowMapping(grm),
      partialMatrix(partial), numGlobalRows(nGR), numGlobalCols(nGC), numGlobalNonZeros(nGNZ),
      rank_(rank), size_(size), hasCOO(true) {
  if (r.size() != c.size() || r.size() != v.size())
    throw std::invalid_argument("Row indices, column indices, and values vectors must have the same size");
  if (nR <= 0 || nC <= 0) throw std::invalid_argument("Matrix dimensions must be positive");
  if (!grm.empty() && static_cast<int>(grm.size()) != nR)
    throw std::invalid_argument("Global row mapping size must match number of rows");
  validate_coo(r, c, v, nR, nC);
  numNonZeros = static_cast<int>(v.size());
}

template <typename T>
HBDIA<T>::HBDIA(std::vector<int>&& r, std::vector<int>&& c, std::vector<T>&& v, int nR, int nC)
    : values(std::move(v)), rowIndices(std::move(r)), colIndices(std::move(c)), numRows(nR),
      numCols(nC), hasCOO(true) {
  validate_coo(rowIndices, colIndices, values, nR, nC);
  numNonZeros = static_cast<int>(values.size());
}

template <typename T>
HBDIA<T>::~HBDIA() {
  releaseDevice();
}

template <typename T>
void HBDIA<T>::moveFrom(HBDIA& o) noexcept {
  values = std::move(o.values);
  rowIndices = std::move(o.rowIndices);
  colIndices = std::move(o.colIndices);
  hbdiaData = std::move(o.hbdiaData);
  ptrToBlock = std::move(o.ptrToBlock);
  hostValuesPending_ = std::exchange(o.hostValuesPending_, false);
  pendingValues_ = std::exchange(o.pendingValues_, 0);
  offsetsPerBlock = std::move(o.offsetsPerBlock);
  cpuRowIndices = std::move(o.cpuRowIndices);
  cpuColIndices = std::move(o.cpuColIndices);
  cpuValues = std::move(o.cpuValues);
  cpuRowPtr = std::move(o.cpuRowPtr);
  blockWidth = o.blockWidth;
  threshold = o.threshold;
  max_coo_entries = o.max_coo_entries;
  executionMode = o.executionMode;
  numberDiagonals = o.numberDiagonals;
  histogramBlocks = std::move(o.histogramBlocks);
  histogramNnz = std::move(o.histogramNnz);
  numRows = o.numRows;
  numCols = o.numCols;
  numNonZeros = o.numNonZeros;
  nnzPerRow = std::move(o.nnzPerRow);
  globalRowMapping = std::move(o.globalRowMapping);
  partialMatrix = o.partialMatrix;
  numGlobalRows = o.numGlobalRows;
  numGlobalCols = o.numGlobalCols;
  numGlobalNonZeros = o.numGlobalNonZeros;
  partialMatrixMetadata_ = std::move(o.partialMatrixMetadata_);
  rank_ = o.rank_;
  size_ = o.size_;
  dataRanges = std::move(o.dataRanges);
  vectorOffsets = std::move(o.vectorOffsets);
  localCOOBuilt_ = std::exchange(o.localCOOBuilt_, false);
  localFirstRow_ = o.localFirstRow_;
  haloLeft_ = o.haloLeft_;
  haloRight_ = o.haloRight_;
  haloColumns_ = std::move(o.haloColumns_);
  localCooRows_ = std::move(o.localCooRows_);
  localCooCols_ = std::move(o.localCooCols_);
  localCooValues_ = std::move(o.localCooValues_);
  localCooFolded_ = o.localCooFolded_;
  hasCOO = o.hasCOO;
  hasHBDIA = o.hasHBDIA;
  hbdiaData_d_ = std::exchange(o.hbdiaData_d_, nullptr);
  flattenedOffsets_d_ = std::exchange(o.flattenedOffsets_d_, nullptr);
  blockStartIndices_d_ = std::exchange(o.blockStartIndices_d_, nullptr);
  blockSizes_d_ = std::exchange(o.blockSizes_d_, nullptr);
  cpuRowIndices_d_ = std::exchange(o.cpuRowIndices_d_, nullptr);
  cpuColIndices_d_ = std::exchange(o.cpuColIndices_d_, nullptr);
  cpuValues_d_ = std::exchange(o.cpuValues_d_, nullptr);
  cpuRowPtr_d_ = std::exchange(o.cpuRowPtr_d_, nullptr);
  cooGroupPtr_d_ = std::exchange(o.cooGroupPtr_d_, nullptr);
  runs3Blocks_d_ = std::exchange(o.runs3Blocks_d_, nullptr);
  runs3Count_ = std::exchange(o.runs3Count_, 0);
  sBDIA_ = std::exchange(o.sBDIA_, nullptr);
  sD2H_ = std::exchange(o.sD2H_, nullptr);
  sH2D_ = std::exchange(o.sH2D_, nullptr);
  sCOO_ = std::exchange(o.sCOO_, nullptr);
  sADD_ = std::exchange(o.sADD_, nullptr);
  bdiaEvent_ = std::exchange(o.bdiaEvent_, nullptr);
  cooEvent_ = std::exchange(o.cooEvent_, nullptr);
  streamsInitialized_ = std::exchange(o.streamsInitialized_, false);
  cusparseInitialized_ = std::exchange(o.cusparseInitialized_, false);
  // ptrToBlock points into hbdiaData, whose buffer moved with it.
}

template <typename T>
HBDIA<T>::HBDIA(HBDIA&& o) noexcept {
  moveFrom(o);
}

template <typename T>
HBDIA<T>& HBDIA<T>::operator=(HBDIA&& o) noexcept {
  if (this != &o) {
    releaseDevice();
    moveFrom(o);
  }
  return *this;
}

// ---------------------------------------------------------------------------
// 27-point stencil generator (reference src/Format/HBDIA.cpp:1416-1607).
//
// Same mt19937(RAND_SEED + iteration) stream, same distributions drawn in the
// same order, same loop nest (outer index over nz, bounds-checked against nx),
// same weights 26.1 / -1.1 / -0.1 / -0.01 and the same row ordering: the
// reference sorts an index permutation with an unstable std::sort keyed on the
// row.  std::sort's result depends only on the sequence of comparison
// outcomes, so sorting (row, index) pairs keyed on the row yields the very same
// permutation while touching memory sequentially.

template <typename T>
void HBDIA<T>::create3DStencil27Point(int nx, int ny, int nz, double noise, int iteration) {
  values.clear();
  rowIndices.clear();
  colIndices.clear();
  deleteHBDIAFormat();

  const int totalNodes = nx * ny * nz;
  nnzPerRow.assign(totalNodes, 0);
  auto getIndex = [&](int i, int j, int k) -> int { return k * nx * ny + j * nx + i; };

  std::mt19937 gen(RAND_SEED + iteration);
  std::uniform_int_distribution<int> colDist(0, totalNodes - 1);
  std::uniform_real_distribution<double> valueDist(-1.0, 1.0);
  std::vector<int> noiseEntriesPerRow(totalNodes, 0);
  int totalTargetNoiseEntries = 0;
  if (noise > 0.0) {
    long long stencilEntries = totalNodes * 27;
    totalTargetNoiseEntries = static_cast<int>(static_cast<double>(stencilEntries) * noise);
    std::uniform_int_distribution<int> rowDist(0, totalNodes - 1);
    for (int e = 0; e < totalTargetNoiseEntries; ++e) noiseEntriesPerRow[rowDist(gen)]++;
  }
  const size_t reserve = static_cast<size_t>(totalNodes) * 27 + static_cast<size_t>(totalTargetNoiseEntries);
  rowIndices.reserve(reserve);
  colIndices.reserve(reserve);
  values.reserve(reserve);

  std::vector<int> used;
  used.reserve(64);
  for (int i = 0; i < nz; i++) {
    for (int j = 0; j < ny; j++) {
      for (int k = 0; k < nx; k++) {
        const int center = getIndex(i, j, k);
        for (int di = -1; di <= 1; di++)
          for (int dj = -1; dj <= 1; dj++)
            for (int dk = -1; dk <= 1; dk++) {
              const int ni = i + di, nj = j + dj, nk = k + dk;
              if (ni >= 0 && ni < nx && nj >= 0 && nj < ny && nk >= 0 && nk < nz) {
                rowIndices.push_back(center);
                colIndices.push_back(getIndex(ni, nj, nk));
                nnzPerRow[center]++;
                const int kind = (di != 0) + (dj != 0) + (dk != 0);
                values.push_back(kind == 0   ? static_cast<T>(26.1)
                                 : kind == 1 ? static_cast<T>(-1.1)
                                 : kind == 2 ? static_cast<T>(-0.1)
                                             : static_cast<T>(-0.01));
              }
            }
        const int target = noiseEntriesPerRow[center];
        if (target > 0) {
          used.clear();
          for (int dk = -1; dk <= 1; dk++)
            for (int dj = -1; dj <= 1; dj++)
              for (int di = -1; di <= 1; di++) {
                const int ni = i + di, nj = j + dj, nk = k + dk;
                if (ni >= 0 && ni < nx && nj >= 0 && nj < ny && nk >= 0 && nk < nz)
                  used.push_back(getIndex(ni, nj, nk));
              }
          int added = 0, attempts = 0;
          const int maxAttempts = target * 10 + 100;
          while (added < target && attempts < maxAttempts) {
            attempts++;
            const int col = colDist(gen);
            if (std::find(used.begin(), used.end(), col) != used.end()) continue;
            used.push_back(col);
            rowIndices.push_back(center);
            colIndices.push_back(col);
            values.push_back(static_cast<T>(valueDist(gen)));
            added++;
            nnzPerRow[center]++;
          }
        }
      }
    }
  }

  const size_t n = rowIndices.size();
  std::vector<std::pair<int, uint32_t>> keyed(n);
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < static_cast<int64_t>(n); ++q)
    keyed[q] = {rowIndices[q], static_cast<uint32_t>(q)};
  std::sort(keyed.begin(), keyed.end(),
            [](const std::pair<int, uint32_t>& a, const std::pair<int, uint32_t>& b) {
              return a.first < b.first;
            });
  {
    std::vector<int> sc(n);
    std::vector<T> sv(n);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < static_cast<int64_t>(n); ++q) {
      sc[q] = colIndices[keyed[q].second];
      sv[q] = values[keyed[q].second];
      rowIndices[q] = keyed[q].first;
    }
    colIndices = std::move(sc);
    values = std::move(sv);
  }

  numRows = totalNodes;
  numCols = totalNodes;
  numNonZeros = static_cast<int>(values.size());
  hasCOO = true;
  hasHBDIA = false;
  partialMatrix = false;
  globalRowMapping.clear();
  if (verbose())
    std::cout << "Created 3D " << nx << "x" << ny << "x" << nz << " stencil matrix: " << numRows
              << " rows, " << numNonZeros << " non-zeros" << std::endl;
}

// ---------------------------------------------------------------------------
// COO -> HBDIA conversion (reference src/Format/HBDIA.cpp:507-718).
//
// Result, identical to the reference's for the same input:
//   * nb = ceil(min(numRows, numCols) / bw) blocks, block = row / bw;
//   * per block, offset = col - row (global row for partial matrices) is kept
//     iff its count in the block is >= threshold, or once the running number of
//     fallback entries (blocks ascending, offsets ascending) reached
//     max_coo_entries;
//   * kept offsets ascending; each non-empty block owns |offsets| * bw zeroed
//     slots; an entry is stored by assignment at [offset index][lane]
//     (duplicates: the later entry wins);
//   * every other entry goes to the fallback COO in input order, with the
//     reference's cpuRowPtr gap-filling rule;
//   * numberDiagonals, histogramBlocks and histogramNnz as in the reference.
// Implementation: entries are grouped by block (directly when the input is
// block-ordered, else by a stable counting sort), each block is analysed and
// filled independently in parallel, and only the global max_coo_entries rule
// runs as one sequential pass over (block, offset) pairs.

template <typename T>
void HBDIA<T>::convertToHBDIAFormat(int bw, int thr, int maxCoo, bool COOisUnique,
                                    ExecutionMode execMode) {
  Phases ph;
  if (!convertToHBDIAFormatOnHost(bw, thr, maxCoo, COOisUnique, execMode)) return;
  ph.mark("host conversion");
  if (!partialMatrix) prepareForGPU();
  ph.mark("upload (prepareForGPU)");
}

template <typename T>
bool HBDIA<T>::convertToHBDIAFormatOnHost(int bw, int thr, int maxCoo, bool /*COOisUnique*/,
                                          ExecutionMode execMode) {
  if (verbose())
    std::cout << "blockWidth: " << bw << ", threshold: " << thr << ", execMode: "
              << (execMode == ExecutionMode::GPU_COO ? "GPU_COO" : "CPU_COO") << std::endl;
  if (hasHBDIA) {
    std::cout << "Matrix is already in HBDIA format" << std::endl;
    return false;
  }
  if (!hasCOO) {
    std::cout << "No COO format available for conversion to HBDIA" << std::endl;
    return false;
  }
  if (values.empty()) {
    std::cout << "Cannot convert empty matrix to HBDIA format" << std::endl;
    return false;
  }
  if (bw <= 0) throw std::invalid_argument("blockWidth must be positive");
  blockWidth = bw;
  threshold = thr;
  max_coo_entries = maxCoo;
  executionMode = execMode;

  const int nb = (std::min(numRows, numCols) + bw - 1) / bw;
  const int64_t nnz = static_cast<int64_t>(values.size());
  const int* R = rowIndices.data();
  const int* C = colIndices.data();
  const T* V = values.data();
  const bool useGlobal = partialMatrix && !globalRowMapping.empty();
  const int* G = globalRowMapping.data();
  auto keyOf = [&](int64_t i) -> int {
    const int b = R[i] / bw;
    return b < nb ? b : nb;
  };
  auto offsetOf = [&](int64_t i) -> int { return C[i] - (useGlobal ? G[R[i]] : R[i]); };

  // 1. group entries by block key (nb = "no block": straight to the fallback)
  bool ordered = true;
#pragma omp parallel for schedule(static) reduction(&& : ordered)
  for (int64_t i = 1; i < nnz; ++i)
    if (keyOf(i) < keyOf(i - 1)) ordered = false;

  std::vector<int64_t> bstart(static_cast<size_t>(nb) + 2, 0);
  std::vector<int64_t> order;  // entry indices grouped by key, stable; empty when `ordered`
  if (ordered) {
    // bstart[k] = first entry with key >= k
    const int k0 = keyOf(0);
    for (int k = 0; k <= k0; ++k) bstart[k] = 0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 1; i < nnz; ++i) {
      const int a = keyOf(i - 1), b = keyOf(i);
      for (int k = a + 1; k <= b; ++k) bstart[k] = i;
    }
    for (int k = keyOf(nnz - 1) + 1; k <= nb + 1; ++k) bstart[k] = nnz;
  } else {
    const int nth = omp_get_max_threads();
    std::vector<std::vector<int64_t>> cnt(nth, std::vector<int64_t>(static_cast<size_t>(nb) + 1, 0));
    const int64_t chunk = (nnz + nth - 1) / nth;
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int t = 0; t < nth; ++t) {
      const int64_t lo = std::min(nnz, t * chunk), hi = std::min(nnz, lo + chunk);
      for (int64_t i = lo; i < hi; ++i) cnt[t][keyOf(i)]++;
    }
    int64_t run = 0;
    for (int k = 0; k <= nb; ++k) {
      bstart[k] = run;
      for (int t = 0; t < nth; ++t) {
        const int64_t c = cnt[t][k];
        cnt[t][k] = run;
        run += c;
      }
    }
    bstart[nb + 1] = run;
    order.resize(nnz);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int t = 0; t < nth; ++t) {
      const int64_t lo = std::min(nnz, t * chunk), hi = std::min(nnz, lo + chunk);
      for (int64_t i = lo; i < hi; ++i) order[cnt[t][keyOf(i)]++] = i;
    }
  }
  auto entryAt = [&](int64_t j) -> int64_t { return ordered ? j : order[j]; };

  // 2. per block: sorted distinct offsets with their counts
  std::vector<std::vector<std::pair<int, int>>> pairs(nb);
#pragma omp parallel
  {
    std::vector<int> tmp;
#pragma omp for schedule(dynamic, 256)
    for (int b = 0; b < nb; ++b) {
      tmp.clear();
      for (int64_t j = bstart[b]; j < bstart[b + 1]; ++j) tmp.push_back(offsetOf(entryAt(j)));
      std::sort(tmp.begin(), tmp.end());
      auto& pb = pairs[b];
      for (size_t q = 0; q < tmp.size();) {
        size_t e = q;
        while (e < tmp.size() && tmp[e] == tmp[q]) ++e;
        pb.emplace_back(tmp[q], static_cast<int>(e - q));
        q = e;
      }
    }
  }

  // 3. threshold / max_coo_entries rule, blocks and offsets ascending
  offsetsPerBlock.assign(nb, {});
  size_t storageRequired = 0;
  int64_t cpu_entries = 0;
  std::vector<int> allOffsets;
  for (int b = 0; b < nb; ++b) {
    for (const auto& p : pairs[b]) {
      allOffsets.push_back(p.first);
      if (p.second >= thr || cpu_entries >= maxCoo) {
        offsetsPerBlock[b].push_back(p.first);
        storageRequired += bw;
      } else {
        cpu_entries += p.second;
      }
    }
  }
  std::vector<std::vector<std::pair<int, int>>>().swap(pairs);
  std::sort(allOffsets.begin(), allOffsets.end());
  numberDiagonals = static_cast<int>(std::unique(allOffsets.begin(), allOffsets.end()) - allOffsets.begin());
  std::vector<int>().swap(allOffsets);

  if (storageRequired > static_cast<size_t>(1e11))
    throw std::runtime_error("Excessive memory requirement for HBDIA format");

  // 4. zeroed contiguous storage and block pointers
  hbdia_host::assign_zeros(hbdiaData, storageRequired);
  ptrToBlock.assign(nb, nullptr);
  std::vector<size_t> blockBase(nb, 0);
  {
    size_t pos = 0;
    for (int b = 0; b < nb; ++b) {
      if (!offsetsPerBlock[b].empty()) {
        ptrToBlock[b] = hbdiaData.data() + pos;
        blockBase[b] = pos;
        pos += offsetsPerBlock[b].size() * static_cast<size_t>(bw);
      }
    }
  }

  // 5. histograms
  int maxDiag = 0;
  for (const auto& o : offsetsPerBlock) maxDiag = std::max(maxDiag, static_cast<int>(o.size()));
  histogramBlocks.assign(maxDiag + 1, 0);
  histogramNnz.assign(maxDiag + 1, 0);
  for (const auto& o : offsetsPerBlock) histogramBlocks[o.size()]++;

  // 6. placement, per block in input order; mark fallback entries
  std::vector<uint8_t> toCoo(nnz, 0);
  const int nth = omp_get_max_threads();
  std::vector<std::vector<int64_t>> hnnz(nth, std::vector<int64_t>(maxDiag + 1, 0));
#pragma omp parallel num_threads(nth)
  {
    auto& hl = hnnz[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 256)
    for (int b = 0; b <= nb; ++b) {
      const int64_t j0 = bstart[b], j1 = bstart[b + 1];
      if (b == nb || offsetsPerBlock[b].empty()) {
        for (int64_t j = j0; j < j1; ++j) toCoo[entryAt(j)] = 1;
        continue;
      }
      const auto& ob = offsetsPerBlock[b];
      T* base = hbdiaData.data() + blockBase[b];
      for (int64_t j = j0; j < j1; ++j) {
        const int64_t i = entryAt(j);
        const auto it = std::lower_bound(ob.begin(), ob.end(), offsetOf(i));
        if (it != ob.end() && *it == offsetOf(i)) {
          const size_t oi = static_cast<size_t>(it - ob.begin());
          base[oi * bw + static_cast<size_t>(R[i] % bw)] = V[i];
          hl[ob.size()]++;
        } else {
          toCoo[i] = 1;
        }
      }
    }
  }
  for (int d = 0; d <= maxDiag; ++d) {
    int64_t s = 0;
    for (int t = 0; t < nth; ++t) s += hnnz[t][d];
    histogramNnz[d] = static_cast<int>(s);
  }

  // 7. fallback COO in input order
  {
    const int64_t chunk = (nnz + nth - 1) / nth;
    std::vector<int64_t> cnt(nth + 1, 0);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int t = 0; t < nth; ++t) {
      const int64_t lo = std::min(nnz, t * chunk), hi = std::min(nnz, lo + chunk);
      int64_t c = 0;
      for (int64_t i = lo; i < hi; ++i) c += toCoo[i];
      cnt[t + 1] = c;
    }
    for (int t = 0; t < nth; ++t) cnt[t + 1] += cnt[t];
    cpuRowIndices.assign(cnt[nth], 0);
    cpuColIndices.assign(cnt[nth], 0);
    cpuValues.assign(cnt[nth], T(0));
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int t = 0; t < nth; ++t) {
      const int64_t lo = std::min(nnz, t * chunk), hi = std::min(nnz, lo + chunk);
      int64_t w = cnt[t];
      for (int64_t i = lo; i < hi; ++i)
        if (toCoo[i]) {
          cpuRowIndices[w] = R[i];
          cpuColIndices[w] = C[i];
          cpuValues[w] = V[i];
          ++w;
        }
    }
  }

  // 8. cpuRowPtr with the reference's gap-filling rule
  fill_cpu_row_ptr(cpuRowIndices, numRows, cpuRowPtr);

  if (verbose())
    std::cout << "HBDIA conversion complete: GPU storage " << (nnz - (int64_t)cpuValues.size())
              << " entries, CPU fallback " << cpuValues.size() << " entries" << std::endl;
  hasHBDIA = true;
  return true;
}

// ---------------------------------------------------------------------------
// conversion on the device: the same rules and results as the host conversion
// above (reference src/Format/HBDIA.cpp:507-718), computed by hbdia_convert_coo.

template <typename T>
bool HBDIA<T>::convertToHBDIAFormatOnDevice(int bw, int thr, int maxCoo, bool /*COOisUnique*/,
                                            ExecutionMode execMode, bool hostValues) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "values are moved as 4- or 8-byte words");
  if (hasHBDIA) {
    std::cout << "Matrix is already in HBDIA format" << std::endl;
    return false;
  }
  if (!hasCOO) {
    std::cout << "No COO format available for conversion to HBDIA" << std::endl;
    return false;
  }
  if (values.empty()) {
    std::cout << "Cannot convert empty matrix to HBDIA format" << std::endl;
    return false;
  }
  if (bw <= 0) throw std::invalid_argument("blockWidth must be positive");
  if (values.size() > static_cast<size_t>(0x7fffffff))
    throw std::runtime_error("device conversion takes at most 2^31-1 entries");
  const bool useGlobal = partialMatrix && !globalRowMapping.empty();

  Phases ph;
  cleanupGPUData();
  initializeStreams();
  int *R = nullptr, *C = nullptr, *G = nullptr;
  T* V = nullptr;
  struct Inputs {
    int **r, **c, **g;
    T** v;
    ~Inputs() {
      dev_free(*r);
      dev_free(*c);
      dev_free(*g);
      dev_free(*v);
    }
  } inputs{&R, &C, &G, &V};
  dev_upload(&R, rowIndices.data(), rowIndices.size());
  dev_upload(&C, colIndices.data(), colIndices.size());
  dev_upload(&V, values.data(), values.size());
  if (useGlobal) dev_upload(&G, globalRowMapping.data(), globalRowMapping.size());
  ph.mark("upload COO");

  hbdia_convert_args_t a{};
  a.rows = R;
  a.cols = C;
  a.values = V;
  a.global_rows = G;
  a.nnz = static_cast<int64_t>(values.size());
  a.num_rows = numRows;
  a.num_cols = numCols;
  a.block_width = bw;
  a.threshold = thr;
  a.max_coo_entries = maxCoo;
  a.elem_size = static_cast<int32_t>(sizeof(T));
  a.want_row_sorted_coo = execMode == ExecutionMode::GPU_COO ? 1 : 0;
  a.want_runs3 = 1;
  ConvertResult res;
  CHECK_HBDIA(hbdia_convert_coo(&a, &res.r, sBDIA_));
  ph.mark("device conversion");
  const hbdia_convert_result_t& r = res.r;
  const int nb = r.num_blocks;

  // host state, as the host conversion leaves it
  std::vector<int> flat, starts, placed;
  dev_download(flat, r.offsets, static_cast<size_t>(r.num_pairs));
  dev_download(starts, r.block_start, static_cast<size_t>(nb) + 1);
  dev_download(placed, r.block_placed, static_cast<size_t>(nb));
  std::vector<T> data, cooV;
  std::vector<int> cooR, cooC;
  const size_t nvals = static_cast<size_t>(r.num_pairs) * static_cast<size_t>(bw);
  const bool deferValues = !hostValues && !partialMatrix;
  if (!deferValues) dev_download(data, r.values, nvals);
  dev_download(cooR, r.coo_rows, static_cast<size_t>(r.num_coo));
  dev_download(cooC, r.coo_cols, static_cast<size_t>(r.num_coo));
  dev_download(cooV, r.coo_values, static_cast<size_t>(r.num_coo));
  ph.mark("download result");

  blockWidth = bw;
  threshold = thr;
  max_coo_entries = maxCoo;
  executionMode = execMode;
  numberDiagonals = r.number_diagonals;
  hbdiaData = std::move(data);
  hostValuesPending_ = deferValues;
  pendingValues_ = deferValues ? nvals : 0;
  offsetsPerBlock.assign(nb, {});
  ptrToBlock.assign(nb, nullptr);
  int maxDiag = 0;
  for (int b = 0; b < nb; ++b) {
    offsetsPerBlock[b].assign(flat.begin() + starts[b], flat.begin() + starts[b + 1]);
    if (!deferValues && starts[b + 1] > starts[b])
      ptrToBlock[b] = hbdiaData.data() + static_cast<size_t>(starts[b]) * bw;
    maxDiag = std::max(maxDiag, starts[b + 1] - starts[b]);
  }
  histogramBlocks.assign(maxDiag + 1, 0);
  histogramNnz.assign(maxDiag + 1, 0);
  for (int b = 0; b < nb; ++b) {
    const int k = starts[b + 1] - starts[b];
    histogramBlocks[k]++;
    histogramNnz[k] += k ? placed[b] : 0;
  }
  cpuRowIndices = std::move(cooR);
  cpuColIndices = std::move(cooC);
  cpuValues = std::move(cooV);
  fill_cpu_row_ptr(cpuRowIndices, numRows, cpuRowPtr);
  hasHBDIA = true;
  ph.mark("host structures");
  if (partialMatrix) {  // the reference uploads partial matrices later, with their halo layout
    cleanupGPUData();
    return true;
  }

  // device state, as prepareForGPU leaves it
  hbdiaData_d_ = static_cast<T*>(res.take(res.r.values));
  flattenedOffsets_d_ = res.take(res.r.offsets);
  blockStartIndices_d_ = res.take(res.r.block_start);
  blockSizes_d_ = res.take(res.r.block_sizes);
  runs3Blocks_d_ = res.take(res.r.runs3_blocks);
  runs3Count_ = r.runs3_count;
  if (executionMode == ExecutionMode::GPU_COO && r.num_coo > 0) {
    if (r.coo_sorted) {
      cpuRowIndices_d_ = res.take(res.r.coo_rows);
      cpuColIndices_d_ = res.take(res.r.coo_cols);
      cpuValues_d_ = static_cast<T*>(res.take(res.r.coo_values));
    } else {
      cpuRowIndices_d_ = res.take(res.r.coo_rows_sorted);
      cpuColIndices_d_ = res.take(res.r.coo_cols_sorted);
      cpuValues_d_ = static_cast<T*>(res.take(res.r.coo_values_sorted));
    }
    dev_upload(&cpuRowPtr_d_, cpuRowPtr.data(), cpuRowPtr.size());
    const int64_t groups = (static_cast<int64_t>(numRows) + HBDIA_ROW_GROUP - 1) / HBDIA_ROW_GROUP;
    void* p = nullptr;
    CHECK_HBDIA(hbdia_malloc(&p, static_cast<size_t>(groups + 1) * sizeof(int)));
    cooGroupPtr_d_ = static_cast<int*>(p);
    CHECK_HBDIA(hbdia_coo_group_ptr(cpuRowIndices_d_, r.num_coo, numRows, cooGroupPtr_d_, sBDIA_));
    CHECK_HBDIA(hbdia_stream_synchronize(sBDIA_));
  }
  ph.mark("device structures");
  return true;
}

// ---------------------------------------------------------------------------
// duplicates (reference src/Format/HBDIA.cpp:372-412): std::map semantics --
// sorted by (row, col), duplicates summed in input order.

template <typename T>
void HBDIA<T>::removeCOODuplicates() {
  if (!hasCOO) {
    std::cout << "No COO format available for duplicate removal" << std::endl;
    return;
  }
  if (values.empty()) return;
  const size_t n = values.size();
  std::vector<uint32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0u);
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
    return rowIndices[a] != rowIndices[b] ? rowIndices[a] < rowIndices[b] : colIndices[a] < colIndices[b];
  });
  std::vector<int> nr, nc;
  std::vector<T> nv;
  nr.reserve(n);
  nc.reserve(n);
  nv.reserve(n);
  for (size_t q = 0; q < n; ++q) {
    const uint32_t i = idx[q];
    if (!nr.empty() && nr.back() == rowIndices[i] && nc.back() == colIndices[i]) {
      nv.back() += values[i];
    } else {
      nr.push_back(rowIndices[i]);
      nc.push_back(colIndices[i]);
      nv.push_back(T(0) + values[i]);
    }
  }
  rowIndices = std::move(nr);
  colIndices = std::move(nc);
  values = std::move(nv);
  numNonZeros = static_cast<int>(values.size());
}

// ---------------------------------------------------------------------------
// Matrix Market input (reference src/Format/HBDIA.cpp:161-370)

namespace {

// One entry line as the reference parses it (`is >> row >> col`, then
// `is >> value`, 1 when that fails); lines without two integers do not count.
template <typename T>
bool parseMtxEntry(const std::string& line, int& row, int& col, T& value) {
  std::istringstream is(line);
  if (!(is >> row >> col)) return false;
  if (!(is >> value)) value = T(1);
  return true;
}

struct MtxFile {
  bool ok = false, symmetric = false;
  int rows = 0, cols = 0, nnz = 0;
  std::vector<std::string> lines;  // the lines after the size line
};

MtxFile readMtx(const std::string& filename) {
  MtxFile f;
  std::ifstream file(filename);
  if (!file.is_open()) {
    std::cerr << "Error: Cannot open file " << filename << std::endl;
    return f;
  }
  std::string line;
  while (std::getline(file, line)) {
    if (line[0] == '%') {
      if (line.find("symmetric") != std::string::npos) f.symmetric = true;
    } else {
      break;  // the first non-comment line holds the sizes
    }
  }
  std::istringstream header(line);
  header >> f.rows >> f.cols >> f.nnz;
  while (std::getline(file, line)) f.lines.push_back(std::move(line));
  f.ok = true;
  return f;
}

// Parses the entry lines in parallel and keeps, in file order, the first
// `limit` lines holding two integers.
template <typename T>
void parseMtxEntries(const MtxFile& f, int limit, std::vector<int>& r, std::vector<int>& c, std::vector<T>& v,
                     std::vector<uint8_t>& good) {
  const int64_t n = static_cast<int64_t>(f.lines.size());
  r.assign(n, 0);
  c.assign(n, 0);
  v.assign(n, T(0));
  good.assign(n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; ++i) good[i] = parseMtxEntry<T>(f.lines[i], r[i], c[i], v[i]);
  int taken = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!good[i]) continue;
    if (taken >= limit) good[i] = 0;
    else ++taken;
  }
}

}  // namespace

template <typename T>
bool HBDIA<T>::loadMTX(const std::string& filename) {
  MtxFile f = readMtx(filename);
  if (!f.ok) return false;
  if (f.symmetric && verbose()) std::cout << "Detected symmetric matrix - will expand to full matrix" << std::endl;
  numRows = f.rows;
  numCols = f.cols;
  std::vector<int> r, c;
  std::vector<T> v;
  std::vector<uint8_t> good;
  parseMtxEntries<T>(f, f.nnz, r, c, v, good);
  int original = 0;  // entries are appended to any already held, as the reference does
  for (size_t i = 0; i < good.size(); ++i) {
    if (!good[i]) continue;
    const int row = r[i] - 1, col = c[i] - 1;  // 1-based in the file
    rowIndices.push_back(row);
    colIndices.push_back(col);
    values.push_back(v[i]);
    ++original;
    if (f.symmetric && row != col) {
      rowIndices.push_back(col);
      colIndices.push_back(row);
      values.push_back(v[i]);
    }
  }
  numNonZeros = static_cast<int>(values.size());
  hasCOO = true;
  // sort by (row, col) with the reference's call: std::sort on an index array
  // (its order of equal keys is libstdc++'s, so it is reproduced exactly)
  std::vector<size_t> indices(values.size());
  std::iota(indices.begin(), indices.end(), size_t{0});
  std::sort(indices.begin(), indices.end(), [&](size_t a, size_t b) {
    if (rowIndices[a] != rowIndices[b]) return rowIndices[a] < rowIndices[b];
    return colIndices[a] < colIndices[b];
  });
  std::vector<T> sv(values.size());
  std::vector<int> sr(values.size()), sc(values.size());
  for (size_t i = 0; i < indices.size(); ++i) {
    sv[i] = values[indices[i]];
    sr[i] = rowIndices[indices[i]];
    sc[i] = colIndices[indices[i]];
  }
  values = std::move(sv);
  rowIndices = std::move(sr);
  colIndices = std::move(sc);
  return original > 0;
}

template <typename T>
bool HBDIA<T>::loadMTX(const std::string& filename, int nRows, int nCols) {
  if (nRows <= 0 || nCols <= 0) {
    std::cerr << "Error: Matrix dimensions must be positive" << std::endl;
    return false;
  }
  MtxFile f = readMtx(filename);
  if (!f.ok) return false;
  numRows = nRows;
  numCols = nCols;
  std::vector<int> r, c;
  std::vector<T> v;
  std::vector<uint8_t> good;
  parseMtxEntries<T>(f, f.nnz, r, c, v, good);
  int original = 0;
  for (size_t i = 0; i < good.size(); ++i) {
    if (!good[i]) continue;
    const int row = r[i] - 1, col = c[i] - 1;
    if (row < 0 || row >= nRows)
      throw std::out_of_range("Row index " + std::to_string(row) + " from file is out of range [0, " +
                              std::to_string(nRows - 1) + "]");
    if (col < 0 || col >= nCols)
      throw std::out_of_range("Column index " + std::to_string(col) + " from file is out of range [0, " +
                              std::to_string(nCols - 1) + "]");
    rowIndices.push_back(row);
    colIndices.push_back(col);
    values.push_back(v[i]);
    ++original;
    if (f.symmetric && row != col) {
      rowIndices.push_back(col);
      colIndices.push_back(row);
      values.push_back(v[i]);
    }
  }
  numNonZeros = static_cast<int>(values.size());
  hasCOO = true;
  return original > 0;
}

// ---------------------------------------------------------------------------
// deletion

template <typename T>
void HBDIA<T>::deleteCOOFormat() {
  std::vector<T>().swap(values);
  std::vector<int>().swap(rowIndices);
  std::vector<int>().swap(colIndices);
  hasCOO = false;
}

template <typename T>
void HBDIA<T>::deleteHBDIAFormat() {
  releaseDevice();
  std::vector<T>().swap(hbdiaData);
  ptrToBlock.clear();
  offsetsPerBlock.clear();
  cpuRowIndices.clear();
  cpuColIndices.clear();
  cpuValues.clear();
  cpuRowPtr.clear();
  histogramBlocks.clear();
  histogramNnz.clear();
  numberDiagonals = 0;
  hasHBDIA = false;
  hostValuesPending_ = false;
  pendingValues_ = 0;
  dataRanges.clear();
  vectorOffsets.clear();
  localCOOBuilt_ = false;
  haloLeft_ = haloRight_ = 0;
  haloColumns_.clear();
  localCooRows_.clear();
  localCooCols_.clear();
  localCooValues_.clear();
  localCooFolded_ = 0;
}

template <typename T>
void HBDIA<T>::deleteMatrix() {
  deleteCOOFormat();
  deleteHBDIAFormat();
  nnzPerRow.clear();
  numRows = numCols = numNonZeros = 0;
}

// ---------------------------------------------------------------------------
// device upload (reference src/Format/HBDIA.cpp:720-816) and streams (:858-876)

template <typename T>
void HBDIA<T>::materializeHostValues() const {
  if (!hostValuesPending_) return;
  if (!hbdiaData_d_ && pendingValues_ > 0)
    throw std::logic_error("HBDIA: blocked values are neither on the host nor on the device");
  std::vector<T> data;
  dev_download(data, hbdiaData_d_, pendingValues_);
  hbdiaData = std::move(data);
  ptrToBlock.assign(offsetsPerBlock.size(), nullptr);
  size_t pos = 0;
  for (size_t b = 0; b < offsetsPerBlock.size(); ++b) {
    if (!offsetsPerBlock[b].empty()) ptrToBlock[b] = hbdiaData.data() + pos;
    pos += offsetsPerBlock[b].size() * static_cast<size_t>(blockWidth);
  }
  hostValuesPending_ = false;
}

template <typename T>
void HBDIA<T>::prepareForGPU() {
  if (!hasHBDIA) return;
  materializeHostValues();  // the upload below starts from the host values
  cleanupGPUData();
  // a partial matrix runs rank-local: halo columns go through the COO part
  if (partialMatrix) buildLocalCOO();
  const std::vector<int>& fbRows = partialMatrix ? localCooRows_ : cpuRowIndices;
  const std::vector<int>& fbCols = partialMatrix ? localCooCols_ : cpuColIndices;
  const std::vector<T>& fbValues = partialMatrix ? localCooValues_ : cpuValues;

  const int nb = static_cast<int>(offsetsPerBlock.size());
  std::vector<int> flattenedOffsets, blockStartIndices, blockSizes;
  blockStartIndices.reserve(static_cast<size_t>(nb) + 1);
  blockSizes.reserve(nb);
  size_t total = 0;
  for (const auto& o : offsetsPerBlock) total += o.size();
  if (total > static_cast<size_t>(0x7fffffff))
    throw std::runtime_error("more than 2^31-1 (block, offset) pairs");
  flattenedOffsets.reserve(total);
  for (int b = 0; b < nb; ++b) {
    blockStartIndices.push_back(static_cast<int>(flattenedOffsets.size()));
    blockSizes.push_back(static_cast<int>(offsetsPerBlock[b].size()));
    flattenedOffsets.insert(flattenedOffsets.end(), offsetsPerBlock[b].begin(), offsetsPerBlock[b].end());
  }
  blockStartIndices.push_back(static_cast<int>(flattenedOffsets.size()));
  // layout hint for the kernels: the blocks whose offsets are runs of three
  // consecutive values (a multiple of three offsets, each triple consecutive)
  std::vector<uint32_t> runs3Bits;
  runs3Count_ = 0;
  if ((blockWidth % HBDIA_ROW_GROUP == 0 || blockWidth == HBDIA_ROW_GROUP / 2) && nb > 0) {
    runs3Bits.assign((static_cast<size_t>(nb) + 31) / 32, 0u);
    for (int b = 0; b < nb; ++b) {
      const auto& o = offsetsPerBlock[b];
      bool runs3 = !o.empty() && o.size() % 3 == 0;
      for (size_t t = 0; runs3 && t + 2 < o.size(); t += 3)
        runs3 = o[t + 1] == o[t] + 1 && o[t + 2] == o[t + 1] + 1;
      if (runs3) {
        runs3Bits[b >> 5] |= 1u << (b & 31);
        ++runs3Count_;
      }
    }
  }
  // the kernels read a chunk's offsets as one wide load: keep the tail readable
  std::vector<int> paddedOffsets(flattenedOffsets);
  paddedOffsets.resize(flattenedOffsets.size() + HBDIA_OFFSETS_PAD, 0);
  initializeStreams();

  dev_upload(&hbdiaData_d_, hbdiaData.data(), hbdiaData.size());
  dev_upload(&flattenedOffsets_d_, paddedOffsets.data(), paddedOffsets.size());
  if (runs3Count_ > 0) dev_upload(&runs3Blocks_d_, runs3Bits.data(), runs3Bits.size());
  dev_upload(&blockStartIndices_d_, blockStartIndices.data(), blockStartIndices.size());
  dev_upload(&blockSizes_d_, blockSizes.data(), blockSizes.size());

  if (executionMode == ExecutionMode::GPU_COO && !fbValues.empty()) {
    // The fused kernel needs row-sorted entries; keep the per-row entry order.
    const bool sorted = std::is_sorted(fbRows.begin(), fbRows.end());
    if (sorted) {
      dev_upload(&cpuRowIndices_d_, fbRows.data(), fbRows.size());
      dev_upload(&cpuColIndices_d_, fbCols.data(), fbCols.size());
      dev_upload(&cpuValues_d_, fbValues.data(), fbValues.size());
    } else {
      std::vector<uint32_t> idx(fbValues.size());
      std::iota(idx.begin(), idx.end(), 0u);
      std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return fbRows[a] < fbRows[b]; });
      std::vector<int> r(idx.size()), c(idx.size());
      std::vector<T> v(idx.size());
      for (size_t q = 0; q < idx.size(); ++q) {
        r[q] = fbRows[idx[q]];
        c[q] = fbCols[idx[q]];
        v[q] = fbValues[idx[q]];
      }
      dev_upload(&cpuRowIndices_d_, r.data(), r.size());
      dev_upload(&cpuColIndices_d_, c.data(), c.size());
      dev_upload(&cpuValues_d_, v.data(), v.size());
    }
    if (!partialMatrix) dev_upload(&cpuRowPtr_d_, cpuRowPtr.data(), cpuRowPtr.size());
    const int64_t groups = (static_cast<int64_t>(numRows) + HBDIA_ROW_GROUP - 1) / HBDIA_ROW_GROUP;
    void* p = nullptr;
    CHECK_HBDIA(hbdia_malloc(&p, static_cast<size_t>(groups + 1) * sizeof(int)));
    cooGroupPtr_d_ = static_cast<int*>(p);
    CHECK_HBDIA(hbdia_coo_group_ptr(cpuRowIndices_d_, static_cast<int64_t>(fbValues.size()), numRows,
                                    cooGroupPtr_d_, sBDIA_));
    CHECK_HBDIA(hbdia_stream_synchronize(sBDIA_));
  }
}

template <typename T>
void HBDIA<T>::releaseDevice() noexcept {
  if (streamsInitialized_) {
    (void)hbdia_stream_synchronize(sBDIA_);
    (void)hbdia_stream_synchronize(sADD_);
  }
  dev_free(hbdiaData_d_);
  dev_free(flattenedOffsets_d_);
  dev_free(runs3Blocks_d_);
  runs3Count_ = 0;
  dev_free(blockStartIndices_d_);
  dev_free(blockSizes_d_);
  dev_free(cpuRowIndices_d_);
  dev_free(cpuColIndices_d_);
  dev_free(cpuValues_d_);
  dev_free(cpuRowPtr_d_);
  dev_free(cooGroupPtr_d_);
  if (streamsInitialized_) {
    (void)hbdia_stream_destroy(sBDIA_);
    (void)hbdia_stream_destroy(sD2H_);
    (void)hbdia_stream_destroy(sH2D_);
    (void)hbdia_stream_destroy(sCOO_);
    (void)hbdia_stream_destroy(sADD_);
    (void)hbdia_event_destroy(bdiaEvent_);
    (void)hbdia_event_destroy(cooEvent_);
    sBDIA_ = sD2H_ = sH2D_ = sCOO_ = sADD_ = nullptr;
    bdiaEvent_ = cooEvent_ = nullptr;
    streamsInitialized_ = false;
  }
  cusparseInitialized_ = false;
}

template <typename T>
void HBDIA<T>::cleanupGPUData() {
  materializeHostValues();  // device-only values must survive the release
  releaseDevice();
}

template <typename T>
void HBDIA<T>::initializeStreams() {
  if (streamsInitialized_) return;
  CHECK_HBDIA(hbdia_stream_create(&sBDIA_));
  CHECK_HBDIA(hbdia_stream_create(&sD2H_));
  CHECK_HBDIA(hbdia_stream_create(&sH2D_));
  CHECK_HBDIA(hbdia_stream_create(&sCOO_));
  CHECK_HBDIA(hbdia_stream_create(&sADD_));
  CHECK_HBDIA(hbdia_event_create(&bdiaEvent_));
  CHECK_HBDIA(hbdia_event_create(&cooEvent_));
  streamsInitialized_ = true;
}

template <typename T>
void HBDIA<T>::cleanupStreams() {
  if (!streamsInitialized_) return;
  (void)hbdia_stream_synchronize(sBDIA_);
  (void)hbdia_stream_synchronize(sADD_);
  (void)hbdia_stream_destroy(sBDIA_);
  (void)hbdia_stream_destroy(sD2H_);
  (void)hbdia_stream_destroy(sH2D_);
  (void)hbdia_stream_destroy(sCOO_);
  (void)hbdia_stream_destroy(sADD_);
  (void)hbdia_event_destroy(bdiaEvent_);
  (void)hbdia_event_destroy(cooEvent_);
  sBDIA_ = sD2H_ = sH2D_ = sCOO_ = sADD_ = nullptr;
  bdiaEvent_ = cooEvent_ = nullptr;
  streamsInitialized_ = false;
}

template <typename T>
void HBDIA<T>::synchronize() const {
  if (!streamsInitialized_) {
    CHECK_HBDIA(hbdia_device_synchronize());
    return;
  }
  CHECK_HBDIA(hbdia_stream_synchronize(sBDIA_));
  CHECK_HBDIA(hbdia_stream_synchronize(sD2H_));
  CHECK_HBDIA(hbdia_stream_synchronize(sH2D_));
  CHECK_HBDIA(hbdia_stream_synchronize(sCOO_));
  CHECK_HBDIA(hbdia_stream_synchronize(sADD_));
}

template <typename T>
size_t HBDIA<T>::storedBytes() const {
  size_t nOff = 0;
  for (const auto& o : offsetsPerBlock) nOff += o.size();
  return getNumStoredValues() * sizeof(T) + nOff * sizeof(int) + (offsetsPerBlock.size() + 1) * sizeof(int) +
         cpuValues.size() * (sizeof(T) + 2 * sizeof(int));
}

// ---------------------------------------------------------------------------
// printing (short dumps; the reference's HBDIAPrinter is a debug aid)

template <typename T>
void HBDIA<T>::print() const {
  std::cout << "HBDIA<" << sizeof(T) * 8 << "-bit> " << numRows << "x" << numCols << ", nnz "
            << numNonZeros << (hasCOO ? ", COO" : "") << (hasHBDIA ? ", HBDIA" : "");
  if (hasHBDIA)
    std::cout << " (bw " << blockWidth << ", thr " << threshold << ", " << offsetsPerBlock.size()
              << " blocks, " << getNumStoredValues() << " stored values, " << cpuValues.size()
              << " fallback entries)";
  std::cout << std::endl;
}

template <typename T>
void HBDIA<T>::printCOO() const {
  const size_t n = std::min<size_t>(values.size(), 1000);
  for (size_t i = 0; i < n; ++i)
    std::cout << "(" << rowIndices[i] << ", " << colIndices[i] << ") " << values[i] << "\n";
  if (n < values.size()) std::cout << "... " << values.size() - n << " more\n";
  std::cout.flush();
}

template <typename T>
void HBDIA<T>::printHBDIA() const {
  if (!hasHBDIA) {
    std::cout << "Matrix is not in HBDIA format" << std::endl;
    return;
  }
  materializeHostValues();
  const size_t nbShow = std::min<size_t>(offsetsPerBlock.size(), 64);
  for (size_t b = 0; b < nbShow; ++b) {
    std::cout << "block " << b << " offsets:";
    for (int o : offsetsPerBlock[b]) std::cout << " " << o;
    std::cout << "\n";
    if (!ptrToBlock[b]) continue;
    for (size_t oi = 0; oi < offsetsPerBlock[b].size(); ++oi) {
      std::cout << "  [" << offsetsPerBlock[b][oi] << "]";
      for (int l = 0; l < blockWidth; ++l) std::cout << " " << ptrToBlock[b][oi * blockWidth + l];
      std::cout << "\n";
    }
  }
  std::cout << "fallback entries: " << cpuValues.size() << std::endl;
}

template class HBDIA<double>;
template class HBDIA<float>;
template class HBDIA<int>;
