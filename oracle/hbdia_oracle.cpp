// hbdia_oracle.cpp -- CPU restatement of the HBDIA SpMV path (the oracle).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
// oracle/build/libhbdia_oracle.so and may be loaded only by tests/, by
// __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  No product
// code links, loads or calls it.
//
// Every function restates one piece of the reference algorithm and cites the
// reference file:line it follows (paths relative to the reference checkout).
// Floating-point rules:
//   * the reference device kernel accumulates `sum += v * x` in ascending offset
//     order; nvcc contracts this to a fused multiply-add by default, so the
//     restatement uses std::fma in the same order;
//   * the reference CPU COO loop is built with -O3 -ffp-contract=fast
//     (CMakeLists.txt:39), which contracts `row_sum += val * x` to an FMA; the
//     restatement again uses std::fma, sequentially in entry order.
// This file is compiled with -ffp-contract=off so that no other contraction
// happens behind these explicit calls.
//
// Parity pins (tests/test_oracle.py):
//   generator   -> closed-form nnz (3n-2)^3, weight multisets, row sortedness,
//                  noise-entry fixture, duplicate-freedom of noise columns;
//   conversion  -> closed-form per-block offset counts for stencils, the
//                  hand-derived layout of the reference's print-test matrix,
//                  partition invariant (every entry lands exactly once);
//   SpMV        -> closed forms y = 18.22 (r+1) on interior rows of the 8^3
//                  stencil with x = i+1, y(0) = -73.14, y(511) = 11610.51,
//                  A*1 = row sums, long-double brute-force COO SpMV.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <numeric>
#include <random>
#include <set>
#include <vector>
#include <omp.h>

namespace {

template <typename T> inline T fma_t(T a, T b, T c);
template <> inline double fma_t<double>(double a, double b, double c) { return std::fma(a, b, c); }
template <> inline float fma_t<float>(float a, float b, float c) { return std::fmaf(a, b, c); }

template <typename T> struct Coo {
  std::vector<int> rows, cols;
  std::vector<T> vals;
  int nrows = 0, ncols = 0;
};

// Restates HBDIA<T>::create3DStencil27Point (src/Format/HBDIA.cpp:1416-1607):
// same mt19937(42+iteration) stream (:1448), same distribution objects and call
// order (noise rows :1457-1466, noise columns and values :1518-1566), same loop
// nest (the outer index walks nz but is bounds-checked against nx, :1474-1490),
// same weights (:1500-1512) and the same unstable std::sort of a permutation by
// row (:1572-1588).
template <typename T>
Coo<T>* stencil27(int nx, int ny, int nz, double noise, int iteration) {
  auto* m = new Coo<T>();
  const int totalNodes = nx * ny * nz;
  auto getIndex = [&](int i, int j, int k) -> int { return k * nx * ny + j * nx + i; };
  std::vector<int>& rowIndices = m->rows;
  std::vector<int>& colIndices = m->cols;
  std::vector<T>& values = m->vals;

  std::mt19937 gen(42 + iteration);  // RAND_SEED + iteration, include/types.hpp:20
  std::uniform_int_distribution<int> colDist(0, totalNodes - 1);
  std::uniform_real_distribution<double> valueDist(-1.0, 1.0);
  std::vector<int> noiseEntriesPerRow(totalNodes, 0);
  if (noise > 0.0) {
    long long stencilEntries = totalNodes * 27;
    int totalTarget = static_cast<int>(static_cast<double>(stencilEntries) * noise);
    std::uniform_int_distribution<int> rowDist(0, totalNodes - 1);
    for (int e = 0; e < totalTarget; ++e) noiseEntriesPerRow[rowDist(gen)]++;
  }
  for (int i = 0; i < nz; i++) {
    for (int j = 0; j < ny; j++) {
      for (int k = 0; k < nx; k++) {
        const int centerIdx = getIndex(i, j, k);
        for (int di = -1; di <= 1; di++)
          for (int dj = -1; dj <= 1; dj++)
            for (int dk = -1; dk <= 1; dk++) {
              int ni = i + di, nj = j + dj, nk = k + dk;
              if (ni >= 0 && ni < nx && nj >= 0 && nj < ny && nk >= 0 && nk < nz) {
                rowIndices.push_back(centerIdx);
                colIndices.push_back(getIndex(ni, nj, nk));
                int nzc = (di != 0) + (dj != 0) + (dk != 0);
                if (nzc == 0) values.push_back(static_cast<T>(26.1));
                else if (nzc == 1) values.push_back(static_cast<T>(-1.1));
                else if (nzc == 2) values.push_back(static_cast<T>(-0.1));
                else values.push_back(static_cast<T>(-0.01));
              }
            }
        int target = noiseEntriesPerRow[centerIdx];
        if (target > 0) {
          std::set<int> usedCols;
          for (int dk = -1; dk <= 1; dk++)
            for (int dj = -1; dj <= 1; dj++)
              for (int di = -1; di <= 1; di++) {
                int ni = i + di, nj = j + dj, nk = k + dk;
                if (ni >= 0 && ni < nx && nj >= 0 && nj < ny && nk >= 0 && nk < nz)
                  usedCols.insert(getIndex(ni, nj, nk));
              }
          int added = 0, attempts = 0;
          const int maxAttempts = target * 10 + 100;
          while (added < target && attempts < maxAttempts) {
            attempts++;
            int col = colDist(gen);
            if (usedCols.find(col) != usedCols.end()) continue;
            usedCols.insert(col);
            rowIndices.push_back(centerIdx);
            colIndices.push_back(col);
            values.push_back(static_cast<T>(valueDist(gen)));
            added++;
          }
        }
      }
    }
  }
  std::vector<size_t> perm(rowIndices.size());
  std::iota(perm.begin(), perm.end(), 0);
  std::sort(perm.begin(), perm.end(),
            [&](size_t a, size_t b) { return rowIndices[a] < rowIndices[b]; });
  std::vector<int> sr(perm.size()), sc(perm.size());
  std::vector<T> sv(perm.size());
  for (size_t i = 0; i < perm.size(); i++) {
    sr[i] = rowIndices[perm[i]];
    sc[i] = colIndices[perm[i]];
    sv[i] = values[perm[i]];
  }
  m->rows = std::move(sr);
  m->cols = std::move(sc);
  m->vals = std::move(sv);
  m->nrows = totalNodes;
  m->ncols = totalNodes;
  return m;
}

template <typename T> struct Hb {
  int bw = 0, nb = 0, numberDiagonals = 0;
  std::vector<T> data;                 // hbdiaData
  std::vector<std::vector<int>> offs;  // offsetsPerBlock
  std::vector<int> flatOffs, blkStart; // prepareForGPU flattening
  std::vector<int> cooR, cooC, rowPtr;
  std::vector<T> cooV;
  std::vector<int> histBlocks, histNnz;
};

// Restates HBDIA<T>::convertToHBDIAFormat (src/Format/HBDIA.cpp:507-718),
// followed by the flattening of prepareForGPU (:736-765).  G, when given, is
// the global row mapping of a partial matrix: offsets are col - G[row] (:543-553,
// :648-660) while block and lane stay local.  std::map per block (:540-560), threshold / max_coo_entries rule
// (:571-583), zero-filled contiguous storage (:600-621), histograms (:623-640),
// placement by assignment or COO fallback in input order with the cpuRowPtr
// gap-filling rule (:643-708).
template <typename T>
Hb<T>* convert(const int* rowIndices, const int* colIndices, const T* values, int64_t nnz,
               int numRows, int numCols, int blockWidth, int threshold, int max_coo_entries,
               const int* G = nullptr) {
  auto grow = [&](int r) { return G ? G[r] : r; };
  auto* h = new Hb<T>();
  h->bw = blockWidth;
  const int maxNumBlocks = (std::min(numRows, numCols) + blockWidth - 1) / blockWidth;
  h->nb = maxNumBlocks;
  std::vector<std::map<int, int>> cnt(maxNumBlocks);
  for (int64_t i = 0; i < nnz; ++i) {
    int block = rowIndices[i] / blockWidth;
    int offset = colIndices[i] - grow(rowIndices[i]);
    if (block < maxNumBlocks) cnt[block][offset]++;
  }
  std::set<int> uniq;
  size_t storage = 0;
  int cpu_entries = 0;
  h->offs.assign(maxNumBlocks, {});
  for (int b = 0; b < maxNumBlocks; ++b) {
    for (const auto& p : cnt[b]) {
      uniq.insert(p.first);
      if (p.second >= threshold || cpu_entries >= max_coo_entries) {
        h->offs[b].push_back(p.first);
        storage += blockWidth;
      } else {
        cpu_entries += p.second;
      }
    }
    std::sort(h->offs[b].begin(), h->offs[b].end());
  }
  h->numberDiagonals = static_cast<int>(uniq.size());
  h->data.assign(storage, T(0));
  std::vector<size_t> ptr(maxNumBlocks, SIZE_MAX);
  size_t pos = 0;
  for (int b = 0; b < maxNumBlocks; ++b)
    if (!h->offs[b].empty()) { ptr[b] = pos; pos += h->offs[b].size() * blockWidth; }
  int maxDiag = 0;
  for (auto& o : h->offs) maxDiag = std::max(maxDiag, (int)o.size());
  h->histBlocks.assign(maxDiag + 1, 0);
  h->histNnz.assign(maxDiag + 1, 0);
  for (auto& o : h->offs) h->histBlocks[o.size()]++;
  h->rowPtr.assign(numRows + 1, 0);
  int currentRow = -1, nextOffset = 0;
  auto to_coo = [&](int r, int c, T v) {
    if (r != currentRow) {
      for (int q = currentRow + 1; q <= r; q++) h->rowPtr[q] = nextOffset;
      currentRow = r;
    }
    h->cooR.push_back(r);
    h->cooC.push_back(c);
    h->cooV.push_back(v);
    nextOffset++;
  };
  for (int64_t i = 0; i < nnz; ++i) {
    int r = rowIndices[i], c = colIndices[i];
    T v = values[i];
    int block = r / blockWidth, offset = c - grow(r), lane = r % blockWidth;
    if (block < maxNumBlocks && !h->offs[block].empty()) {
      auto& ob = h->offs[block];
      auto it = std::find(ob.begin(), ob.end(), offset);
      if (it != ob.end()) {
        int oi = static_cast<int>(it - ob.begin());
        h->data[ptr[block] + (size_t)oi * blockWidth + lane] = v;
        h->histNnz[ob.size()]++;
      } else {
        to_coo(r, c, v);
      }
    } else {
      to_coo(r, c, v);
    }
  }
  for (int q = currentRow + 1; q <= numRows; q++) h->rowPtr[q] = nextOffset;
  for (int b = 0; b < maxNumBlocks; ++b) {
    h->blkStart.push_back(static_cast<int>(h->flatOffs.size()));
    for (int o : h->offs[b]) h->flatOffs.push_back(o);
  }
  h->blkStart.push_back(static_cast<int>(h->flatOffs.size()));
  return h;
}

// Restates bdia_spmv_kernel (src/Operations/HBDIASpMV.cu:7-74), one row per
// iteration: rows past the last block or in an empty block are not written
// (:25-36); zero values are skipped (:50-53); columns outside [0, numRows) are
// skipped (:62-67); the sum runs over ascending offset index (:42) as an FMA.
template <typename T>
void bdia_spmv(const T* hbdiaData, const int* flattenedOffsets, const int* blockStartIndices,
               const T* x, T* y, int64_t numRows, int numBlocks, int blockWidth) {
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < numRows; ++row) {
    int64_t currentBlock = row / blockWidth;
    if (currentBlock >= numBlocks) continue;
    int64_t blockStart = blockStartIndices[currentBlock];
    int blockSize = blockStartIndices[currentBlock + 1] - (int)blockStart;
    if (blockSize == 0) continue;
    int64_t lane = row % blockWidth;
    T sum = T(0);
    for (int o = 0; o < blockSize; ++o) {
      T v = hbdiaData[blockStart * blockWidth + (int64_t)o * blockWidth + lane];
      if (v == T(0)) continue;
      int64_t col = row + flattenedOffsets[blockStart + o];
      if (col < 0 || col >= numRows) continue;
      sum = fma_t<T>(v, x[col], sum);
    }
    y[row] = sum;
  }
}

// Restates cpu_coo_spmv_kernel (src/Operations/HBDIASpMV.cpp:35-99): static
// row split over threads, lower_bound on the row-sorted fallback rows, a
// sequential per-row sum in entry order, and a store only for rows that have
// entries.
template <typename T>
void cpu_coo_spmv(const int* rows, const int* cols, const T* vals, int64_t nnz, const T* x,
                  int numRows, T* res) {
  if (nnz == 0) return;
#pragma omp parallel
  {
    const int nt = omp_get_num_threads(), tid = omp_get_thread_num();
    const int rpt = (numRows + nt - 1) / nt;
    const int start_row = tid * rpt;
    const int end_row = std::min(start_row + rpt, numRows);
    size_t s = 0, e = (size_t)nnz;
    if (start_row > 0) s = std::lower_bound(rows, rows + nnz, start_row) - rows;
    if (end_row < numRows) e = std::lower_bound(rows, rows + nnz, end_row) - rows;
    T row_sum = T(0);
    int cur = -1;
    for (size_t i = s; i < e; ++i) {
      const int r = rows[i];
      if (r != cur) {
        if (cur >= start_row && cur < end_row) res[cur] = row_sum;
        cur = r;
        row_sum = T(0);
      }
      row_sum = fma_t<T>(vals[i], x[cols[i]], row_sum);
    }
    if (cur >= start_row && cur < end_row) res[cur] = row_sum;
  }
}

// Restates hbdiaSpMV (src/Operations/HBDIASpMV.cpp:186-251) for one GPU with
// separate (non-unified) vectors: BDIA kernel into y, COO result into the
// persistent COO buffer (only rows with entries are overwritten), then
// vector_add_kernel y[i] += coo[i] over all rows (HBDIASpMV.cu:76-88), but
// only when the fallback is non-empty (:224, :232).
template <typename T>
void hbdia_spmv(const T* vals, const int* offs, const int* blkStart, int nb, int bw,
                const int* cooR, const int* cooC, const T* cooV, int64_t nnzCoo, const T* x,
                T* y, T* cooBuf, int64_t numRows) {
  bdia_spmv<T>(vals, offs, blkStart, x, y, numRows, nb, bw);
  if (nnzCoo > 0) {
    cpu_coo_spmv<T>(cooR, cooC, cooV, nnzCoo, x, (int)numRows, cooBuf);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < numRows; ++i) y[i] += cooBuf[i];
  }
}

}  // namespace

extern "C" {

// ---- generator ----
void* orc_stencil27_f64(int nx, int ny, int nz, double noise, int iteration) {
  return stencil27<double>(nx, ny, nz, noise, iteration);
}
void* orc_stencil27_f32(int nx, int ny, int nz, double noise, int iteration) {
  return stencil27<float>(nx, ny, nz, noise, iteration);
}
int64_t orc_coo_nnz_f64(void* h) { return (int64_t)static_cast<Coo<double>*>(h)->vals.size(); }
int64_t orc_coo_nnz_f32(void* h) { return (int64_t)static_cast<Coo<float>*>(h)->vals.size(); }
int orc_coo_nrows_f64(void* h) { return static_cast<Coo<double>*>(h)->nrows; }
int orc_coo_nrows_f32(void* h) { return static_cast<Coo<float>*>(h)->nrows; }
void orc_coo_get_f64(void* h, int* r, int* c, double* v) {
  auto* m = static_cast<Coo<double>*>(h);
  std::copy(m->rows.begin(), m->rows.end(), r);
  std::copy(m->cols.begin(), m->cols.end(), c);
  std::copy(m->vals.begin(), m->vals.end(), v);
}
void orc_coo_get_f32(void* h, int* r, int* c, float* v) {
  auto* m = static_cast<Coo<float>*>(h);
  std::copy(m->rows.begin(), m->rows.end(), r);
  std::copy(m->cols.begin(), m->cols.end(), c);
  std::copy(m->vals.begin(), m->vals.end(), v);
}
void orc_coo_free_f64(void* h) { delete static_cast<Coo<double>*>(h); }
void orc_coo_free_f32(void* h) { delete static_cast<Coo<float>*>(h); }

// ---- conversion ----
void* orc_convert_f64(const int* r, const int* c, const double* v, int64_t nnz, int nrows,
                      int ncols, int bw, int thr, int max_coo) {
  return convert<double>(r, c, v, nnz, nrows, ncols, bw, thr, max_coo);
}
void* orc_convert_f32(const int* r, const int* c, const float* v, int64_t nnz, int nrows,
                      int ncols, int bw, int thr, int max_coo) {
  return convert<float>(r, c, v, nnz, nrows, ncols, bw, thr, max_coo);
}

void* orc_convert_g_f64(const int* r, const int* c, const double* v, int64_t nnz, int nrows, int ncols, int bw,
                        int thr, int max_coo, const int* G) {
  return convert<double>(r, c, v, nnz, nrows, ncols, bw, thr, max_coo, G);
}
void* orc_convert_g_f32(const int* r, const int* c, const float* v, int64_t nnz, int nrows, int ncols, int bw,
                        int thr, int max_coo, const int* G) {
  return convert<float>(r, c, v, nnz, nrows, ncols, bw, thr, max_coo, G);
}
// sizes[0]=num_blocks [1]=num_values [2]=num_offsets [3]=num_coo [4]=numberDiagonals
// [5]=hist_len [6]=rowptr_len
#define ORC_SIZES(T, h)                                             \
  auto* m = static_cast<Hb<T>*>(h);                                 \
  sizes[0] = m->nb;                                                 \
  sizes[1] = (int64_t)m->data.size();                               \
  sizes[2] = (int64_t)m->flatOffs.size();                           \
  sizes[3] = (int64_t)m->cooV.size();                               \
  sizes[4] = m->numberDiagonals;                                    \
  sizes[5] = (int64_t)m->histBlocks.size();                         \
  sizes[6] = (int64_t)m->rowPtr.size();
void orc_hb_sizes_f64(void* h, int64_t* sizes) { ORC_SIZES(double, h) }
void orc_hb_sizes_f32(void* h, int64_t* sizes) { ORC_SIZES(float, h) }
#define ORC_GET(T, h)                                                        \
  auto* m = static_cast<Hb<T>*>(h);                                          \
  std::copy(m->data.begin(), m->data.end(), vals);                           \
  std::copy(m->flatOffs.begin(), m->flatOffs.end(), offs);                   \
  std::copy(m->blkStart.begin(), m->blkStart.end(), blk);                    \
  std::copy(m->cooR.begin(), m->cooR.end(), cr);                             \
  std::copy(m->cooC.begin(), m->cooC.end(), cc);                             \
  std::copy(m->cooV.begin(), m->cooV.end(), cv);                             \
  std::copy(m->rowPtr.begin(), m->rowPtr.end(), rp);                         \
  std::copy(m->histBlocks.begin(), m->histBlocks.end(), hb);                 \
  std::copy(m->histNnz.begin(), m->histNnz.end(), hn);
void orc_hb_get_f64(void* h, double* vals, int* offs, int* blk, int* cr, int* cc, double* cv,
                    int* rp, int* hb, int* hn) { ORC_GET(double, h) }
void orc_hb_get_f32(void* h, float* vals, int* offs, int* blk, int* cr, int* cc, float* cv,
                    int* rp, int* hb, int* hn) { ORC_GET(float, h) }
void orc_hb_free_f64(void* h) { delete static_cast<Hb<double>*>(h); }
void orc_hb_free_f32(void* h) { delete static_cast<Hb<float>*>(h); }

// ---- SpMV ----
void orc_bdia_spmv_f64(const double* vals, const int* offs, const int* blk, const double* x,
                       double* y, int64_t nrows, int nb, int bw) {
  bdia_spmv<double>(vals, offs, blk, x, y, nrows, nb, bw);
}
void orc_bdia_spmv_f32(const float* vals, const int* offs, const int* blk, const float* x,
                       float* y, int64_t nrows, int nb, int bw) {
  bdia_spmv<float>(vals, offs, blk, x, y, nrows, nb, bw);
}
void orc_cpu_coo_spmv_f64(const int* r, const int* c, const double* v, int64_t nnz,
                          const double* x, int nrows, double* res) {
  cpu_coo_spmv<double>(r, c, v, nnz, x, nrows, res);
}
void orc_cpu_coo_spmv_f32(const int* r, const int* c, const float* v, int64_t nnz,
                          const float* x, int nrows, float* res) {
  cpu_coo_spmv<float>(r, c, v, nnz, x, nrows, res);
}
void orc_hbdia_spmv_f64(const double* vals, const int* offs, const int* blk, int nb, int bw,
                        const int* cr, const int* cc, const double* cv, int64_t nnzCoo,
                        const double* x, double* y, double* cooBuf, int64_t nrows) {
  hbdia_spmv<double>(vals, offs, blk, nb, bw, cr, cc, cv, nnzCoo, x, y, cooBuf, nrows);
}
void orc_hbdia_spmv_f32(const float* vals, const int* offs, const int* blk, int nb, int bw,
                        const int* cr, const int* cc, const float* cv, int64_t nnzCoo,
                        const float* x, float* y, float* cooBuf, int64_t nrows) {
  hbdia_spmv<float>(vals, offs, blk, nb, bw, cr, cc, cv, nnzCoo, x, y, cooBuf, nrows);
}

// Independent check: y = A x over the raw COO in long double, summing
// duplicates, with no knowledge of the blocked layout.
void orc_coo_spmv_ld(const int* r, const int* c, const double* v, int64_t nnz, const double* x,
                     int64_t nrows, double* y) {
  std::vector<long double> acc(nrows, 0.0L);
  for (int64_t i = 0; i < nnz; ++i)
    acc[r[i]] += (long double)v[i] * (long double)x[c[i]];
  for (int64_t i = 0; i < nrows; ++i) y[i] = (double)acc[i];
}

int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
