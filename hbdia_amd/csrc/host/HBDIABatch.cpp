// HBDIABatch.cpp -- pooled many-matrix HBDIA SpMV and its autotuner.
#include "Operations/HBDIABatch.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>

#include "types.hpp"

namespace {

template <typename T>
struct BatchOps;

template <>
struct BatchOps<double> {
  static int batch(const hbdia_batch_t* b, const hbdia_batch_schedule_t* s, const double* x, double* y,
                   hbdia_stream_t st) {
    return hbdia_batch_spmv_f64(b, s, x, y, st);
  }
  static constexpr auto fused = hbdia_hbdia_spmv_f64;
  static constexpr auto bdia = hbdia_bdia_spmv_f64;
};

template <>
struct BatchOps<float> {
  static int batch(const hbdia_batch_t* b, const hbdia_batch_schedule_t* s, const float* x, float* y,
                   hbdia_stream_t st) {
    return hbdia_batch_spmv_f32(b, s, x, y, st);
  }
  static constexpr auto fused = hbdia_hbdia_spmv_f32;
  static constexpr auto bdia = hbdia_bdia_spmv_f32;
};

template <typename U>
void* upload(const std::vector<U>& v) {
  void* p = nullptr;
  const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(U);
  CHECK_HBDIA(hbdia_malloc(&p, bytes));
  if (!v.empty()) CHECK_HBDIA(hbdia_memcpy(p, v.data(), v.size() * sizeof(U)));
  return p;
}

constexpr int64_t kInt32Max = std::numeric_limits<int32_t>::max();

}  // namespace

template <typename T>
HBDIABatch<T>::HBDIABatch(const std::vector<const HBDIA<T>*>& ms) {
  const int M = static_cast<int>(ms.size());
  desc_.resize(M);
  int64_t nvals = 0, noffs = 0, nblk = 0, ncoo = 0, ngp = 0;
  for (int m = 0; m < M; ++m) {
    const HBDIA<T>* A = ms[m];
    if (!A) throw std::invalid_argument("HBDIABatch: null matrix");
    if (!A->getHasHBDIA()) throw std::invalid_argument("HBDIABatch: matrix " + std::to_string(m) + " not converted");
    if (A->isPartialMatrix())
      throw std::invalid_argument("HBDIABatch: matrix " + std::to_string(m) + " is a rank-local (partial) matrix");
    if (A->getNumRows() != A->getNumCols())
      throw std::invalid_argument("HBDIABatch: matrix " + std::to_string(m) + " is not square");
    const int bw = A->getBlockWidth();
    if (bw <= 0 || bw % HBDIA_ROW_GROUP != 0)
      throw std::invalid_argument("HBDIABatch: block widths must be multiples of 64");
    const int n = A->getNumRows();
    const int64_t groups = (static_cast<int64_t>(n) + HBDIA_ROW_GROUP - 1) / HBDIA_ROW_GROUP;
    int64_t pairs = 0;
    for (const auto& o : A->getOffsetsPerBlock()) pairs += static_cast<int64_t>(o.size());
    hbdia_batch_desc_t& d = desc_[m];
    d.values = nvals;
    d.rows = totalRows_;
    d.first_tile = numTiles_;
    d.offsets = static_cast<int32_t>(noffs);
    d.block_start = static_cast<int32_t>(nblk);
    d.coo = static_cast<int32_t>(ncoo);
    d.coo_nnz = static_cast<int32_t>(A->getCpuValues().size());
    d.group_ptr = static_cast<int32_t>(ngp);
    d.num_rows = n;
    d.num_blocks = A->getNumBlocks();
    d.block_width = bw;
    nvals += static_cast<int64_t>(A->getNumStoredValues());
    noffs += pairs;
    nblk += d.num_blocks + 1;
    ncoo += d.coo_nnz;
    ngp += groups + 1;
    numTiles_ += groups;
    totalRows_ += n;
    if (noffs > kInt32Max || nblk > kInt32Max || ncoo > kInt32Max || ngp > kInt32Max || numTiles_ > kInt32Max ||
        totalRows_ > kInt32Max)
      throw std::invalid_argument("HBDIABatch: pooled index arrays exceed 2^31 entries");
  }
  cooNnz_ = ncoo;

  // pools with the padding the batch kernel reads past the last entry
  std::vector<T> values(static_cast<size_t>(nvals) + HBDIA_ROW_GROUP, T(0));
  std::vector<int32_t> offsets(static_cast<size_t>(noffs) + HBDIA_OFFSETS_PAD, 0);
  std::vector<int32_t> blockStart(static_cast<size_t>(nblk));
  std::vector<int32_t> cooRows(static_cast<size_t>(ncoo) + HBDIA_BATCH_COO_PAD, 0);
  std::vector<int32_t> cooCols(static_cast<size_t>(ncoo) + HBDIA_BATCH_COO_PAD, 0);
  std::vector<T> cooVals(static_cast<size_t>(ncoo) + HBDIA_BATCH_COO_PAD, T(0));
  std::vector<int32_t> groupPtr(static_cast<size_t>(ngp));
  std::vector<hbdia_batch_tile_t> tiles(static_cast<size_t>(numTiles_));
  tileWork_.assign(static_cast<size_t>(numTiles_) + 1, 0);
  int64_t nonzeros = 0;
  bool x32 = true;
  int64_t allLines = 0, longLines = 0;  // value lines of all blocks, of blocks with long lists

#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nonzeros) reduction(&& : x32)
  for (int m = 0; m < M; ++m) {
    const HBDIA<T>& A = *ms[m];
    const hbdia_batch_desc_t& d = desc_[m];
    const auto& data = A.getHBDIAData();
    std::copy(data.begin(), data.end(), values.begin() + d.values);
    for (const T& v : data) nonzeros += (v != T(0));
    // flattened offsets and block starts relative to the matrix's own offsets
    const auto& opb = A.getOffsetsPerBlock();
    int32_t* op = offsets.data() + d.offsets;
    int32_t* bs = blockStart.data() + d.block_start;
    int32_t k = 0;
    for (int b = 0; b < d.num_blocks; ++b) {
      bs[b] = k;
      for (int o : opb[b]) op[k++] = o;
      const int64_t lines = static_cast<int64_t>(opb[b].size());
      allLines += lines;
      if (lines > HBDIA_BATCH_CHUNK) longLines += lines;
    }
    bs[d.num_blocks] = k;
    // fallback: a stable row sort keeps each row's entries in input order
    const auto& r = A.getCpuRowIndices();
    const auto& c = A.getCpuColIndices();
    const auto& v = A.getCpuValues();
    std::vector<uint32_t> idx(r.size());
    std::iota(idx.begin(), idx.end(), 0u);
    if (!std::is_sorted(r.begin(), r.end()))
      std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return r[a] < r[b]; });
    for (size_t q = 0; q < idx.size(); ++q) {
      cooRows[d.coo + q] = r[idx[q]];
      cooCols[d.coo + q] = c[idx[q]];
      cooVals[d.coo + q] = v[idx[q]];
    }
    const int32_t* cr = cooRows.data() + d.coo;
    const int64_t groups = (static_cast<int64_t>(d.num_rows) + HBDIA_ROW_GROUP - 1) / HBDIA_ROW_GROUP;
    int32_t* gp = groupPtr.data() + d.group_ptr;
    for (int64_t g = 0; g <= groups; ++g)
      gp[g] = static_cast<int32_t>(std::lower_bound(cr, cr + d.coo_nnz, static_cast<int32_t>(g * HBDIA_ROW_GROUP)) - cr);
    // tile work (value lines, fallback entries, x and y) for the contiguous split
    for (int64_t g = 0; g < groups; ++g) {
      const int b = static_cast<int>(g * HBDIA_ROW_GROUP / d.block_width);
      const int64_t lines = b < d.num_blocks ? static_cast<int64_t>(opb[b].size()) : 0;
      tileWork_[d.first_tile + g + 1] = lines * HBDIA_ROW_GROUP * static_cast<int64_t>(sizeof(T)) +
                                        static_cast<int64_t>(gp[g + 1] - gp[g]) * (sizeof(T) + 8) +
                                        2 * HBDIA_ROW_GROUP * static_cast<int64_t>(sizeof(T));
    }
    if (static_cast<int64_t>(d.num_rows) * static_cast<int64_t>(sizeof(T)) >= (int64_t{1} << 32)) x32 = false;
  }
  for (int64_t t = 0; t < numTiles_; ++t) tileWork_[t + 1] += tileWork_[t];
  CHECK_HBDIA(hbdia_batch_make_tiles(desc_.data(), M, offsets.data(), blockStart.data(), groupPtr.data(),
                                     static_cast<int32_t>(sizeof(T)), tiles.data(), numTiles_));
  bdiaNnz_ = nonzeros;
  storedBytes_ = nvals * static_cast<int64_t>(sizeof(T)) + noffs * 4 + ncoo * (8 + static_cast<int64_t>(sizeof(T))) +
                 numTiles_ * static_cast<int64_t>(sizeof(hbdia_batch_tile_t));

  try {
    pools_[0] = upload(tiles);
    pools_[1] = nullptr;
    pools_[2] = upload(values);
    pools_[3] = upload(offsets);
    pools_[4] = upload(blockStart);
    pools_[5] = upload(cooRows);
    pools_[6] = upload(cooCols);
    pools_[7] = upload(cooVals);
    pools_[8] = upload(groupPtr);
    CHECK_HBDIA(hbdia_stream_create(&stream_));
    CHECK_HBDIA(hbdia_device_compute_units(&computeUnits_));
  } catch (...) {
    release();
    throw;
  }
  dev_.tiles = static_cast<const hbdia_batch_tile_t*>(pools_[0]);
  dev_.wave_tiles = nullptr;
  dev_.values = pools_[2];
  dev_.offsets = static_cast<const int32_t*>(pools_[3]);
  dev_.has_coo = ncoo > 0;
  dev_.coo_rows = static_cast<const int32_t*>(pools_[5]);
  dev_.coo_cols = static_cast<const int32_t*>(pools_[6]);
  dev_.coo_values = pools_[7];
  dev_.num_tiles = numTiles_;
  dev_.x32 = x32;
  // the kernel that pairs long-list tiles holds more registers (FP32: 5 waves
  // per SIMD instead of 6): taken when those tiles hold most of the values
  dev_.long_lists = 2 * longLines > allLines;
  try {
    setSchedule(BatchSchedule{});
  } catch (...) {
    release();
    throw;
  }
}

template <typename T>
HBDIABatch<T>::~HBDIABatch() {
  release();
}

template <typename T>
void HBDIABatch<T>::release() noexcept {
  if (stream_) (void)hbdia_stream_synchronize(stream_);
  if (graph_) (void)hbdia_graph_destroy(graph_);
  graph_ = nullptr;
  for (void*& p : pools_) {
    if (p) (void)hbdia_free(p);
    p = nullptr;
  }
  if (waveTiles_) (void)hbdia_free(waveTiles_);
  waveTiles_ = nullptr;
  if (stream_) (void)hbdia_stream_destroy(stream_);
  stream_ = nullptr;
}

template <typename T>
int64_t HBDIABatch<T>::algorithmicBytes() const {
  const int64_t s = sizeof(T);
  return bdiaNnz_ * s + cooNnz_ * (s + 8) + 2 * totalRows_ * s;
}

template <typename T>
int HBDIABatch<T>::residentWorkgroups(int kind) const {
  int perCu = 0;
  CHECK_HBDIA(hbdia_batch_occupancy(&dev_, kind, static_cast<int32_t>(sizeof(T)), &perCu));
  return std::max(1, perCu) * std::max(1, computeUnits_);
}

template <typename T>
void HBDIABatch<T>::setSchedule(const BatchSchedule& s) {
  if (s.kind != HBDIA_BATCH_INTERLEAVED && s.kind != HBDIA_BATCH_CONTIGUOUS && s.kind != HBDIA_BATCH_XCD)
    throw std::invalid_argument("HBDIABatch::setSchedule: unknown kind");
  if (s.tilesPerStep < 1 || s.workgroups < 0)
    throw std::invalid_argument("HBDIABatch::setSchedule: tilesPerStep >= 1 and workgroups >= 0");
  int wg = s.workgroups > 0 ? s.workgroups : residentWorkgroups(s.kind);
  if (s.kind == HBDIA_BATCH_XCD) wg = std::max(8, wg / 8 * 8);  // the same number on every XCD
  if (s.kind != HBDIA_BATCH_INTERLEAVED) {
    // contiguous: wave w takes the tiles whose work prefix falls in
    // [w, w+1) * total / W; XCD split: the same cut into W = 8 ranges
    const int64_t W = s.kind == HBDIA_BATCH_XCD ? 8 : static_cast<int64_t>(wg) * 4;
    std::vector<int32_t> bounds(static_cast<size_t>(W) + 1);
    const long double total = static_cast<long double>(tileWork_.back());
    for (int64_t w = 0; w <= W; ++w) {
      const int64_t target = static_cast<int64_t>(std::llround(total * w / W));
      bounds[w] = static_cast<int32_t>(std::lower_bound(tileWork_.begin() + 1, tileWork_.end(), target) -
                                       (tileWork_.begin() + 1));
    }
    bounds[0] = 0;
    bounds[W] = static_cast<int32_t>(numTiles_);
    for (int64_t w = 1; w <= W; ++w) bounds[w] = std::max(bounds[w], bounds[w - 1]);
    void* p = upload(bounds);
    if (waveTiles_) (void)hbdia_free(waveTiles_);
    waveTiles_ = static_cast<int32_t*>(p);
    dev_.wave_tiles = waveTiles_;
  } else {
    dev_.wave_tiles = nullptr;
  }
  schedule_ = s;
  schedule_.workgroups = wg;
  launchWorkgroups_ = wg;
  if (graph_) {  // a recorded launch carries the old schedule
    (void)hbdia_graph_destroy(graph_);
    graph_ = nullptr;
  }
}

template <typename T>
void HBDIABatch<T>::spmv(const T* x, T* y) const {
  const hbdia_batch_schedule_t s{schedule_.kind, schedule_.tilesPerStep, launchWorkgroups_};
  CHECK_HBDIA(BatchOps<T>::batch(&dev_, &s, x, y, stream_));
}

template <typename T>
void HBDIABatch<T>::spmvMatrix(int m, const T* x, T* y, hbdia_stream_t stream) const {
  const hbdia_batch_desc_t& d = desc_.at(m);
  const T* vals = static_cast<const T*>(pools_[2]);
  const int32_t* offs = static_cast<const int32_t*>(pools_[3]);
  const int32_t* blk = static_cast<const int32_t*>(pools_[4]);
  if (d.coo_nnz > 0)
    CHECK_HBDIA(BatchOps<T>::fused(vals + d.values, offs + d.offsets, blk + d.block_start, d.num_blocks, d.block_width,
                                   static_cast<const int32_t*>(pools_[5]) + d.coo,
                                   static_cast<const int32_t*>(pools_[6]) + d.coo,
                                   static_cast<const T*>(pools_[7]) + d.coo,
                                   static_cast<const int32_t*>(pools_[8]) + d.group_ptr, x, y, d.num_rows, stream));
  else
    CHECK_HBDIA(BatchOps<T>::bdia(vals + d.values, offs + d.offsets, blk + d.block_start, x, y, d.num_rows,
                                  d.num_blocks, d.block_width, stream));
}

template <typename T>
void HBDIABatch<T>::spmvPerMatrix(const T* x, T* y) const {
  for (int m = 0; m < numMatrices(); ++m) spmvMatrix(m, x + desc_[m].rows, y + desc_[m].rows, stream_);
}

template <typename T>
void HBDIABatch<T>::captureGraph(const T* x, T* y) {
  if (graph_) {
    (void)hbdia_graph_destroy(graph_);
    graph_ = nullptr;
  }
  CHECK_HBDIA(hbdia_graph_begin_capture(stream_));
  const hbdia_batch_schedule_t s{schedule_.kind, schedule_.tilesPerStep, launchWorkgroups_};
  const int rc = BatchOps<T>::batch(&dev_, &s, x, y, stream_);
  hbdia_graph_t g = nullptr;
  const int rc2 = hbdia_graph_end_capture(stream_, &g);
  CHECK_HBDIA(rc);
  CHECK_HBDIA(rc2);
  graph_ = g;
}

template <typename T>
void HBDIABatch<T>::launchGraph() const {
  if (!graph_) throw std::runtime_error("HBDIABatch::launchGraph: no captured graph");
  CHECK_HBDIA(hbdia_graph_launch(graph_, stream_));
}

template <typename T>
void HBDIABatch<T>::synchronize() const {
  CHECK_HBDIA(hbdia_stream_synchronize(stream_));
}

// ---------------------------------------------------------------------------

namespace {

template <typename T>
void convertAll(const std::vector<HBDIA<T>*>& ms, int bw, int threshold) {
  const int M = static_cast<int>(ms.size());
  std::string err;
#pragma omp parallel for schedule(dynamic, 1)
  for (int m = 0; m < M; ++m) {
    try {
      // conversion refuses a matrix already in HBDIA form: drop the previous layout first
      if (ms[m]->getHasHBDIA()) ms[m]->deleteHBDIAFormat();
      if (!ms[m]->convertToHBDIAFormatOnHost(bw, threshold))
        throw std::runtime_error("matrix " + std::to_string(m) + " could not be converted");
    } catch (const std::exception& e) {
#pragma omp critical
      err = e.what();
    }
  }
  if (!err.empty()) throw std::runtime_error("autotuneBatch: conversion failed: " + err);
}

template <typename T>
double timeLaunches(const HBDIABatch<T>& b, const T* x, T* y, int reps) {
  for (int i = 0; i < 2; ++i) b.spmv(x, y);
  hbdia_event_t e0 = nullptr, e1 = nullptr;
  CHECK_HBDIA(hbdia_event_create(&e0));
  CHECK_HBDIA(hbdia_event_create(&e1));
  float ms = 0.0f;
  try {
    CHECK_HBDIA(hbdia_event_record(e0, b.stream()));
    for (int i = 0; i < reps; ++i) b.spmv(x, y);
    CHECK_HBDIA(hbdia_event_record(e1, b.stream()));
    CHECK_HBDIA(hbdia_event_synchronize(e1));
    CHECK_HBDIA(hbdia_event_elapsed_ms(&ms, e0, e1));
  } catch (...) {
    (void)hbdia_event_destroy(e0);
    (void)hbdia_event_destroy(e1);
    throw;
  }
  (void)hbdia_event_destroy(e0);
  (void)hbdia_event_destroy(e1);
  return static_cast<double>(ms) / reps;
}

}  // namespace

template <typename T>
std::unique_ptr<HBDIABatch<T>> autotuneBatch(const std::vector<HBDIA<T>*>& matrices,
                                             const std::vector<int>& blockWidths, int threshold, int reps,
                                             std::vector<BatchTuneResult>* table) {
  if (matrices.empty()) throw std::invalid_argument("autotuneBatch: no matrices");
  if (blockWidths.empty()) throw std::invalid_argument("autotuneBatch: no block widths");
  for (int bw : blockWidths)
    if (bw <= 0 || bw % HBDIA_ROW_GROUP != 0)
      throw std::invalid_argument("autotuneBatch: block widths must be multiples of 64");
  if (reps < 1) throw std::invalid_argument("autotuneBatch: reps >= 1");
  int64_t rows = 0;
  for (const HBDIA<T>* m : matrices) rows += m->getNumRows();
  std::vector<T> hx(static_cast<size_t>(rows));
  for (int64_t i = 0; i < rows; ++i) hx[i] = static_cast<T>(1.0 + 1e-3 * static_cast<double>(i % 1000));
  void* x = upload(hx);
  void* y = nullptr;
  std::unique_ptr<HBDIABatch<T>> best;
  try {
    CHECK_HBDIA(hbdia_malloc(&y, std::max<size_t>(hx.size(), 1) * sizeof(T)));
    CHECK_HBDIA(hbdia_memset(y, 0, std::max<size_t>(hx.size(), 1) * sizeof(T), nullptr));
    // one pooled batch per block width, kept until the choice is made
    std::vector<std::unique_ptr<HBDIABatch<T>>> batches;
    struct Cand {
      size_t batch;
      BatchSchedule schedule;
      double ms;
    };
    std::vector<Cand> all;
    for (size_t bi = 0; bi < blockWidths.size(); ++bi) {
      const int bw = blockWidths[bi];
      convertAll(matrices, bw, threshold);
      std::vector<const HBDIA<T>*> cm(matrices.begin(), matrices.end());
      batches.push_back(std::make_unique<HBDIABatch<T>>(cm));
      HBDIABatch<T>& batch = *batches.back();
      if (bi == 0) {
        // the clocks ramp up over the first launches after an idle GPU: about
        // 50 ms of launches before any candidate is timed
        double spent = 0.0;
        for (int i = 0; i < 64 && spent < 50.0; ++i)
          spent += reps * timeLaunches(batch, static_cast<const T*>(x), static_cast<T*>(y), reps);
      }
      std::vector<BatchSchedule> cands;
      for (int kind : {HBDIA_BATCH_INTERLEAVED, HBDIA_BATCH_XCD, HBDIA_BATCH_CONTIGUOUS}) {
        const int R = batch.residentWorkgroups(kind);
        for (int wg : {R, std::max(1, R / 2)}) {
          if (kind != HBDIA_BATCH_CONTIGUOUS)
            for (int tps : {1, 2, 4, 8}) cands.push_back(BatchSchedule{kind, tps, wg});
          else
            cands.push_back(BatchSchedule{kind, 1, wg});
        }
      }
      for (const BatchSchedule& c : cands) {
        batch.setSchedule(c);
        const double ms = timeLaunches(batch, static_cast<const T*>(x), static_cast<T*>(y), reps);
        if (table) table->push_back(BatchTuneResult{bw, batch.schedule(), ms});
        all.push_back(Cand{bi, batch.schedule(), ms});
      }
    }
    // the fastest few again, in turn over three rounds, each taking its best
    // round: one timing of `reps` launches moves by a few per cent
    std::stable_sort(all.begin(), all.end(), [](const Cand& p, const Cand& q) { return p.ms < q.ms; });
    const size_t K = std::min<size_t>(6, all.size());
    std::vector<double> again(K, std::numeric_limits<double>::infinity());
    for (int round = 0; round < 3; ++round)
      for (size_t k = 0; k < K; ++k) {
        HBDIABatch<T>& batch = *batches[all[k].batch];
        batch.setSchedule(all[k].schedule);
        again[k] = std::min(again[k], timeLaunches(batch, static_cast<const T*>(x), static_cast<T*>(y), reps));
      }
    const size_t pick = static_cast<size_t>(std::min_element(again.begin(), again.end()) - again.begin());
    const size_t bb = all[pick].batch;
    best = std::move(batches[bb]);
    best->setSchedule(all[pick].schedule);
    batches.clear();
    if (bb + 1 != blockWidths.size()) convertAll(matrices, blockWidths[bb], threshold);
  } catch (...) {
    (void)hbdia_free(x);
    if (y) (void)hbdia_free(y);
    throw;
  }
  (void)hbdia_device_synchronize();
  (void)hbdia_free(x);
  (void)hbdia_free(y);
  return best;
}

template class HBDIABatch<double>;
template class HBDIABatch<float>;
template std::unique_ptr<HBDIABatch<double>> autotuneBatch<double>(const std::vector<HBDIA<double>*>&,
                                                                  const std::vector<int>&, int, int,
                                                                  std::vector<BatchTuneResult>*);
template std::unique_ptr<HBDIABatch<float>> autotuneBatch<float>(const std::vector<HBDIA<float>*>&,
                                                                const std::vector<int>&, int, int,
                                                                std::vector<BatchTuneResult>*);
