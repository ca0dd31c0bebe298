// HBDIABatch.hpp -- many independent HBDIA matrices multiplied in one
// persistent launch, and the autotuner that picks their block width and tile
// schedule.
//
// No reference counterpart: the reference multiplies one matrix per
// hbdiaSpMV call (include/Operations/HBDIASpMV.hpp:8-21).  Each matrix of a
// batch keeps the reference's HBDIA layout (values [block][offset][lane],
// ascending offsets per block, row-sorted fallback), so each matrix's rows of
// y equal what hbdiaSpMV(matrix, x, y, false, true) gives for that matrix
// alone, bit for bit (the fallback, if any, is computed on the device and
// fused, as in hbdiaSpMV's GPU branch).
//
// x and y are pooled: matrix m reads x[rowOffset(m) + c] and writes
// y[rowOffset(m) + r].  Launches run on the batch's own stream and are
// asynchronous; synchronize() before reading y on the host.
#ifndef HBDIA_BATCH_HPP
#define HBDIA_BATCH_HPP

#include <cstdint>
#include <memory>
#include <vector>

#include "../Format/HBDIA.hpp"
#include "../hbdia_hip.h"

// Which waves take which tiles (see hbdia_batch_spmv_* in hbdia_hip.h).
struct BatchSchedule {
  int kind = HBDIA_BATCH_INTERLEAVED;
  int tilesPerStep = 1;
  int workgroups = 0;  // 0: every workgroup the device holds at once
};

template <typename T>
class HBDIABatch {
 public:
  // Pools the host HBDIA arrays of already converted matrices and uploads
  // them; the matrices are not referenced afterwards.  Every block width must
  // be a multiple of 64 and every matrix square.
  explicit HBDIABatch(const std::vector<const HBDIA<T>*>& matrices);
  ~HBDIABatch();
  HBDIABatch(const HBDIABatch&) = delete;
  HBDIABatch& operator=(const HBDIABatch&) = delete;

  int numMatrices() const { return static_cast<int>(desc_.size()); }
  int64_t totalRows() const { return totalRows_; }
  int64_t rowOffset(int m) const { return desc_.at(m).rows; }
  int numRows(int m) const { return desc_.at(m).num_rows; }
  int blockWidth(int m) const { return desc_.at(m).block_width; }
  int64_t numTiles() const { return numTiles_; }
  int64_t bdiaNnz() const { return bdiaNnz_; }  // nonzeros stored in the blocked part
  int64_t cooNnz() const { return cooNnz_; }
  // Bytes a product must move at least: nonzero values, fallback entries with
  // their two int32 indices, x and y once.
  int64_t algorithmicBytes() const;
  int64_t flops() const { return 2 * (bdiaNnz_ + cooNnz_); }
  int64_t storedBytes() const { return storedBytes_; }

  // Workgroups of the selected kernel that the device holds at once.
  int residentWorkgroups(int kind) const;
  void setSchedule(const BatchSchedule& schedule);
  const BatchSchedule& schedule() const { return schedule_; }

  // y = A_m x for every matrix, one persistent launch (x, y device pools).
  void spmv(const T* x, T* y) const;
  // The same with one single-matrix launch per matrix (hbdia_hbdia_spmv_* or
  // hbdia_bdia_spmv_* on the matrix's part of the pools).
  void spmvPerMatrix(const T* x, T* y) const;
  // One single-matrix launch of matrix m on `stream`; x and y point at the
  // matrix's own first row (not at the pools).
  void spmvMatrix(int m, const T* x, T* y, hbdia_stream_t stream) const;
  // Device view of the pools (tile records, values, offsets, fallback) and
  // the placement of matrix m in them.
  const hbdia_batch_t& deviceBatch() const { return dev_; }
  const hbdia_batch_desc_t& desc(int m) const { return desc_.at(m); }
  int64_t tilesOf(int m) const { return (static_cast<int64_t>(desc_.at(m).num_rows) + 63) / 64; }
  // Records spmv(x, y) into a graph on the batch stream; launchGraph replays it.
  void captureGraph(const T* x, T* y);
  void launchGraph() const;
  hbdia_stream_t stream() const { return stream_; }
  void synchronize() const;

 private:
  void release() noexcept;

  std::vector<hbdia_batch_desc_t> desc_;
  std::vector<int64_t> tileWork_;  // prefix of per-tile work, for the contiguous split
  int64_t totalRows_ = 0, numTiles_ = 0, bdiaNnz_ = 0, cooNnz_ = 0, storedBytes_ = 0;
  hbdia_batch_t dev_{};
  BatchSchedule schedule_{};
  int launchWorkgroups_ = 1;
  void* pools_[9] = {};  // tiles, (unused), values, offsets, block_start, coo rows / cols / values, group_ptr
  int32_t* waveTiles_ = nullptr;
  hbdia_stream_t stream_ = nullptr;
  hbdia_graph_t graph_ = nullptr;
  int computeUnits_ = 0;
};

// One candidate the autotuner timed.
struct BatchTuneResult {
  int blockWidth;
  BatchSchedule schedule;
  double ms;  // mean time of one launch
};

// Converts every matrix at each candidate block width (host conversion with
// `threshold`), pools them, times each tile schedule over `reps` launches on
// device-resident x and y (after about 50 ms of launches that let the clocks
// ramp up), times the six fastest again in turn over three rounds, and returns
// the batch of the fastest (block width, schedule) by its best round.  The
// pools of every block width are held on the device until the choice is made.  The matrices are left converted on the host at the
// chosen block width (a previous HBDIA layout, device copies included, is
// dropped).  `table`, if given, receives every candidate's time.
template <typename T>
std::unique_ptr<HBDIABatch<T>> autotuneBatch(const std::vector<HBDIA<T>*>& matrices,
                                             const std::vector<int>& blockWidths, int threshold, int reps,
                                             std::vector<BatchTuneResult>* table);

#endif  // HBDIA_BATCH_HPP
