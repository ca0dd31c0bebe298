// HBDIAMixed.hpp -- reduced-precision storage for an FP64 HBDIA matrix.
//
// No counterpart in the reference (it stores and computes in one type).
// Holds device copies of the blocked-DIA values and of the fallback values
// in FP32 or BF16 next to the FP64 matrix (which stays resident: HBM capacity
// is plentiful, bandwidth is what the product is bound by).  spmv() computes
// y = A x with x and y in FP64: every row is accumulated in FP32 from the
// low-precision values, and rows whose cancellation ratio sum|a x| / |sum a x|
// exceeds the threshold are recomputed in FP64 with exactly the arithmetic of
// hbdiaSpMV's GPU branch.  The ratio is taken against the smallest |y|
// consistent with the FP32 sum and its error bound, so a kept row of n terms
// has relative error at most u_eff(n) * threshold, with
// u_eff(n) = 1.0001 (u_store + (n + 2) 2^-24) (u_store = 2^-24 for FP32,
// 2^-8 + 2^-24 for BF16; the factor covers the FP32 evaluation of the bound).  relErrorBound() states that figure for the matrix's longest row.
// Without an explicit threshold the constructor derives it from a target
// relative error (kDefaultRelTolerance): FP32 storage then keeps nearly every
// row, BF16 storage recomputes most of them, since BF16 values alone carry a
// relative error of 2^-8.
#ifndef HBDIAMIXED_HPP
#define HBDIAMIXED_HPP

#include <cstdint>
#include <vector>

#include "../Format/HBDIA.hpp"
#include "../Format/HBDIAVector.hpp"

enum class StoragePrecision { FP32 = 1, BF16 = 2 };

class HBDIAMixed {
 public:
  static constexpr double kDefaultRelTolerance = 1e-2;
  // The matrix must be converted (GPU_COO) and prepared; it must outlive this object.
  // Threshold ratioForTolerance(matrix, precision, kDefaultRelTolerance).
  HBDIAMixed(const HBDIA<double>& matrix, StoragePrecision precision);
  HBDIAMixed(const HBDIA<double>& matrix, StoragePrecision precision, double cancellationRatio);
  // The largest threshold whose guarantee u_eff(n) * threshold stays within
  // relTol for every row of the matrix (n: its most terms in one row).
  static double ratioForTolerance(const HBDIA<double>& matrix, StoragePrecision precision, double relTol);
  // u_eff(n) for n terms.
  static double effectiveUnitRoundoff(StoragePrecision precision, int64_t terms);
  // Terms of the matrix's longest row: the widest block's offsets plus the most fallback entries of a row.
  static int64_t maxRowTerms(const HBDIA<double>& matrix);
  ~HBDIAMixed();
  HBDIAMixed(const HBDIAMixed&) = delete;
  HBDIAMixed& operator=(const HBDIAMixed&) = delete;

  // y = A x on the matrix's BDIA stream (asynchronous, like hbdiaSpMV).
  void spmv(const HBDIAVector<double>& x, HBDIAVector<double>& y);
  // The same on raw device arrays of numRows doubles.
  void spmv(const double* x, double* y);
  // Rows recomputed in FP64 by the last spmv (waits for it).
  int64_t lastFp64Rows() const;
  // Which rows those were: bit r of word g is row 64 g + r (waits for the last spmv).
  std::vector<uint64_t> lastFp64Masks() const;
  StoragePrecision precision() const { return precision_; }
  double cancellationRatio() const { return ratio_; }
  // Guaranteed relative error of every row kept in FP32: u_eff(maxRowTerms) * cancellationRatio.
  double relErrorBound() const { return effectiveUnitRoundoff(precision_, maxTerms_) * ratio_; }
  // Bytes of the low-precision copies.
  size_t storedBytes() const { return bytes_; }

 private:
  const HBDIA<double>& m_;
  StoragePrecision precision_;
  double ratio_;
  int64_t maxTerms_ = 0;
  void* values_ = nullptr;
  void* cooValues_ = nullptr;
  void* workspace_ = nullptr;
  size_t bytes_ = 0;
};

#endif  // HBDIAMIXED_HPP
