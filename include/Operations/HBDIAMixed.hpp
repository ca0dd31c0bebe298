// HBDIAMixed.hpp -- reduced-precision storage for an FP64 HBDIA matrix.
//
// No counterpart in the reference (it stores and computes in one type).
// Holds device copies of the blocked-DIA values and of the fallback values
// in FP32 or BF16 next to the FP64 matrix (which stays resident: HBM capacity
// is plentiful, bandwidth is what the product is bound by).  spmv() computes
// y = A x with x and y in FP64: every row is accumulated in FP32 from the
// low-precision values, and rows whose cancellation ratio sum|a x| / |sum a x|
// exceeds the threshold (1e4 by default) are recomputed in FP64 with exactly
// the arithmetic of hbdiaSpMV's GPU branch.  The ratio is taken against the
// smallest |y| consistent with the FP32 sum and its error bound, so a kept row
// has relative error below (storage unit roundoff + n 2^-24) * threshold.
#ifndef HBDIAMIXED_HPP
#define HBDIAMIXED_HPP

#include <cstdint>

#include "../Format/HBDIA.hpp"
#include "../Format/HBDIAVector.hpp"

enum class StoragePrecision { FP32 = 1, BF16 = 2 };

class HBDIAMixed {
 public:
  // The matrix must be converted (GPU_COO) and prepared; it must outlive this object.
  HBDIAMixed(const HBDIA<double>& matrix, StoragePrecision precision, double cancellationRatio = 1e4);
  ~HBDIAMixed();
  HBDIAMixed(const HBDIAMixed&) = delete;
  HBDIAMixed& operator=(const HBDIAMixed&) = delete;

  // y = A x on the matrix's BDIA stream (asynchronous, like hbdiaSpMV).
  void spmv(const HBDIAVector<double>& x, HBDIAVector<double>& y);
  // The same on raw device arrays of numRows doubles.
  void spmv(const double* x, double* y);
  // Rows recomputed in FP64 by the last spmv (waits for it).
  int64_t lastFp64Rows() const;
  StoragePrecision precision() const { return precision_; }
  double cancellationRatio() const { return ratio_; }
  // Bytes of the low-precision copies.
  size_t storedBytes() const { return bytes_; }

 private:
  const HBDIA<double>& m_;
  StoragePrecision precision_;
  double ratio_;
  void* values_ = nullptr;
  void* cooValues_ = nullptr;
  void* workspace_ = nullptr;
  size_t bytes_ = 0;
};

#endif  // HBDIAMIXED_HPP
