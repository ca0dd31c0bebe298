// simple_hbdia_example.cpp -- single-GPU use of the drop-in API, the same
// calls as the single-GPU half of the reference's simple_hbdia_example.cpp
// (:15-33): an 8x8x8 27-point stencil, blocked with width 32 and threshold 2,
// x[i] = i + 1, COO fallback on the CPU.
//
// Prints one line per checked quantity and exits non-zero if the closed forms
// of this input do not hold: interior rows give 18.22 (r + 1), row 0 gives
// -73.14 and row 511 gives 11610.51.
#include <cmath>
#include <cstdio>
#include <vector>

#include "Format/HBDIA.hpp"
#include "Format/HBDIAVector.hpp"
#include "Operations/HBDIASpMV.hpp"

int main() {
  HBDIA<DataType> matrix;
  matrix.create3DStencil27Point(8, 8, 8, 0.0);
  matrix.convertToHBDIAFormat(32, 2);

  std::vector<DataType> inputData(matrix.getNumRows());
  for (int i = 0; i < matrix.getNumRows(); i++) inputData[i] = static_cast<DataType>(i + 1);
  HBDIAVector<DataType> vecX(inputData);
  HBDIAVector<DataType> vecY(std::vector<DataType>(matrix.getNumRows(), 0.0));

  const bool execCOOCPU = true, execCOOGPU = false;
  hbdiaSpMV(matrix, vecX, vecY, execCOOCPU, execCOOGPU);
  matrix.synchronize();
  const std::vector<DataType> y = vecY.download();

  int bad = 0;
  double worst = 0.0;
  for (int r = 0; r < 512; ++r) {
    const int i = r % 8, j = (r / 8) % 8, k = r / 64;
    if (i < 1 || i > 6 || j < 1 || j > 6 || k < 1 || k > 6) continue;
    const double want = 18.22 * (r + 1);
    const double rel = std::fabs(y[r] - want) / std::fabs(want);
    worst = std::fmax(worst, rel);
    if (rel > 1e-12) ++bad;
  }
  const double e0 = std::fabs(y[0] + 73.14) / 73.14;
  const double e511 = std::fabs(y[511] - 11610.51) / 11610.51;
  std::printf("rows=%d nnz=%d blocks=%d fallback=%zu\n", matrix.getNumRows(), matrix.getNumNonZeros(),
              matrix.getNumBlocks(), matrix.getCpuValues().size());
  std::printf("interior max rel err %.3e (bad rows %d)\n", worst, bad);
  std::printf("y[0]=%.15g (rel err %.3e)  y[511]=%.15g (rel err %.3e)\n", y[0], e0, y[511], e511);
  const bool ok = bad == 0 && e0 < 1e-12 && e511 < 1e-12;
  std::printf("%s\n", ok ? "PASS" : "FAIL");
  return ok ? 0 : 1;
}
