// Workloads.hpp -- seeded synthetic inputs for the HBDIA benchmarks.
//
// All randomness comes from a counter-based hash (SplitMix64 finaliser over
// (seed, stream, a, b)), so every entry is a pure function of its coordinates:
// the generators run in parallel and give the same matrix for any thread
// count.
//
// stencil + scatter (the 256^3 / 192^3 / 128^3 benchmark matrices):
//   grid n^3, row r = z n^2 + y n + x; points = 27 (3x3x3) or 7 (faces);
//   off-diagonal stencil values U[-1,0); then floor(scatter * points * N)
//   extra entries whose rows are uniform over [0,N) and whose columns are
//   uniform over [0,N), distinct within the row and from the row's stencil
//   columns, values U[-1,0); the diagonal is 1 + sum |off-diagonal| of the
//   row.  Entries are row-sorted: stencil columns ascending, then the extra
//   entries in draw order.
// banded (many independent matrices):
//   n rows, `noffsets` distinct diagonal offsets drawn from [-maxOffset,
//   maxOffset] always including 0, values U[-1,1) off the main diagonal,
//   diagonal = 1 + sum |off| of the row; plus floor(offband * band entries)
//   entries at uniform rows and uniform columns off the band (values U[-1,1),
//   counted in the diagonal's sum too).
#ifndef HBDIA_WORKLOADS_HPP
#define HBDIA_WORKLOADS_HPP

#include <cstdint>
#include <vector>

namespace hbdia {

template <typename T>
struct CooMatrix {
  std::vector<int> rows, cols;
  std::vector<T> vals;
  int numRows = 0, numCols = 0;
};

template <typename T>
void generateStencilScatter(int n, int points, double scatter, uint64_t seed, CooMatrix<T>& out);

template <typename T>
void generateBanded(int n, int noffsets, int maxOffset, double offband, uint64_t seed, CooMatrix<T>& out);

// A set of independent banded matrices (the many-matrix batch): matrix i has
// n_i = floor(exp(ln minRows + U (ln maxRows - ln minRows))) rows (log-uniform),
// noffsets_i uniform in [minOffsets, maxOffsets], and its own seed, all drawn
// from (seed, i); each is generateBanded(n_i, noffsets_i, maxOffset, offband,
// seed_i).
struct BandedBatchSpec {
  int count = 2048;
  int minRows = 1000;
  int maxRows = 100000;
  int minOffsets = 3;
  int maxOffsets = 9;
  int maxOffset = 64;
  double offband = 0.01;
  uint64_t seed = 0xC4B;
};

struct BandedShape {
  int rows;
  int noffsets;
  uint64_t seed;
};

BandedShape bandedBatchShape(const BandedBatchSpec& spec, int i);

// Generates every matrix of the set (in parallel over matrices).
template <typename T>
void generateBandedBatch(const BandedBatchSpec& spec, std::vector<CooMatrix<T>>& out);

// v[i] = lo + (hi - lo) * U[0,1)
template <typename T>
void generateUniformVector(int64_t n, uint64_t seed, double lo, double hi, T* out);

}  // namespace hbdia

#endif  // HBDIA_WORKLOADS_HPP
