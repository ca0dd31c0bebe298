// Workloads.cpp -- seeded synthetic matrices and vectors (see Workloads.hpp).
#include "Workloads.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace hbdia {
namespace {

constexpr uint64_t kStreamRow = 1, kStreamCol = 2, kStreamVal = 3, kStreamVec = 4, kStreamOff = 5, kStreamBatch = 6;

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t hash4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
  uint64_t h = mix64(seed + stream * 0xD1B54A32D192ED03ull);
  h = mix64(h ^ a);
  return mix64(h ^ (b * 0x8CB92BA72F3D8DD7ull));
}

inline double u01(uint64_t h) { return static_cast<double>(h >> 11) * 0x1.0p-53; }

inline uint64_t bounded(uint64_t h, uint64_t n) {
  return static_cast<uint64_t>((static_cast<unsigned __int128>(h) * n) >> 64);
}

struct Grid {
  int64_t n;
  int points;
  // neighbour (dz, dy, dx) in ascending column order
  int nb = 0;
  int dz[27], dy[27], dx[27];
  explicit Grid(int64_t n_, int p) : n(n_), points(p) {
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b)
        for (int c = -1; c <= 1; ++c) {
          const int d = std::abs(a) + std::abs(b) + std::abs(c);
          if (points == 7 && d > 1) continue;
          dz[nb] = a;
          dy[nb] = b;
          dx[nb] = c;
          ++nb;
        }
  }
  bool inStencil(int64_t r, int64_t col) const {
    const int64_t z = r / (n * n), y = (r / n) % n, x = r % n;
    const int64_t cz = col / (n * n), cy = (col / n) % n, cx = col % n;
    const int64_t a = std::llabs(cz - z), b = std::llabs(cy - y), c = std::llabs(cx - x);
    if (a > 1 || b > 1 || c > 1) return false;
    return points == 27 || a + b + c <= 1;
  }
  int stencilCount(int64_t r) const {
    const int64_t z = r / (n * n), y = (r / n) % n, x = r % n;
    int cnt = 0;
    for (int k = 0; k < nb; ++k) {
      const int64_t zz = z + dz[k], yy = y + dy[k], xx = x + dx[k];
      if (zz >= 0 && zz < n && yy >= 0 && yy < n && xx >= 0 && xx < n) ++cnt;
    }
    return cnt;
  }
};

// Columns of the extra entries of row r (rejection sampling, at most
// 10 c + 100 draws as in the reference's noise generator).
template <typename Reject>
int drawExtraColumns(uint64_t seed, int64_t r, int c, uint64_t ncols, Reject reject, int* out) {
  int got = 0;
  const int maxAttempts = c * 10 + 100;
  for (int a = 0; a < maxAttempts && got < c; ++a) {
    const int64_t col = static_cast<int64_t>(bounded(hash4(seed, kStreamCol, static_cast<uint64_t>(r), a), ncols));
    if (reject(col)) continue;
    bool dup = false;
    for (int q = 0; q < got; ++q) dup |= (out[q] == col);
    if (dup) continue;
    out[got++] = static_cast<int>(col);
  }
  return got;
}

template <typename T>
void prefixAndAllocate(const std::vector<int>& perRow, std::vector<int64_t>& rowptr, CooMatrix<T>& out) {
  const int64_t N = static_cast<int64_t>(perRow.size());
  rowptr.assign(N + 1, 0);
  for (int64_t r = 0; r < N; ++r) rowptr[r + 1] = rowptr[r] + perRow[r];
  const int64_t nnz = rowptr[N];
  if (nnz > 0x7fffffffLL) throw std::runtime_error("generator: more than 2^31-1 non-zeros");
  out.rows.resize(nnz);
  out.cols.resize(nnz);
  out.vals.resize(nnz);
}

}  // namespace

template <typename T>
void generateStencilScatter(int n_, int points, double scatter, uint64_t seed, CooMatrix<T>& out) {
  if (n_ <= 0) throw std::invalid_argument("generateStencilScatter: n must be positive");
  if (points != 7 && points != 27) throw std::invalid_argument("generateStencilScatter: points must be 7 or 27");
  const Grid g(n_, points);
  const int64_t n = n_;
  const int64_t N = n * n * n;
  if (N > 0x7fffffffLL) throw std::invalid_argument("generateStencilScatter: grid too large for int32 indices");
  const int64_t extra = static_cast<int64_t>(scatter * static_cast<double>(points) * static_cast<double>(N));

  std::vector<int> extraPerRow(N, 0);
  {
    int* ep = extraPerRow.data();
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < extra; ++e) {
      const int64_t r = static_cast<int64_t>(bounded(hash4(seed, kStreamRow, static_cast<uint64_t>(e), 0), N));
#pragma omp atomic
      ep[r]++;
    }
  }
  // accepted extra entries per row (rejection may fall short on tiny grids)
  std::vector<int> perRow(N);
#pragma omp parallel
  {
    std::vector<int> cols;
#pragma omp for schedule(dynamic, 4096)
    for (int64_t r = 0; r < N; ++r) {
      int got = 0;
      if (extraPerRow[r] > 0) {
        cols.resize(extraPerRow[r]);
        got = drawExtraColumns(seed, r, extraPerRow[r], static_cast<uint64_t>(N),
                               [&](int64_t col) { return g.inStencil(r, col); }, cols.data());
      }
      extraPerRow[r] = got;
      perRow[r] = g.stencilCount(r) + got;
    }
  }
  std::vector<int64_t> rowptr;
  prefixAndAllocate(perRow, rowptr, out);
  out.numRows = static_cast<int>(N);
  out.numCols = static_cast<int>(N);
#pragma omp parallel
  {
    std::vector<int> cols;
    std::vector<double> v;
#pragma omp for schedule(dynamic, 4096)
    for (int64_t r = 0; r < N; ++r) {
      const int64_t z = r / (n * n), y = (r / n) % n, x = r % n;
      int64_t w = rowptr[r];
      int64_t diagPos = -1;
      double absSum = 0.0;
      for (int k = 0; k < g.nb; ++k) {
        const int64_t zz = z + g.dz[k], yy = y + g.dy[k], xx = x + g.dx[k];
        if (zz < 0 || zz >= n || yy < 0 || yy >= n || xx < 0 || xx >= n) continue;
        const int64_t col = (zz * n + yy) * n + xx;
        out.rows[w] = static_cast<int>(r);
        out.cols[w] = static_cast<int>(col);
        if (col == r) {
          diagPos = w;
        } else {
          const double val = u01(hash4(seed, kStreamVal, static_cast<uint64_t>(r), static_cast<uint64_t>(col))) - 1.0;
          out.vals[w] = static_cast<T>(val);
          absSum += std::fabs(val);
        }
        ++w;
      }
      const int c = extraPerRow[r];
      if (c > 0) {
        cols.resize(c);
        drawExtraColumns(seed, r, c, static_cast<uint64_t>(N), [&](int64_t col) { return g.inStencil(r, col); },
                         cols.data());
        for (int q = 0; q < c; ++q) {
          const double val =
              u01(hash4(seed, kStreamVal, static_cast<uint64_t>(r), static_cast<uint64_t>(cols[q]))) - 1.0;
          out.rows[w] = static_cast<int>(r);
          out.cols[w] = cols[q];
          out.vals[w] = static_cast<T>(val);
          absSum += std::fabs(val);
          ++w;
        }
      }
      out.vals[diagPos] = static_cast<T>(1.0 + absSum);
    }
  }
}

template <typename T>
void generateBanded(int n_, int noffsets, int maxOffset, double offband, uint64_t seed, CooMatrix<T>& out) {
  if (n_ <= 0) throw std::invalid_argument("generateBanded: n must be positive");
  if (noffsets < 1 || noffsets > 2 * maxOffset + 1)
    throw std::invalid_argument("generateBanded: noffsets must be in [1, 2*maxOffset+1]");
  const int64_t n = n_;
  // offsets: 0 plus distinct draws from [-maxOffset, maxOffset] \ {0}, ascending
  std::vector<int> offs{0};
  for (uint64_t a = 0; static_cast<int>(offs.size()) < noffsets; ++a) {
    const int o = static_cast<int>(bounded(hash4(seed, kStreamOff, a, 0), 2 * maxOffset + 1)) - maxOffset;
    if (std::find(offs.begin(), offs.end(), o) == offs.end()) offs.push_back(o);
  }
  std::sort(offs.begin(), offs.end());
  auto inBand = [&](int64_t r, int64_t col) {
    return std::binary_search(offs.begin(), offs.end(), static_cast<int>(col - r));
  };
  int64_t bandNnz = 0;
  for (int o : offs) bandNnz += std::max<int64_t>(0, n - (o < 0 ? -o : o));
  const int64_t extra = static_cast<int64_t>(offband * static_cast<double>(bandNnz));
  std::vector<int> extraPerRow(n, 0);
  for (int64_t e = 0; e < extra; ++e)
    extraPerRow[bounded(hash4(seed, kStreamRow, static_cast<uint64_t>(e), 0), n)]++;
  std::vector<int> perRow(n);
  std::vector<int> cols;
  for (int64_t r = 0; r < n; ++r) {
    int band = 0;
    for (int o : offs) band += (r + o >= 0 && r + o < n);
    int got = 0;
    if (extraPerRow[r] > 0) {
      cols.resize(extraPerRow[r]);
      got = drawExtraColumns(seed, r, extraPerRow[r], static_cast<uint64_t>(n),
                             [&](int64_t col) { return inBand(r, col); }, cols.data());
    }
    extraPerRow[r] = got;
    perRow[r] = band + got;
  }
  std::vector<int64_t> rowptr;
  prefixAndAllocate(perRow, rowptr, out);
  out.numRows = static_cast<int>(n);
  out.numCols = static_cast<int>(n);
  for (int64_t r = 0; r < n; ++r) {
    int64_t w = rowptr[r], diagPos = -1;
    double absSum = 0.0;
    for (int o : offs) {
      const int64_t col = r + o;
      if (col < 0 || col >= n) continue;
      out.rows[w] = static_cast<int>(r);
      out.cols[w] = static_cast<int>(col);
      if (o == 0) {
        diagPos = w;
      } else {
        const double val = 2.0 * u01(hash4(seed, kStreamVal, static_cast<uint64_t>(r), static_cast<uint64_t>(col))) - 1.0;
        out.vals[w] = static_cast<T>(val);
        absSum += std::fabs(val);
      }
      ++w;
    }
    const int c = extraPerRow[r];
    if (c > 0) {
      cols.resize(c);
      drawExtraColumns(seed, r, c, static_cast<uint64_t>(n), [&](int64_t col) { return inBand(r, col); },
                       cols.data());
      for (int q = 0; q < c; ++q) {
        const double val =
            2.0 * u01(hash4(seed, kStreamVal, static_cast<uint64_t>(r), static_cast<uint64_t>(cols[q]))) - 1.0;
        out.rows[w] = static_cast<int>(r);
        out.cols[w] = cols[q];
        out.vals[w] = static_cast<T>(val);
        absSum += std::fabs(val);
        ++w;
      }
    }
    out.vals[diagPos] = static_cast<T>(1.0 + absSum);
  }
}

BandedShape bandedBatchShape(const BandedBatchSpec& spec, int i) {
  if (spec.minRows < 1 || spec.maxRows < spec.minRows || spec.minOffsets < 1 || spec.maxOffsets < spec.minOffsets)
    throw std::invalid_argument("bandedBatchShape: invalid spec");
  const uint64_t h = hash4(spec.seed, kStreamBatch, static_cast<uint64_t>(i), 0);
  const double lo = std::log(static_cast<double>(spec.minRows)), hi = std::log(static_cast<double>(spec.maxRows));
  int n = static_cast<int>(std::floor(std::exp(lo + (hi - lo) * u01(h))));
  n = std::min(std::max(n, spec.minRows), spec.maxRows);
  const int span = spec.maxOffsets - spec.minOffsets + 1;
  const int k = spec.minOffsets + static_cast<int>(bounded(hash4(spec.seed, kStreamBatch, static_cast<uint64_t>(i), 1),
                                                           static_cast<uint64_t>(span)));
  return {n, k, hash4(spec.seed, kStreamBatch, static_cast<uint64_t>(i), 2)};
}

template <typename T>
void generateBandedBatch(const BandedBatchSpec& spec, std::vector<CooMatrix<T>>& out) {
  if (spec.count < 0) throw std::invalid_argument("generateBandedBatch: negative count");
  out.assign(static_cast<size_t>(spec.count), CooMatrix<T>{});
  std::vector<BandedShape> shapes(static_cast<size_t>(spec.count));
  for (int i = 0; i < spec.count; ++i) shapes[i] = bandedBatchShape(spec, i);
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < spec.count; ++i)
    generateBanded<T>(shapes[i].rows, shapes[i].noffsets, spec.maxOffset, spec.offband, shapes[i].seed, out[i]);
}

template <typename T>
void generateUniformVector(int64_t n, uint64_t seed, double lo, double hi, T* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    out[i] = static_cast<T>(lo + (hi - lo) * u01(hash4(seed, kStreamVec, static_cast<uint64_t>(i), 0)));
}

template void generateStencilScatter<double>(int, int, double, uint64_t, CooMatrix<double>&);
template void generateStencilScatter<float>(int, int, double, uint64_t, CooMatrix<float>&);
template void generateBanded<double>(int, int, int, double, uint64_t, CooMatrix<double>&);
template void generateBanded<float>(int, int, int, double, uint64_t, CooMatrix<float>&);
template void generateBandedBatch<double>(const BandedBatchSpec&, std::vector<CooMatrix<double>>&);
template void generateBandedBatch<float>(const BandedBatchSpec&, std::vector<CooMatrix<float>>&);
template void generateUniformVector<double>(int64_t, uint64_t, double, double, double*);
template void generateUniformVector<float>(int64_t, uint64_t, double, double, float*);

}  // namespace hbdia
