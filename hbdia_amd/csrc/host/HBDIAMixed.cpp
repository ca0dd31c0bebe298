// HBDIAMixed.cpp -- reduced-precision storage (see Operations/HBDIAMixed.hpp).
#include "Operations/HBDIAMixed.hpp"

#include <algorithm>
#include <stdexcept>
#include <vector>

namespace {
size_t elem_bytes(StoragePrecision p) { return p == StoragePrecision::FP32 ? 4 : 2; }
int vdtype(StoragePrecision p) { return p == StoragePrecision::FP32 ? HBDIA_VALUES_F32 : HBDIA_VALUES_BF16; }
}  // namespace

double HBDIAMixed::effectiveUnitRoundoff(StoragePrecision precision, int64_t terms) {
  const double u24 = 0x1p-24;
  const double store = precision == StoragePrecision::FP32 ? u24 : 0x1p-8 + u24;
  // the kernels evaluate the bound in FP32 and widen it by 1.0001 for that rounding
  return (store + static_cast<double>(terms + 2) * u24) * 1.0001;
}

int64_t HBDIAMixed::maxRowTerms(const HBDIA<double>& matrix) {
  int64_t blocked = 0;
  for (const auto& o : matrix.getOffsetsPerBlock()) blocked = std::max<int64_t>(blocked, static_cast<int64_t>(o.size()));
  // the fallback is row-sorted: its longest run of one row
  const std::vector<int>& rows = matrix.getCpuRowIndices();
  int64_t coo = 0;
  for (size_t i = 0, j; i < rows.size(); i = j) {
    for (j = i + 1; j < rows.size() && rows[j] == rows[i]; ++j) {
    }
    coo = std::max<int64_t>(coo, static_cast<int64_t>(j - i));
  }
  return blocked + coo;
}

double HBDIAMixed::ratioForTolerance(const HBDIA<double>& matrix, StoragePrecision precision, double relTol) {
  if (!(relTol > 0.0)) throw std::invalid_argument("HBDIAMixed: the relative tolerance must be positive");
  return relTol / effectiveUnitRoundoff(precision, maxRowTerms(matrix));
}

HBDIAMixed::HBDIAMixed(const HBDIA<double>& matrix, StoragePrecision precision)
    : HBDIAMixed(matrix, precision, ratioForTolerance(matrix, precision, kDefaultRelTolerance)) {}

HBDIAMixed::HBDIAMixed(const HBDIA<double>& matrix, StoragePrecision precision, double cancellationRatio)
    : m_(matrix), precision_(precision), ratio_(cancellationRatio), maxTerms_(maxRowTerms(matrix)) {
  if (!matrix.isHBDIAFormat() || !matrix.areStreamsInitialized())
    throw std::runtime_error("HBDIAMixed: matrix is not converted and on the device");
  if (matrix.isPartialMatrix())
    throw std::invalid_argument("HBDIAMixed: rank-local (partial) matrices run through hbdiaSpMV only");
  if (matrix.getBlockWidth() % HBDIA_ROW_GROUP != 0)
    throw std::invalid_argument("HBDIAMixed: block width must be a multiple of 64");
  if (!(cancellationRatio > 0.0)) throw std::invalid_argument("HBDIAMixed: cancellation ratio must be positive");
  const bool hasCOO = !matrix.getCpuValues().empty();
  if (hasCOO && !matrix.hasDeviceCOO())
    throw std::runtime_error("HBDIAMixed: the fallback is not on the device (convert with ExecutionMode::GPU_COO)");
  const size_t es = elem_bytes(precision);
  const size_t nv = matrix.getNumStoredValues(), nc = matrix.getCpuValues().size();
  hbdia_stream_t s = matrix.getBDIAStream();
  try {
    if (nv) {
      CHECK_HBDIA(hbdia_malloc(&values_, nv * es));
      CHECK_HBDIA(hbdia_convert_values(matrix.getHBDIADataDevice(), values_, static_cast<int64_t>(nv),
                                       vdtype(precision), s));
    }
    if (nc) {
      CHECK_HBDIA(hbdia_malloc(&cooValues_, nc * es));
      CHECK_HBDIA(hbdia_convert_values(matrix.getCpuValuesDevice(), cooValues_, static_cast<int64_t>(nc),
                                       vdtype(precision), s));
    }
    CHECK_HBDIA(hbdia_malloc(&workspace_, hbdia_mixed_workspace_bytes(matrix.getNumRows())));
    CHECK_HBDIA(hbdia_memset(workspace_, 0, hbdia_mixed_workspace_bytes(matrix.getNumRows()), s));
    CHECK_HBDIA(hbdia_stream_synchronize(s));
  } catch (...) {
    if (values_) (void)hbdia_free(values_);
    if (cooValues_) (void)hbdia_free(cooValues_);
    if (workspace_) (void)hbdia_free(workspace_);
    throw;
  }
  bytes_ = (nv + nc) * es;
}

HBDIAMixed::~HBDIAMixed() {
  (void)hbdia_stream_synchronize(m_.getBDIAStream());
  if (values_) (void)hbdia_free(values_);
  if (cooValues_) (void)hbdia_free(cooValues_);
  if (workspace_) (void)hbdia_free(workspace_);
}

void HBDIAMixed::spmv(const double* x, double* y) {
  const bool hasCOO = !m_.getCpuValues().empty();
  // With no stored values (every entry in the fallback) the kernels never
  // read the value arrays; any non-null device pointer satisfies the ABI.
  const void* vlp = values_ ? values_ : static_cast<const void*>(x);
  const double* v64 = m_.getHBDIADataDevice() ? m_.getHBDIADataDevice() : x;
  CHECK_HBDIA(hbdia_hbdia_spmv_mixed_ex(
      vdtype(precision_), vlp, v64,
      m_.getFlattenedOffsetsDevice(), m_.getBlockStartIndicesDevice(), m_.getNumBlocks(), m_.getBlockWidth(),
      hasCOO ? m_.getCpuRowIndicesDevice() : nullptr, hasCOO ? m_.getCpuColIndicesDevice() : nullptr,
      hasCOO ? cooValues_ : nullptr, hasCOO ? m_.getCpuValuesDevice() : nullptr,
      hasCOO ? m_.getCooGroupPtrDevice() : nullptr, x, y, m_.getNumRows(), ratio_, workspace_,
      m_.getRuns3BlocksDevice(), m_.getBDIAStream()));
  CHECK_HBDIA(hbdia_event_record(m_.getBDIAEvent(), m_.getBDIAStream()));
}

void HBDIAMixed::spmv(const HBDIAVector<double>& x, HBDIAVector<double>& y) {
  if (y.getLocalSize() < static_cast<size_t>(m_.getNumRows()) || x.getTotalSize() < static_cast<size_t>(m_.getNumRows()))
    throw std::invalid_argument("HBDIAMixed::spmv: vector shorter than the number of rows");
  if (x.getTotalSize() < static_cast<size_t>(m_.getNumCols()))
    throw std::invalid_argument("HBDIAMixed::spmv: input vector shorter than the number of columns");
  spmv(x.getDeviceDataPtr(), y.getDeviceLocalPtr());
}

std::vector<uint64_t> HBDIAMixed::lastFp64Masks() const {
  const int64_t groups = (static_cast<int64_t>(m_.getNumRows()) + HBDIA_ROW_GROUP - 1) / HBDIA_ROW_GROUP;
  std::vector<uint64_t> masks(static_cast<size_t>(groups));
  CHECK_HBDIA(hbdia_stream_synchronize(m_.getBDIAStream()));
  CHECK_HBDIA(hbdia_memcpy(masks.data(), workspace_, masks.size() * sizeof(uint64_t)));
  return masks;
}

int64_t HBDIAMixed::lastFp64Rows() const {
  int64_t n = 0;
  for (uint64_t w : lastFp64Masks()) n += __builtin_popcountll(w);
  return n;
}
