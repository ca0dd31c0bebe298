// HBDIAVector.cpp -- device-resident vectors (reference src/Format/HBDIAVector.cpp).
#include "Format/HBDIAVector.hpp"

#include <algorithm>
#include <iostream>
#include <stdexcept>
#include <utility>

template <typename T>
HBDIAVector<T>::HBDIAVector() = default;

template <typename T>
HBDIAVector<T>::HBDIAVector(const std::vector<T>& localData) : localVector_(localData) {
  size_local_ = localData.size();
  setupUnifiedMemory();
}

template <typename T>
HBDIAVector<T>::HBDIAVector(const std::vector<T>& localData, bool unifiedMemory, bool unifiedMemoryMalloc,
                            bool unifiedMemoryManagedMalloc, bool unifiedNumaAllocOnNode)
    : localVector_(localData) {
  unifiedMemory_ = unifiedMemory;
  unifiedMemoryMalloc_ = unifiedMemoryMalloc;
  unifiedMemoryManagedMalloc_ = unifiedMemoryManagedMalloc;
  unifiedNumaAllocOnNode_ = unifiedNumaAllocOnNode;
  size_local_ = localData.size();
  setupUnifiedMemory();
}

template <typename T>
HBDIAVector<T>::HBDIAVector(const std::vector<T>& localData, const HBDIA<T>& matrix, int rank, int size)
    : localVector_(localData) {
  size_local_ = localData.size();
  if (matrix.isPartialMatrix() && matrix.hasPartialMatrixMetadata()) {
    const auto& md = matrix.getPartialMatrixMetadata();
    for (int p = 0; p < size && p < static_cast<int>(md.processDataRanges.size()); ++p) {
      if (p == rank) continue;
      for (const auto& range : md.processDataRanges[p]) {
        const size_t n = static_cast<size_t>(std::get<1>(range) - std::get<0>(range));
        if (p < rank) size_recv_left_ += n;
        else size_recv_right_ += n;
      }
    }
  }
  setupUnifiedMemory();
}

template <typename T>
HBDIAVector<T>::~HBDIAVector() {
  cleanupUnifiedMemory();
}

template <typename T>
void HBDIAVector<T>::moveFrom(HBDIAVector& o) noexcept {
  localVector_ = std::move(o.localVector_);
  size_recv_left_ = std::exchange(o.size_recv_left_, 0);
  size_recv_right_ = std::exchange(o.size_recv_right_, 0);
  size_local_ = std::exchange(o.size_local_, 0);
  unifiedMemory_ = o.unifiedMemory_;
  unifiedMemoryMalloc_ = o.unifiedMemoryMalloc_;
  unifiedMemoryManagedMalloc_ = o.unifiedMemoryManagedMalloc_;
  unifiedNumaAllocOnNode_ = o.unifiedNumaAllocOnNode_;
  unified_data_ptr_ = std::exchange(o.unified_data_ptr_, nullptr);
  unified_local_ptr_ = std::exchange(o.unified_local_ptr_, nullptr);
  data_ptr_d_ = std::exchange(o.data_ptr_d_, nullptr);
  left_ptr_d_ = std::exchange(o.left_ptr_d_, nullptr);
  local_ptr_d_ = std::exchange(o.local_ptr_d_, nullptr);
  right_ptr_d_ = std::exchange(o.right_ptr_d_, nullptr);
  coo_result_ptr_d_ = std::exchange(o.coo_result_ptr_d_, nullptr);
  data_ptr_h_ = std::exchange(o.data_ptr_h_, nullptr);
  left_ptr_h_ = std::exchange(o.left_ptr_h_, nullptr);
  local_ptr_h_ = std::exchange(o.local_ptr_h_, nullptr);
  right_ptr_h_ = std::exchange(o.right_ptr_h_, nullptr);
  coo_result_ptr_h_ = std::exchange(o.coo_result_ptr_h_, nullptr);
}

template <typename T>
HBDIAVector<T>::HBDIAVector(HBDIAVector&& o) noexcept {
  moveFrom(o);
}

template <typename T>
HBDIAVector<T>& HBDIAVector<T>::operator=(HBDIAVector&& o) noexcept {
  if (this != &o) {
    cleanupUnifiedMemory();
    moveFrom(o);
  }
  return *this;
}

// Reference src/Format/HBDIAVector.cpp:77-154.  All buffers start zeroed; the
// local segment receives localData.
template <typename T>
void HBDIAVector<T>::setupUnifiedMemory() {
  const size_t total = size_recv_left_ + size_local_ + size_recv_right_;
  const size_t bytes = total * sizeof(T), cooBytes = size_local_ * sizeof(T);
  void* p = nullptr;
  void* q = nullptr;
  if (unifiedMemory_) {
    if (unifiedMemoryManagedMalloc_) {
      CHECK_HBDIA(hbdia_malloc_managed(&p, bytes));
      CHECK_HBDIA(hbdia_malloc_managed(&q, cooBytes));
    } else if (unifiedMemoryMalloc_ || unifiedNumaAllocOnNode_) {
      CHECK_HBDIA(hbdia_malloc_host(&p, bytes));
      CHECK_HBDIA(hbdia_malloc_host(&q, cooBytes));
    } else {
      throw std::runtime_error("Failed to setup unified memory for HBDIAVector");
    }
  } else {
    CHECK_HBDIA(hbdia_malloc(&p, bytes));
    CHECK_HBDIA(hbdia_malloc(&q, cooBytes));
  }
  data_ptr_d_ = static_cast<T*>(p);
  coo_result_ptr_d_ = static_cast<T*>(q);
  left_ptr_d_ = data_ptr_d_;
  local_ptr_d_ = data_ptr_d_ ? data_ptr_d_ + size_recv_left_ : nullptr;
  right_ptr_d_ = data_ptr_d_ ? data_ptr_d_ + size_recv_left_ + size_local_ : nullptr;
  CHECK_HBDIA(hbdia_memset(data_ptr_d_, 0, bytes, nullptr));
  CHECK_HBDIA(hbdia_memset(coo_result_ptr_d_, 0, cooBytes, nullptr));
  CHECK_HBDIA(hbdia_device_synchronize());
  CHECK_HBDIA(hbdia_memcpy(local_ptr_d_, localVector_.data(), cooBytes));
  if (unifiedMemory_) {
    unified_data_ptr_ = data_ptr_d_;
    unified_local_ptr_ = local_ptr_d_;
  } else {
    void* h = nullptr;
    void* hc = nullptr;
    CHECK_HBDIA(hbdia_malloc_host(&h, bytes));
    CHECK_HBDIA(hbdia_malloc_host(&hc, cooBytes));
    data_ptr_h_ = static_cast<T*>(h);
    coo_result_ptr_h_ = static_cast<T*>(hc);
    if (data_ptr_h_) std::fill(data_ptr_h_, data_ptr_h_ + total, T{});
    if (coo_result_ptr_h_) std::fill(coo_result_ptr_h_, coo_result_ptr_h_ + size_local_, T{});
    left_ptr_h_ = data_ptr_h_;
    local_ptr_h_ = data_ptr_h_ ? data_ptr_h_ + size_recv_left_ : nullptr;
    right_ptr_h_ = data_ptr_h_ ? data_ptr_h_ + size_recv_left_ + size_local_ : nullptr;
  }
}

template <typename T>
void HBDIAVector<T>::cleanupUnifiedMemory() noexcept {
  if (unifiedMemory_ && (unifiedMemoryMalloc_ || unifiedNumaAllocOnNode_) && !unifiedMemoryManagedMalloc_) {
    if (data_ptr_d_) (void)hbdia_free_host(data_ptr_d_);
    if (coo_result_ptr_d_) (void)hbdia_free_host(coo_result_ptr_d_);
  } else {
    if (data_ptr_d_) (void)hbdia_free(data_ptr_d_);
    if (coo_result_ptr_d_) (void)hbdia_free(coo_result_ptr_d_);
  }
  if (data_ptr_h_) (void)hbdia_free_host(data_ptr_h_);
  if (coo_result_ptr_h_) (void)hbdia_free_host(coo_result_ptr_h_);
  data_ptr_d_ = left_ptr_d_ = local_ptr_d_ = right_ptr_d_ = coo_result_ptr_d_ = nullptr;
  data_ptr_h_ = left_ptr_h_ = local_ptr_h_ = right_ptr_h_ = coo_result_ptr_h_ = nullptr;
  unified_data_ptr_ = unified_local_ptr_ = nullptr;
}

template <typename T>
void HBDIAVector<T>::upload(const std::vector<T>& data) {
  if (data.size() != size_local_)
    throw std::invalid_argument("HBDIAVector::upload: size " + std::to_string(data.size()) +
                                " != local size " + std::to_string(size_local_));
  localVector_ = data;
  CHECK_HBDIA(hbdia_memcpy(local_ptr_d_, data.data(), size_local_ * sizeof(T)));
}

template <typename T>
void HBDIAVector<T>::uploadHalo(const std::vector<T>& left, const std::vector<T>& right) {
  if (left.size() != size_recv_left_ || right.size() != size_recv_right_)
    throw std::invalid_argument("HBDIAVector::uploadHalo: halo sizes " + std::to_string(left.size()) + "+" +
                                std::to_string(right.size()) + " != " + std::to_string(size_recv_left_) + "+" +
                                std::to_string(size_recv_right_));
  if (!left.empty()) {
    if (left_ptr_h_) std::copy(left.begin(), left.end(), left_ptr_h_);
    CHECK_HBDIA(hbdia_memcpy(left_ptr_d_, left.data(), left.size() * sizeof(T)));
  }
  if (!right.empty()) {
    if (right_ptr_h_) std::copy(right.begin(), right.end(), right_ptr_h_);
    CHECK_HBDIA(hbdia_memcpy(right_ptr_d_, right.data(), right.size() * sizeof(T)));
  }
}

template <typename T>
void HBDIAVector<T>::download(std::vector<T>& out) const {
  out.resize(size_local_);
  CHECK_HBDIA(hbdia_memcpy(out.data(), local_ptr_d_, size_local_ * sizeof(T)));
}

template <typename T>
std::vector<T> HBDIAVector<T>::downloadAll() const {
  std::vector<T> out(getTotalSize());
  if (!out.empty()) CHECK_HBDIA(hbdia_memcpy(out.data(), data_ptr_d_, out.size() * sizeof(T)));
  return out;
}

template <typename T>
std::vector<T> HBDIAVector<T>::download() const {
  std::vector<T> out;
  download(out);
  return out;
}

template <typename T>
void HBDIAVector<T>::print(const std::string& vectorName) const {
  std::vector<T> v = download();
  std::cout << vectorName << " (" << size_local_ << " local, halo " << size_recv_left_ << "+"
            << size_recv_right_ << "):";
  const size_t n = std::min<size_t>(v.size(), 32);
  for (size_t i = 0; i < n; ++i) std::cout << " " << v[i];
  if (n < v.size()) std::cout << " ...";
  std::cout << std::endl;
}

template class HBDIAVector<double>;
template class HBDIAVector<float>;
template class HBDIAVector<int>;
