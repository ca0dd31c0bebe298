// HBDIAVector.cpp -- device-resident vectors (reference src/Format/HBDIAVector.cpp).
#include "Format/HBDIAVector.hpp"

#include <algorithm>
#include <iostream>
#include <stdexcept>
#include <utility>

template <typename T>
HBDIAVector<T>::HBDIAVector() = default;

template <typename T>
HBDIAVector<T>::HBDIAVector(const std::vector<T>& localData) : hostInit_(localData) {
  localSize_ = localData.size();
  setupUnifiedMemory();
}

template <typename T>
HBDIAVector<T>::HBDIAVector(const std::vector<T>& localData, bool unifiedMemory, bool unifiedMemoryMalloc,
                            bool unifiedMemoryManagedMalloc, bool unifiedNumaAllocOnNode)
    : hostInit_(localData) {
  unified_ = unifiedMemory;
  unifiedPinned_ = unifiedMemoryMalloc;
  unifiedManaged_ = unifiedMemoryManagedMalloc;
  unifiedNuma_ = unifiedNumaAllocOnNode;
  localSize_ = localData.size();
  setupUnifiedMemory();
}

template <typename T>
HBDIAVector<T>::HBDIAVector(const std::vector<T>& localData, const HBDIA<T>& matrix, int rank, int size)
    : hostInit_(localData) {
  localSize_ = localData.size();
  if (matrix.isPartialMatrix() && matrix.hasPartialMatrixMetadata()) {
    const auto& md = matrix.getPartialMatrixMetadata();
    for (int p = 0; p < size && p < static_cast<int>(md.processDataRanges.size()); ++p) {
      if (p == rank) continue;
      for (const auto& range : md.processDataRanges[p]) {
        const size_t n = static_cast<size_t>(std::get<1>(range) - std::get<0>(range));
        if (p < rank) haloLeft_ += n;
        else haloRight_ += n;
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
  hostInit_ = std::move(o.hostInit_);
  haloLeft_ = std::exchange(o.haloLeft_, 0);
  haloRight_ = std::exchange(o.haloRight_, 0);
  localSize_ = std::exchange(o.localSize_, 0);
  unified_ = o.unified_;
  unifiedPinned_ = o.unifiedPinned_;
  unifiedManaged_ = o.unifiedManaged_;
  unifiedNuma_ = o.unifiedNuma_;
  unifiedBase_ = std::exchange(o.unifiedBase_, nullptr);
  unifiedLocal_ = std::exchange(o.unifiedLocal_, nullptr);
  devBase_ = std::exchange(o.devBase_, nullptr);
  devLeft_ = std::exchange(o.devLeft_, nullptr);
  devLocal_ = std::exchange(o.devLocal_, nullptr);
  devRight_ = std::exchange(o.devRight_, nullptr);
  devCooResult_ = std::exchange(o.devCooResult_, nullptr);
  hostBase_ = std::exchange(o.hostBase_, nullptr);
  hostLeft_ = std::exchange(o.hostLeft_, nullptr);
  hostLocal_ = std::exchange(o.hostLocal_, nullptr);
  hostRight_ = std::exchange(o.hostRight_, nullptr);
  hostCooResult_ = std::exchange(o.hostCooResult_, nullptr);
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
  const size_t total = haloLeft_ + localSize_ + haloRight_;
  const size_t bytes = total * sizeof(T), cooBytes = localSize_ * sizeof(T);
  void* p = nullptr;
  void* q = nullptr;
  if (unified_) {
    if (unifiedManaged_) {
      CHECK_HBDIA(hbdia_malloc_managed(&p, bytes));
      CHECK_HBDIA(hbdia_malloc_managed(&q, cooBytes));
    } else if (unifiedPinned_ || unifiedNuma_) {
      CHECK_HBDIA(hbdia_malloc_host(&p, bytes));
      CHECK_HBDIA(hbdia_malloc_host(&q, cooBytes));
    } else {
      throw std::runtime_error("Failed to setup unified memory for HBDIAVector");
    }
  } else {
    CHECK_HBDIA(hbdia_malloc(&p, bytes));
    CHECK_HBDIA(hbdia_malloc(&q, cooBytes));
  }
  devBase_ = static_cast<T*>(p);
  devCooResult_ = static_cast<T*>(q);
  devLeft_ = devBase_;
  devLocal_ = devBase_ ? devBase_ + haloLeft_ : nullptr;
  devRight_ = devBase_ ? devBase_ + haloLeft_ + localSize_ : nullptr;
  CHECK_HBDIA(hbdia_memset(devBase_, 0, bytes, nullptr));
  CHECK_HBDIA(hbdia_memset(devCooResult_, 0, cooBytes, nullptr));
  CHECK_HBDIA(hbdia_device_synchronize());
  CHECK_HBDIA(hbdia_memcpy(devLocal_, hostInit_.data(), cooBytes));
  if (unified_) {
    unifiedBase_ = devBase_;
    unifiedLocal_ = devLocal_;
  } else {
    void* h = nullptr;
    void* hc = nullptr;
    CHECK_HBDIA(hbdia_malloc_host(&h, bytes));
    CHECK_HBDIA(hbdia_malloc_host(&hc, cooBytes));
    hostBase_ = static_cast<T*>(h);
    hostCooResult_ = static_cast<T*>(hc);
    if (hostBase_) std::fill(hostBase_, hostBase_ + total, T{});
    if (hostCooResult_) std::fill(hostCooResult_, hostCooResult_ + localSize_, T{});
    hostLeft_ = hostBase_;
    hostLocal_ = hostBase_ ? hostBase_ + haloLeft_ : nullptr;
    hostRight_ = hostBase_ ? hostBase_ + haloLeft_ + localSize_ : nullptr;
  }
}

template <typename T>
void HBDIAVector<T>::cleanupUnifiedMemory() noexcept {
  if (unified_ && (unifiedPinned_ || unifiedNuma_) && !unifiedManaged_) {
    if (devBase_) (void)hbdia_free_host(devBase_);
    if (devCooResult_) (void)hbdia_free_host(devCooResult_);
  } else {
    if (devBase_) (void)hbdia_free(devBase_);
    if (devCooResult_) (void)hbdia_free(devCooResult_);
  }
  if (hostBase_) (void)hbdia_free_host(hostBase_);
  if (hostCooResult_) (void)hbdia_free_host(hostCooResult_);
  devBase_ = devLeft_ = devLocal_ = devRight_ = devCooResult_ = nullptr;
  hostBase_ = hostLeft_ = hostLocal_ = hostRight_ = hostCooResult_ = nullptr;
  unifiedBase_ = unifiedLocal_ = nullptr;
}

template <typename T>
void HBDIAVector<T>::upload(const std::vector<T>& data) {
  if (data.size() != localSize_)
    throw std::invalid_argument("HBDIAVector::upload: size " + std::to_string(data.size()) +
                                " != local size " + std::to_string(localSize_));
  hostInit_ = data;
  CHECK_HBDIA(hbdia_memcpy(devLocal_, data.data(), localSize_ * sizeof(T)));
}

template <typename T>
void HBDIAVector<T>::uploadHalo(const std::vector<T>& left, const std::vector<T>& right) {
  if (left.size() != haloLeft_ || right.size() != haloRight_)
    throw std::invalid_argument("HBDIAVector::uploadHalo: halo sizes " + std::to_string(left.size()) + "+" +
                                std::to_string(right.size()) + " != " + std::to_string(haloLeft_) + "+" +
                                std::to_string(haloRight_));
  if (!left.empty()) {
    if (hostLeft_) std::copy(left.begin(), left.end(), hostLeft_);
    CHECK_HBDIA(hbdia_memcpy(devLeft_, left.data(), left.size() * sizeof(T)));
  }
  if (!right.empty()) {
    if (hostRight_) std::copy(right.begin(), right.end(), hostRight_);
    CHECK_HBDIA(hbdia_memcpy(devRight_, right.data(), right.size() * sizeof(T)));
  }
}

template <typename T>
void HBDIAVector<T>::download(std::vector<T>& out) const {
  out.resize(localSize_);
  CHECK_HBDIA(hbdia_memcpy(out.data(), devLocal_, localSize_ * sizeof(T)));
}

template <typename T>
std::vector<T> HBDIAVector<T>::downloadAll() const {
  std::vector<T> out(getTotalSize());
  if (!out.empty()) CHECK_HBDIA(hbdia_memcpy(out.data(), devBase_, out.size() * sizeof(T)));
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
  std::cout << vectorName << " (" << localSize_ << " local, halo " << haloLeft_ << "+"
            << haloRight_ << "):";
  const size_t n = std::min<size_t>(v.size(), 32);
  for (size_t i = 0; i < n; ++i) std::cout << " " << v[i];
  if (n < v.size()) std::cout << " ...";
  std::cout << std::endl;
}

template class HBDIAVector<double>;
template class HBDIAVector<float>;
template class HBDIAVector<int>;
