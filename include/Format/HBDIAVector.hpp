// HBDIAVector.hpp -- device-resident vector for the HBDIA SpMV.
//
// Drop-in for the reference's HBDIAVector<T> (include/Format/HBDIAVector.hpp:23-109,
// src/Format/HBDIAVector.cpp:12-196).  Layout [left halo | local | right halo]
// (both halos are empty on one GPU) plus a COO-result buffer of the local
// size.  Default mode: device memory from hbdia_malloc plus a pinned host
// mirror that only the CPU-COO path uses.  The "unified" modes map to memory
// the GPU can address directly: unifiedMemoryManagedMalloc -> hipMallocManaged;
// unifiedMemoryMalloc and unifiedNumaAllocOnNode -> pinned host memory
// (plain malloc memory is not GPU-addressable without XNACK, and NUMA
// placement is left to the OS).
//
// Unlike the reference, the type is move-only (an implicit copy would free
// the buffers twice) and offers upload()/download(), since setLocalVector()
// alone never reaches the device (reference HBDIAVector.hpp:38).
#ifndef HBDIAVECTOR_HPP
#define HBDIAVECTOR_HPP

#include <string>
#include <vector>

#include "HBDIA.hpp"

template <typename T>
class HBDIAVector {
 public:
  HBDIAVector();
  explicit HBDIAVector(const std::vector<T>& localData);
  HBDIAVector(const std::vector<T>& localData, bool unifiedMemory, bool unifiedMemoryMalloc,
              bool unifiedMemoryManagedMalloc, bool unifiedNumaAllocOnNode);
  // Halo-sized constructor of the reference (HBDIAVector.cpp:38-70); on one GPU
  // the halos come from the matrix's partial metadata and are normally empty.
  HBDIAVector(const std::vector<T>& localData, const HBDIA<T>& matrix, int rank, int size);
  ~HBDIAVector();

  HBDIAVector(const HBDIAVector&) = delete;
  HBDIAVector& operator=(const HBDIAVector&) = delete;
  HBDIAVector(HBDIAVector&& o) noexcept;
  HBDIAVector& operator=(HBDIAVector&& o) noexcept;

  const std::vector<T>& getLocalVector() const { return localVector_; }
  void setLocalVector(const std::vector<T>& data) { localVector_ = data; }

  T* getLocalDataPtr() const { return unified_local_ptr_; }
  size_t getLocalSize() const { return size_local_; }
  size_t getTotalSize() const { return size_recv_left_ + size_local_ + size_recv_right_; }

  T* getUnifiedDataPtr() const { return unified_data_ptr_; }
  T* getDeviceDataPtr() const { return data_ptr_d_; }
  T* getDeviceLocalPtr() const { return local_ptr_d_; }
  T* getDeviceLeftPtr() const { return left_ptr_d_; }
  T* getDeviceRightPtr() const { return right_ptr_d_; }
  T* getDeviceCOOResultsPtr() const { return coo_result_ptr_d_; }
  T* getHostDataPtr() const { return data_ptr_h_; }
  T* getHostLocalPtr() const { return local_ptr_h_; }
  T* getHostLeftPtr() const { return left_ptr_h_; }
  T* getHostRightPtr() const { return right_ptr_h_; }
  T* getHostCOOResultsPtr() const { return coo_result_ptr_h_; }

  bool isUnifiedMemory() const { return unifiedMemory_; }
  bool isUnifiedMemoryMalloc() const { return unifiedMemoryMalloc_; }
  bool isUnifiedMemoryManagedMalloc() const { return unifiedMemoryManagedMalloc_; }
  bool isUnifiedNumaAllocOnNode() const { return unifiedNumaAllocOnNode_; }

  size_t getLeftSize() const { return size_recv_left_; }
  size_t getRightSize() const { return size_recv_right_; }

  // Synchronous copies of the local segment.
  void upload(const std::vector<T>& data);
  // Synchronous copy of both halo segments (host mirror and device); the
  // sizes must equal getLeftSize() and getRightSize().
  void uploadHalo(const std::vector<T>& left, const std::vector<T>& right);
  void download(std::vector<T>& out) const;
  std::vector<T> download() const;
  // The whole device buffer, [left halo | local | right halo].
  std::vector<T> downloadAll() const;

  void print(const std::string& vectorName = "HBDIAVector") const;

 private:
  void setupUnifiedMemory();
  void cleanupUnifiedMemory() noexcept;
  void moveFrom(HBDIAVector& o) noexcept;

  std::vector<T> localVector_;
  size_t size_recv_left_ = 0;
  size_t size_recv_right_ = 0;
  size_t size_local_ = 0;

  bool unifiedMemory_ = false;
  bool unifiedMemoryMalloc_ = false;
  bool unifiedMemoryManagedMalloc_ = false;
  bool unifiedNumaAllocOnNode_ = false;

  T* unified_data_ptr_ = nullptr;
  T* unified_local_ptr_ = nullptr;

  T* data_ptr_d_ = nullptr;
  T* left_ptr_d_ = nullptr;
  T* local_ptr_d_ = nullptr;
  T* right_ptr_d_ = nullptr;
  T* coo_result_ptr_d_ = nullptr;

  T* data_ptr_h_ = nullptr;
  T* left_ptr_h_ = nullptr;
  T* local_ptr_h_ = nullptr;
  T* right_ptr_h_ = nullptr;
  T* coo_result_ptr_h_ = nullptr;
};

#endif  // HBDIAVECTOR_HPP
