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

  const std::vector<T>& getLocalVector() const { return hostInit_; }
  void setLocalVector(const std::vector<T>& data) { hostInit_ = data; }

  T* getLocalDataPtr() const { return unifiedLocal_; }
  size_t getLocalSize() const { return localSize_; }
  size_t getTotalSize() const { return haloLeft_ + localSize_ + haloRight_; }

  T* getUnifiedDataPtr() const { return unifiedBase_; }
  T* getDeviceDataPtr() const { return devBase_; }
  T* getDeviceLocalPtr() const { return devLocal_; }
  T* getDeviceLeftPtr() const { return devLeft_; }
  T* getDeviceRightPtr() const { return devRight_; }
  T* getDeviceCOOResultsPtr() const { return devCooResult_; }
  T* getHostDataPtr() const { return hostBase_; }
  T* getHostLocalPtr() const { return hostLocal_; }
  T* getHostLeftPtr() const { return hostLeft_; }
  T* getHostRightPtr() const { return hostRight_; }
  T* getHostCOOResultsPtr() const { return hostCooResult_; }

  bool isUnifiedMemory() const { return unified_; }
  bool isUnifiedMemoryMalloc() const { return unifiedPinned_; }
  bool isUnifiedMemoryManagedMalloc() const { return unifiedManaged_; }
  bool isUnifiedNumaAllocOnNode() const { return unifiedNuma_; }

  size_t getLeftSize() const { return haloLeft_; }
  size_t getRightSize() const { return haloRight_; }

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

  std::vector<T> hostInit_;
  size_t haloLeft_ = 0;
  size_t haloRight_ = 0;
  size_t localSize_ = 0;

  bool unified_ = false;
  bool unifiedPinned_ = false;
  bool unifiedManaged_ = false;
  bool unifiedNuma_ = false;

  T* unifiedBase_ = nullptr;
  T* unifiedLocal_ = nullptr;

  T* devBase_ = nullptr;
  T* devLeft_ = nullptr;
  T* devLocal_ = nullptr;
  T* devRight_ = nullptr;
  T* devCooResult_ = nullptr;

  T* hostBase_ = nullptr;
  T* hostLeft_ = nullptr;
  T* hostLocal_ = nullptr;
  T* hostRight_ = nullptr;
  T* hostCooResult_ = nullptr;
};

#endif  // HBDIAVECTOR_HPP
