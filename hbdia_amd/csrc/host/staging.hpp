// staging.hpp -- large host <-> device copies and large host buffers.
//
// Copies of pageable host memory go through two pinned staging buffers of
// 64 MiB: the host side fills or drains one buffer with all OpenMP threads
// while the DMA engine moves the other, so a copy runs at the slower of the
// two instead of the driver's pageable rate.  Small copies go direct.
#pragma once

#include <cstddef>
#include <vector>

namespace hbdia_host {

// Synchronous copies between pageable host memory and device memory.
void upload_staged(void* dst_device, const void* src_host, size_t bytes);
void download_staged(void* dst_host, const void* src_device, size_t bytes);

// Asks the kernel for transparent huge pages on the 2 MiB-aligned part of
// [p, p + bytes) (no effect where they are off); fewer, larger page faults
// when the range is first written.
void advise_huge_pages(void* p, size_t bytes);

// v = n zeros, in fresh storage advised for huge pages when it is large.
template <typename U>
void assign_zeros(std::vector<U>& v, size_t n) {
  if (n * sizeof(U) >= (size_t(64) << 20)) {
    std::vector<U>().swap(v);
    v.reserve(n);
    advise_huge_pages(v.data(), n * sizeof(U));
  }
  v.assign(n, U(0));
}

}  // namespace hbdia_host
