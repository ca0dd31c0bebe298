// Staging.cpp -- pinned double-buffered host <-> device copies (staging.hpp).
#include "staging.hpp"

#include <sys/mman.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "hbdia_hip.h"
#include "types.hpp"

namespace hbdia_host {
namespace {

constexpr size_t kChunk = size_t(64) << 20;
constexpr size_t kDirect = size_t(32) << 20;  // smaller copies go direct
constexpr size_t kPiece = size_t(1) << 20;    // per-thread piece of a host copy

// Two pinned buffers, a stream and an event per buffer, created on first use
// and kept for the life of the process (never freed: the HIP runtime may be
// gone by the time static destructors run).
struct Stager {
  std::mutex mu;
  void* buf[2] = {nullptr, nullptr};
  hbdia_stream_t stream = nullptr;
  hbdia_event_t done[2] = {nullptr, nullptr};

  void init() {
    if (stream) return;
    for (int b = 0; b < 2; ++b) {
      CHECK_HBDIA(hbdia_malloc_host(&buf[b], kChunk));
      CHECK_HBDIA(hbdia_event_create(&done[b]));
    }
    CHECK_HBDIA(hbdia_stream_create(&stream));
  }
};

Stager& stager() {
  static Stager* s = new Stager();
  return *s;
}

void parallel_copy(void* dst, const void* src, size_t bytes) {
  const int64_t pieces = static_cast<int64_t>((bytes + kPiece - 1) / kPiece);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < pieces; ++k) {
    const size_t off = static_cast<size_t>(k) * kPiece;
    std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, std::min(kPiece, bytes - off));
  }
}

}  // namespace

void upload_staged(void* dst, const void* src, size_t bytes) {
  if (bytes < kDirect) {
    if (bytes) CHECK_HBDIA(hbdia_memcpy(dst, src, bytes));
    return;
  }
  Stager& st = stager();
  std::lock_guard<std::mutex> lock(st.mu);
  st.init();
  const size_t chunks = (bytes + kChunk - 1) / kChunk;
  for (size_t k = 0; k < chunks; ++k) {
    const int b = static_cast<int>(k & 1);
    const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
    if (k >= 2) CHECK_HBDIA(hbdia_event_synchronize(st.done[b]));  // the copy of chunk k-2 left this buffer
    parallel_copy(st.buf[b], static_cast<const char*>(src) + off, len);
    CHECK_HBDIA(hbdia_memcpy_h2d(static_cast<char*>(dst) + off, st.buf[b], len, st.stream));
    CHECK_HBDIA(hbdia_event_record(st.done[b], st.stream));
  }
  CHECK_HBDIA(hbdia_stream_synchronize(st.stream));
}

void download_staged(void* dst, const void* src, size_t bytes) {
  if (bytes < kDirect) {
    if (bytes) CHECK_HBDIA(hbdia_memcpy(dst, src, bytes));
    return;
  }
  Stager& st = stager();
  std::lock_guard<std::mutex> lock(st.mu);
  st.init();
  const size_t chunks = (bytes + kChunk - 1) / kChunk;
  auto drain = [&](size_t k) {  // chunk k: staged -> dst
    const int b = static_cast<int>(k & 1);
    const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
    CHECK_HBDIA(hbdia_event_synchronize(st.done[b]));
    parallel_copy(static_cast<char*>(dst) + off, st.buf[b], len);
  };
  for (size_t k = 0; k < chunks; ++k) {
    // buffer k & 1 was drained (chunk k-2) in the previous iteration
    const int b = static_cast<int>(k & 1);
    const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
    CHECK_HBDIA(hbdia_memcpy_d2h(st.buf[b], static_cast<const char*>(src) + off, len, st.stream));
    CHECK_HBDIA(hbdia_event_record(st.done[b], st.stream));
    if (k >= 1) drain(k - 1);
  }
  drain(chunks - 1);
}

void advise_huge_pages(void* p, size_t bytes) {
  constexpr uintptr_t kHuge = uintptr_t(2) << 20;
  const uintptr_t a = (reinterpret_cast<uintptr_t>(p) + kHuge - 1) & ~(kHuge - 1);
  const uintptr_t e = (reinterpret_cast<uintptr_t>(p) + bytes) & ~(kHuge - 1);
  if (e > a) (void)madvise(reinterpret_cast<void*>(a), e - a, MADV_HUGEPAGE);
}

}  // namespace hbdia_host
