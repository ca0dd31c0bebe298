// hbdia_internal.hpp -- helpers shared by the HIP translation units of
// libhbdia_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace hbdia_detail {

// The thread's last error, returned by hbdia_last_error().
extern thread_local std::string g_last_error;
// Records msg as the last error and returns code.
int fail(int code, const std::string& msg);
// HBDIA_SUCCESS for hipSuccess, else records "what: <HIP error>" and returns
// HBDIA_ERROR_OUT_OF_MEMORY or HBDIA_ERROR_DEVICE.
int check_hip(hipError_t e, const char* what);

template <typename T> __device__ __forceinline__ T fma_t(T a, T b, T c);
template <> __device__ __forceinline__ double fma_t<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
template <> __device__ __forceinline__ float fma_t<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

// ok ? fma(a, b, c) : c, computed as fma(a', b', c) with a' = -0 and b' = +0
// when !ok: c + (-0) is c for every c, -0 and NaN included (a sum that
// underflowed to -0 keeps its sign across a skipped entry, as the reference's
// skip does).  Masking the operands instead of selecting the result keeps
// every load of a chunk unconditional: with a select the compiler sinks the x
// load behind the test of the loaded value, a second memory latency per chunk.
__device__ __forceinline__ double masked_fma(bool ok, double a, double b, double c) {
  const uint64_t m = ok ? ~0ull : 0ull;
  const uint64_t sign = 0x8000000000000000ull;
  const double am = __builtin_bit_cast(double, (__builtin_bit_cast(uint64_t, a) & m) | (~m & sign));
  const double bm = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, b) & m);
  return __builtin_fma(am, bm, c);
}
__device__ __forceinline__ float masked_fma(bool ok, float a, float b, float c) {
  const uint32_t m = ok ? ~0u : 0u;
  const float am = __builtin_bit_cast(float, (__builtin_bit_cast(uint32_t, a) & m) | (~m & 0x80000000u));
  const float bm = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, b) & m);
  return __builtin_fmaf(am, bm, c);
}

// Workgroup b of a launch runs on XCD b % 8.  xcd_chunk_work maps it to the
// work item so that each XCD takes runs of 2^kshift consecutive workgroups:
// XCD k, in its j-th turn, gets workgroups [(8j + k) 2^kshift, (8j + k + 1) 2^kshift).
// Rows whose x lines the same groups read -- a stencil's neighbouring planes --
// then share one XCD's L2 instead of being fetched by three.  The workgroups
// past the last whole round of 8 * 2^kshift keep the identity, so every XCD
// gets the same share.  kshift = 0 is the identity.
__device__ __forceinline__ unsigned xcd_chunk_work(unsigned b, unsigned nwg, int kshift) {
  const unsigned full = (nwg >> (3 + kshift)) << (3 + kshift);
  if (b >= full) return b;
  const unsigned xcd = b & 7u, t = b >> 3;
  return ((((t >> kshift) << 3) | xcd) << kshift) | (t & ((1u << kshift) - 1u));
}

// log2 of the XCD run length for a launch of nwg workgroups: the largest power
// of two up to 1024 that still leaves every XCD at least 8 runs.  At 256^3 a
// z-plane is 256 workgroups; runs of 256 to 1024 measured 3 % faster than row
// order (648-652 vs 670-674 us), runs of 8192 (an eighth of the rows per XCD)
// 7 % slower.
inline int xcd_chunk_shift(int64_t nwg) {
  int k = 0;
  while (k < 10 && (int64_t(1) << (k + 1)) * 64 <= nwg) ++k;
  return k;
}

}  // namespace hbdia_detail
