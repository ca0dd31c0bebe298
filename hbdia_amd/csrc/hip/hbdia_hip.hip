// hbdia_hip.hip -- device side of the HBDIA SpMV path for MI355X (gfx950),
// exported through the C ABI declared in include/hbdia_hip.h.
//
// Kernel design (HBM-bound; no MFMA, SpMV is ~0.2 flop/byte):
//   * one 64-lane wavefront per group of 64 consecutive rows; four waves per
//     256-thread workgroup;
//   * workgroups run in row order, so that all eight XCDs work on
//     neighbouring rows at once (measured faster on MI355X than giving each
//     XCD a contiguous eighth of the rows);
//   * when the block width is a multiple of 64 the block, its offset list and
//     its start are wave-uniform (scalar loads); values of one (block, offset)
//     pair are one contiguous 512-byte FP64 line per wave, loaded with the
//     nontemporal hint so that the value stream does not evict x;
//   * offsets are consumed in branch-free chunks of 9 so that 9 value loads
//     and 9 x loads are in flight per wave before the first FMA; value and x
//     addresses are 32-bit byte offsets from wave-uniform bases;
//   * the COO fallback is fused: the entries of the 64-row group are loaded
//     cooperatively (one per lane), x is gathered in parallel, and each lane
//     then sums its own row's entries sequentially in entry order out of LDS.
//     This keeps the reference CPU summation order, is deterministic, and
//     writes y exactly once with no separate vector-add pass.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>

#include "hbdia_hip.h"
#include "hbdia_internal.hpp"

namespace hbdia_detail {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return HBDIA_SUCCESS;
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory) return fail(HBDIA_ERROR_OUT_OF_MEMORY, m);
  return fail(HBDIA_ERROR_DEVICE, m);
}

}  // namespace hbdia_detail

namespace {

using hbdia_detail::check_hip;
using hbdia_detail::fail;
using hbdia_detail::fma_t;
using hbdia_detail::g_last_error;
using hbdia_detail::masked_fma;
using hbdia_detail::xcd_chunk_shift;
using hbdia_detail::xcd_chunk_work;

inline hipStream_t hs(hbdia_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;
constexpr int kWavesPerWG = 4;
constexpr int kWG = kWave * kWavesPerWG;
constexpr int kChunk = 9;
constexpr int kRow2Chunk = 9;


__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum of the COO entries of this wave's 64-row group for the calling lane's
// row, in entry order, as sequential FMAs (the reference CPU COO order).
// Entries [p0, p1) are row-sorted and all lie in rows [row0, row0 + 64).
template <typename T>
__device__ __forceinline__ T coo_group_sum_from(const int* __restrict__ cr, const int* __restrict__ cc,
                                                const T* __restrict__ cv, const T* __restrict__ x,
                                                int p0, int p1, int64_t row0, int lane, T* s_v,
                                                T* s_x, int* s_lo, int* s_hi, T cs, bool* any) {
  bool has = false;
  for (int base = p0; base < p1; base += kWave) {
    const int i = base + lane;
    const bool ok = i < p1;
    int r = -1;
    T vv = T(0), xx = T(0);
    if (ok) {
      r = cr[i] - static_cast<int>(row0);
      vv = cv[i];
      xx = x[cc[i]];
      if (r < 0 || r >= kWave) r = -1;  // unsorted input: drop rather than write outside LDS
    }
    const int rp = __shfl_up(r, 1);
    const int rn = __shfl_down(r, 1);
    s_lo[lane] = 0;
    s_hi[lane] = 0;
    s_v[lane] = vv;
    s_x[lane] = xx;
    wave_sync();
    if (ok && r >= 0 && (lane == 0 || rp != r)) s_lo[r] = lane;
    if (ok && r >= 0 && (lane == kWave - 1 || i + 1 >= p1 || rn != r)) s_hi[r] = lane + 1;
    wave_sync();
    const int a = s_lo[lane], e = s_hi[lane];
    for (int k = a; k < e; ++k) cs = fma_t<T>(s_v[k], s_x[k], cs);
    has |= (e > a);
    wave_sync();
  }
  *any = has;
  return cs;
}

// The same sums for a group with many entries (more than one 64-entry
// chunk): each lane finds its row's entries [lo, hi) by a lower-bound search
// over the group's row-sorted rows, then sums them in entry order, four loads
// of each kind in flight at a time.  Same FMA sequence as the chunked loop, so
// the same result bit for bit, without a wave barrier per chunk.  Entries of
// rows outside the group match no lane and are skipped.
template <typename T>
__device__ __forceinline__ T coo_group_sum_direct(const int* __restrict__ cr, const int* __restrict__ cc,
                                                  const T* __restrict__ cv, const T* __restrict__ x,
                                                  int p0, int p1, int64_t row0, int lane, bool* any) {
  const int target = static_cast<int>(row0) + lane;
  // lower bound of target in cr[p0, p1) (the first entry whose row is not below
  // target, or p1): from a guess that assumes equal entries per row, gallop to
  // a bracket [lo, bh] that holds it, then bisect.  Rows of equal length (the
  // usual dense group) find it in one or two probes.  Every probe lies in
  // [p0, p1).
  const int n = p1 - p0;
  const int g = p0 + static_cast<int>((static_cast<int64_t>(lane) * n) >> 6);
  int lo, bh;
  if (cr[g] >= target) {
    bh = g;  // cr[bh] >= target
    lo = p0;
    for (int step = 1;; step <<= 1) {
      const int probe = bh - step;
      if (probe < p0) break;
      if (cr[probe] < target) {
        lo = probe + 1;
        break;
      }
      bh = probe;
    }
  } else {
    lo = g + 1;  // cr[lo - 1] < target
    bh = p1;
    for (int step = 1;; step <<= 1) {
      const int probe = lo + step - 1;
      if (probe >= p1) break;
      if (cr[probe] >= target) {
        bh = probe;
        break;
      }
      lo = probe + 1;
    }
  }
  while (lo < bh) {
    const int mid = (lo + bh) >> 1;
    if (cr[mid] < target) lo = mid + 1;
    else bh = mid;
  }
  // the next lane's lower bound ends this lane's rows; the last lane's is p1
  const int nxt = __shfl_down(lo, 1);
  int hi = lane == kWave - 1 ? p1 : nxt;
  if (hi < lo) hi = lo;
  T cs = T(0);
  int i = lo;
  for (; i + 4 <= hi; i += 4) {
    const int c0 = cc[i], c1 = cc[i + 1], c2 = cc[i + 2], c3 = cc[i + 3];
    const T v0 = cv[i], v1 = cv[i + 1], v2 = cv[i + 2], v3 = cv[i + 3];
    const T x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
    cs = fma_t<T>(v0, x0, cs);
    cs = fma_t<T>(v1, x1, cs);
    cs = fma_t<T>(v2, x2, cs);
    cs = fma_t<T>(v3, x3, cs);
  }
  for (; i < hi; ++i) cs = fma_t<T>(cv[i], x[cc[i]], cs);
  *any = hi > lo;
  return cs;
}

template <typename T>
__device__ __forceinline__ T coo_group_sum(const int* __restrict__ cr, const int* __restrict__ cc,
                                           const T* __restrict__ cv, const T* __restrict__ x,
                                           int p0, int p1, int64_t row0, int lane, T* s_v,
                                           T* s_x, int* s_lo, int* s_hi, bool* any) {
  if (p1 - p0 > kWave) return coo_group_sum_direct<T>(cr, cc, cv, x, p0, p1, row0, lane, any);
  return coo_group_sum_from<T>(cr, cc, cv, x, p0, p1, row0, lane, s_v, s_x, s_lo, s_hi, T(0), any);
}

// The fallback sums of a 128-row wave's two groups when their entries
// [p0, p2) fit one 64-entry pass (group A rows [row0, row0 + 64) holds
// [p0, p1), group B the next 64 rows [p1, p2)): one load of rows, columns,
// values and x for both groups, then each lane sums its row of each group in
// entry order (FMA).  The sums equal coo_group_sum's per group bit for bit.
// s_lo and s_hi hold 128 entries.
template <typename T, bool ADJ = false>
__device__ __forceinline__ void coo_pair_sum(const int* __restrict__ cr, const int* __restrict__ cc,
                                             const T* __restrict__ cv, const T* __restrict__ x, int p0, int p1,
                                             int p2, int64_t row0, int lane, T* s_v, T* s_x, int* s_lo, int* s_hi,
                                             T& csA, T& csB) {
  const int i = p0 + lane;
  const bool ok = i < p2;
  int r = -1;
  T vv = T(0), xx = T(0);
  if (ok) {
    r = cr[i] - static_cast<int>(row0);
    vv = cv[i];
    xx = x[cc[i]];
    // an entry outside its own group's rows (unsorted input) is dropped, as the
    // per-group sums drop it, rather than written outside LDS
    if (r < 0 || r >= 2 * kWave || ((i < p1) != (r < kWave))) r = -1;
  }
  const int rp = __shfl_up(r, 1);
  const int rn = __shfl_down(r, 1);
  s_lo[lane] = 0;
  s_lo[kWave + lane] = 0;
  s_hi[lane] = 0;
  s_hi[kWave + lane] = 0;
  s_v[lane] = vv;
  s_x[lane] = xx;
  wave_sync();
  if (ok && r >= 0 && (lane == 0 || rp != r)) s_lo[r] = lane;
  if (ok && r >= 0 && (lane == kWave - 1 || i + 1 >= p2 || rn != r)) s_hi[r] = lane + 1;
  wave_sync();
  T a_sum = T(0), b_sum = T(0);
  // csA, csB: rows row0 + lane and row0 + 64 + lane, or with ADJ the two-row
  // layout's rows row0 + 2 lane and row0 + 2 lane + 1
  const int ia = ADJ ? 2 * lane : lane, ib = ADJ ? 2 * lane + 1 : kWave + lane;
  const int a0 = s_lo[ia], e0 = s_hi[ia];
  for (int k = a0; k < e0; ++k) a_sum = fma_t<T>(s_v[k], s_x[k], a_sum);
  const int a1 = s_lo[ib], e1 = s_hi[ib];
  for (int k = a1; k < e1; ++k) b_sum = fma_t<T>(s_v[k], s_x[k], b_sum);
  wave_sync();
  csA = a_sum;
  csB = b_sum;
}

// Element `i` of a wave-uniform base as base + 32-bit byte offset, so that the
// load uses the scalar-base addressing form (one VGPR per address).
template <typename T>
__device__ __forceinline__ const T* at_u32(const T* base, uint32_t i) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * static_cast<uint32_t>(sizeof(T)));
}

// Loads x[c] for c already clamped into [0, nrows).  X32: nrows * sizeof(T) < 2^32.
template <typename T, bool X32>
__device__ __forceinline__ T load_x(const T* __restrict__ x, int64_t c) {
  if constexpr (X32) return *at_u32(x, static_cast<uint32_t>(c));
  else return x[c];
}

// Blocked-DIA sum of one row whose 64-row wave shares one offset list (one
// block of a width that is a multiple of 64, or two width-32 blocks with equal
// offsets).  `vb` is a wave-uniform value address and vlane the lane's index
// from it: offset o of the lane is vb[o*bw + vlane].
// Offsets are consumed in chunks of kChunk with no control flow inside a
// chunk: the tail of the last chunk re-reads the last valid line (a cache hit)
// and x addresses are clamped into [0, nrows), so every load of a chunk can
// be in flight at once.  A chunk's offsets are read as kChunk consecutive
// entries (one wide scalar load; entries past the block's count belong to the
// next block or to the padding behind the array and only feed clamped, unused
// x loads).  The FMA chain runs over ascending offsets and skips
// zero values and columns outside [0, nrows), exactly as the reference kernel.
template <typename T, bool X32>
__device__ __forceinline__ T bdia_sum_uniform(const T* __restrict__ vb, const int* __restrict__ op, int cnt,
                                              int bw, int vlane, const T* __restrict__ x, int64_t row,
                                              int64_t nrows, bool active) {
  T sum = T(0);
  const int64_t last = nrows - 1;
  for (int o0 = 0; o0 < cnt; o0 += kChunk) {
    int off[kChunk];
    T v[kChunk], xl[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) off[k] = op[o0 + k];
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      v[k] = __builtin_nontemporal_load(at_u32(vb, static_cast<uint32_t>(min(o0 + k, cnt - 1) * bw + vlane)));
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      int64_t c = row + off[k];
      c = c < 0 ? 0 : (c > last ? last : c);
      xl[k] = load_x<T, X32>(x, c);
    }
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const int64_t c = row + off[k];
      const bool ok = (o0 + k < cnt) & active & (v[k] != T(0)) & (c >= 0) & (c < nrows);
      sum = masked_fma(ok, v[k], xl[k], sum);
    }
  }
  return sum;
}

// Whole-wave lane shifts (DPP wave_shr:1 / wave_shl:1, 32 bits at a time):
// shr1 gives lane i the value of lane i-1 and lane 0 `first`; shl1 gives
// lane i the value of lane i+1 and lane 63 `last`.
__device__ __forceinline__ int dpp_shr1(int v, int first) {
  return __builtin_amdgcn_update_dpp(first, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_shl1(int v, int last) {
  return __builtin_amdgcn_update_dpp(last, v, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ float wave_shr1(float v, float first) {
  return __builtin_bit_cast(float, dpp_shr1(__builtin_bit_cast(int, v), __builtin_bit_cast(int, first)));
}
__device__ __forceinline__ float wave_shl1(float v, float last) {
  return __builtin_bit_cast(float, dpp_shl1(__builtin_bit_cast(int, v), __builtin_bit_cast(int, last)));
}
__device__ __forceinline__ double wave_shr1(double v, double first) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v), f = __builtin_bit_cast(uint64_t, first);
  const uint32_t lo = static_cast<uint32_t>(dpp_shr1(static_cast<int>(b), static_cast<int>(f)));
  const uint32_t hi = static_cast<uint32_t>(dpp_shr1(static_cast<int>(b >> 32), static_cast<int>(f >> 32)));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ double wave_shl1(double v, double last) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v), l = __builtin_bit_cast(uint64_t, last);
  const uint32_t lo = static_cast<uint32_t>(dpp_shl1(static_cast<int>(b), static_cast<int>(l)));
  const uint32_t hi = static_cast<uint32_t>(dpp_shl1(static_cast<int>(b >> 32), static_cast<int>(l >> 32)));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// x for R runs of three consecutive offsets (o-1, o, o+1), centres off[3w+1]:
// xl[3w + d] = x[row + o - 1 + d] for the calling lane, clamped into
// [0, last].  One vector load per run at its centre; the neighbours are that
// vector shifted by one lane, and the two values shifted in, x[row0+o-1] at
// lane 0 and x[row0+o+64] at lane 63, have wave-uniform addresses and are
// scalar loads.  They are issued after the vector loads: a scalar-load wait
// placed before a vector load would hold that load back.
template <typename T, bool X32, int R>
__device__ __forceinline__ void runs3_x(const T* __restrict__ x, const int* off, int64_t row0, int64_t row,
                                        int64_t last, T* xl) {
  T mid[R], first[R], after[R];
#pragma unroll
  for (int w = 0; w < R; ++w) {
    int64_t c = row + off[3 * w + 1];
    c = c < 0 ? 0 : (c > last ? last : c);
    mid[w] = load_x<T, X32>(x, c);
  }
#pragma unroll
  for (int w = 0; w < R; ++w) {
    const int64_t o = off[3 * w + 1];
    int64_t cf = row0 + o - 1, ce = row0 + o + kWave;
    cf = cf < 0 ? 0 : (cf > last ? last : cf);
    ce = ce < 0 ? 0 : (ce > last ? last : ce);
    first[w] = load_x<T, X32>(x, cf);
    after[w] = load_x<T, X32>(x, ce);
  }
#pragma unroll
  for (int w = 0; w < R; ++w) {
    xl[3 * w] = wave_shr1(mid[w], first[w]);
    xl[3 * w + 1] = mid[w];
    xl[3 * w + 2] = wave_shl1(mid[w], after[w]);
  }
}

// bdia_sum_uniform for a block whose offsets are runs of three consecutive
// values, as a 27-point stencil's are: a chunk of nine offsets is three runs
// (runs3_x).  The x values, skips and FMA order are exactly those of
// bdia_sum_uniform, with a third of its vector x loads.
template <typename T, bool X32>
__device__ __forceinline__ T bdia_sum_runs3(const T* __restrict__ vb, const int* __restrict__ op, int cnt, int bw,
                                            int vlane, const T* __restrict__ x, int64_t row0, int64_t row,
                                            int64_t nrows, bool active) {
  static_assert(kChunk % 3 == 0, "a chunk holds whole runs");
  T sum = T(0);
  const int64_t last = nrows - 1;
  for (int o0 = 0; o0 < cnt; o0 += kChunk) {
    int off[kChunk];
    T v[kChunk], xl[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) off[k] = op[o0 + k];  // padded array
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      v[k] = __builtin_nontemporal_load(at_u32(vb, static_cast<uint32_t>(min(o0 + k, cnt - 1) * bw + vlane)));
    // runs past the block's count read padding: clamped, their products unused
    runs3_x<T, X32, kChunk / 3>(x, off, row0, row, last, xl);
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const int64_t c = row + off[k];
      const bool ok = (o0 + k < cnt) & active & (v[k] != T(0)) & (c >= 0) & (c < nrows);
      sum = masked_fma(ok, v[k], xl[k], sum);
    }
  }
  return sum;
}

// The same for a row whose block is per lane (block width not a multiple of
// 64): `vp` is the lane's address in its block's first offset line, `op` the
// lane's offset list; lanes leave the loop at their own count.
template <typename T, bool X32, int CH = kChunk>
__device__ __forceinline__ T bdia_sum_lane(const T* __restrict__ vp, const int* __restrict__ op, int cnt,
                                           int64_t bw, const T* __restrict__ x, int64_t row, int64_t nrows) {
  T sum = T(0);
  const int64_t last = nrows - 1;
  for (int o0 = 0; o0 < cnt; o0 += CH) {
    int off[CH];
    T v[CH], xl[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) off[k] = op[min(o0 + k, cnt - 1)];
#pragma unroll
    for (int k = 0; k < CH; ++k)
      v[k] = __builtin_nontemporal_load(vp + static_cast<int64_t>(min(o0 + k, cnt - 1)) * bw);
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      int64_t c = row + off[k];
      c = c < 0 ? 0 : (c > last ? last : c);
      xl[k] = load_x<T, X32>(x, c);
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int64_t c = row + off[k];
      const bool ok = (o0 + k < cnt) & (v[k] != T(0)) & (c >= 0) & (c < nrows);
      sum = masked_fma(ok, v[k], xl[k], sum);
    }
  }
  return sum;
}


// The wave-uniform sums read a block's offsets in whole chunks (one wide
// scalar load each; the two-row sums read the first 3 kChunk entries at once),
// so they read up to 3 kChunk - 2 entries past the block's last offset.  In
// k_hbdia (every launch without the runs-of-three hints) a block that ends
// that close to the end of the offsets array (total entries) takes the
// per-lane sums instead, which clamp every offset read and have the same
// arithmetic, so those launches need no padding behind the array.  The
// two-row kernels, launched only with the hints, keep the wide reads with no
// test (a test there costs registers the kernels do not have): callers that
// pass hints also pass HBDIA_OFFSETS_PAD entries of padding.
__device__ __forceinline__ bool reads_past_end(int start, int cnt, int total) {
  // offsets counts are below 2^31 - 3 kChunk, so the sum does not overflow
  return start + max(cnt + 3 * kChunk - 2, 3 * kChunk) > total;
}

// The blocked part of one 64-row group (rows row0 + lane) for a block width
// that is not a multiple of 64; start and cnt are the calling lane's block's
// (zero past the last row).
template <typename T, bool X32, bool RUNS3, bool CHECK_END>
__device__ __forceinline__ T group_bdia_nonuniform(const T* __restrict__ vals, const int* __restrict__ offs,
                                                   const int* __restrict__ blk, const T* __restrict__ x,
                                                   int64_t row0, int lane, int64_t nrows, int nb, int bw, int64_t b,
                                                   int start, int cnt, const uint32_t* __restrict__ runs3_blocks,
                                                   int total) {
  const int64_t row = row0 + lane;
  const bool active = row < nrows;
  T sum = T(0);
  // Block width 32 (the reference's default): the wave's rows are blocks
  // b0 = row0 / 32 and b0 + 1.  When both exist and hold the same offsets,
  // the wave runs the uniform sums with a per-lane value index (block
  // b0 + 1's lines follow block b0's: lane l >= 32 reads line
  // pcnt + o at l - 32); otherwise each lane runs its own block.
  bool pair = false;
  int pstart = 0, pcnt = 0;
  int64_t b0 = 0;
  if (bw == 32) {
    b0 = row0 / 32;
    if (b0 + 1 < nb) {
      const int s0 = blk[b0], s1 = blk[b0 + 1], e1 = blk[b0 + 2];
      if (s1 - s0 == e1 - s1 && s1 > s0 && s1 - s0 <= kWave && !(CHECK_END && reads_past_end(s0, s1 - s0, total))) {
        const int c0 = s1 - s0;
        const int oa = lane < c0 ? offs[s0 + lane] : 0;
        const int ob = lane < c0 ? offs[s1 + lane] : 0;
        pair = __ballot(oa != ob) == 0;
        pstart = s0;
        pcnt = c0;
      }
    }
  }
  if (pair) {
    const T* vb = vals + static_cast<int64_t>(pstart) * 32;
    const int vlane = lane + (lane >= 32 ? (pcnt - 1) * 32 : 0);
    bool runs = false;
    if constexpr (RUNS3) runs = (runs3_blocks[b0 >> 5] >> (b0 & 31)) & 1u;
    if (runs) sum = bdia_sum_runs3<T, X32>(vb, offs + pstart, pcnt, 32, vlane, x, row0, row, nrows, active);
    else sum = bdia_sum_uniform<T, X32>(vb, offs + pstart, pcnt, 32, vlane, x, row, nrows, active);
  } else if (cnt > 0) {
    const T* vp = vals + static_cast<int64_t>(start) * bw + (row - b * bw);
    sum = bdia_sum_lane<T, X32>(vp, offs + start, cnt, bw, x, row, nrows);
  }
  return sum;
}

// Fused HBDIA SpMV.  UNIFORM: block_width % 64 == 0, so a wave's 64 rows lie
// in one block.  HAS_COO: the matrix has a non-empty COO fallback, and then
// every row receives y = (bdia or old y) + coo, as the reference's
// vector_add_kernel does for all rows.  X32: x fits 32-bit byte offsets.
// Workgroups take rows in runs per XCD (xcd_chunk_work).
template <typename T, bool UNIFORM, bool HAS_COO, bool X32, bool RUNS3>
__global__ __launch_bounds__(kWG) void k_hbdia(
    const T* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk,
    const T* __restrict__ x, T* __restrict__ y, int64_t nrows, int nb, int bw,
    const int* __restrict__ cr, const int* __restrict__ cc, const T* __restrict__ cv,
    const int* __restrict__ gptr, const uint32_t* __restrict__ runs3_blocks, int kshift) {
  __shared__ T s_v[kWavesPerWG][kWave];
  __shared__ T s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][kWave];
  __shared__ int s_hi[kWavesPerWG][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned work = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const int64_t grp = static_cast<int64_t>(work) * kWavesPerWG + wv;
  const int64_t row0 = grp * kWave;
  if (row0 >= nrows) return;
  const int64_t row = row0 + lane;
  const bool active = row < nrows;

  int64_t b;
  if constexpr (UNIFORM) b = row0 / bw;
  else b = row / bw;
  int start = 0, cnt = 0;
  if (b < nb && (UNIFORM || active)) {
    start = blk[b];
    cnt = blk[b + 1] - start;
  }
  const int total = blk[nb];
  T sum = T(0);
  if constexpr (UNIFORM) {
    // a block whose values exceed 32-bit byte offsets, or whose chunks would
    // read past the offsets array, takes the per-lane path
    const bool small = static_cast<int64_t>(cnt) * bw * static_cast<int64_t>(sizeof(T)) < (1ll << 32);
    if (cnt > 0 && small && !reads_past_end(start, cnt, total)) {
      const T* vb = vals + static_cast<int64_t>(start) * bw + (row0 - b * bw);
      bool runs = false;
      if constexpr (RUNS3) runs = (runs3_blocks[b >> 5] >> (b & 31)) & 1u;
      if (runs) sum = bdia_sum_runs3<T, X32>(vb, offs + start, cnt, bw, lane, x, row0, row, nrows, active);
      else sum = bdia_sum_uniform<T, X32>(vb, offs + start, cnt, bw, lane, x, row, nrows, active);
    } else if (cnt > 0 && active) {
      const T* vp = vals + static_cast<int64_t>(start) * bw + (row - b * bw);
      sum = bdia_sum_lane<T, X32>(vp, offs + start, cnt, bw, x, row, nrows);
    }
  } else {
    sum = group_bdia_nonuniform<T, X32, RUNS3, true>(vals, offs, blk, x, row0, lane, nrows, nb, bw, b, start, cnt,
                                                     runs3_blocks, total);
  }
  if constexpr (HAS_COO) {
    const int p0 = gptr[grp], p1 = gptr[grp + 1];
    T cs = T(0);
    bool any = false;
    if (p0 < p1)
      cs = coo_group_sum<T>(cr, cc, cv, x, p0, p1, row0, lane, s_v[wv], s_x[wv], s_lo[wv],
                            s_hi[wv], &any);
    if (active) {
      const T base = cnt > 0 ? sum : y[row];
      __builtin_nontemporal_store(base + cs, y + row);
    }
  } else {
    if (active && cnt > 0) __builtin_nontemporal_store(sum, y + row);
  }
}

// ---------------------------------------------------------------------------
// FP64, block width a multiple of 64, runs-of-three hints: 128-row waves with
// two adjacent rows per lane, so that a lane's two values of an offset, its
// two x values of a run's centre and (through LDS) its results move as
// 16-byte accesses.  Measured at 256^3 (no fallback): 601 vs 620 us.

typedef double dbl2a8 __attribute__((ext_vector_type(2), aligned(8)));
// Two adjacent elements as one access: 16 bytes for FP64, 8 for FP32; the
// alignment of one element, for loads at any x index.
typedef float flt2a4 __attribute__((ext_vector_type(2), aligned(4)));
template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef dbl2a8 type; };
template <> struct Vec2<float> { typedef flt2a4 type; };

// Rows row and row + 1 (row = row0 + 2 lane) of a 128-row wave whose rows share
// one offset list of runs of three.  vb + vlane addresses the lane's two values
// of the list's first offset; offset o of them is at vb + vlane + o * bw.  The x
// values, skips and FMA order per row are exactly those of bdia_sum_runs3.
template <typename T, int CH>
__device__ __forceinline__ void bdia_sum_runs3_row2(const T* __restrict__ vb, int vlane,
                                                    const int* __restrict__ op, int cnt, int bw,
                                                    const T* __restrict__ x, int64_t row0, int64_t row,
                                                    int64_t nrows, bool act_a, bool act_b, T& sa,
                                                    T& sb) {
  using V2 = typename Vec2<T>::type;
  const int64_t last = nrows - 1;
  sa = T(0);
  sb = T(0);
  for (int o0 = 0; o0 < cnt; o0 += CH) {
    int off[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) off[k] = op[o0 + k];  // padded array
    V2 v[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k)
      v[k] = __builtin_nontemporal_load(
          reinterpret_cast<const V2*>(vb + static_cast<uint32_t>(min(o0 + k, cnt - 1) * bw + vlane)));
    // every run's centre pair (one 16-byte load) and edge values, before any FMA
    V2 m[CH / 3];
    T first[CH / 3], after[CH / 3];
#pragma unroll
    for (int w = 0; w < CH / 3; ++w) {
      // runs past the block's count read padding: clamped, their products unused
      const int64_t o = off[3 * w + 1];
      if (row0 + o - 1 >= 0 && row0 + o + 2 * kWave <= last) {  // wave-uniform: the window lies inside x
        m[w] = *reinterpret_cast<const V2*>(x + (row + o));
      } else {
        int64_t c0 = row + o, c1 = row + 1 + o;
        c0 = c0 < 0 ? 0 : (c0 > last ? last : c0);
        c1 = c1 < 0 ? 0 : (c1 > last ? last : c1);
        m[w].x = x[c0];
        m[w].y = x[c1];
      }
    }
#pragma unroll
    for (int w = 0; w < CH / 3; ++w) {
      const int64_t o = off[3 * w + 1];
      int64_t cf = row0 + o - 1, ce = row0 + o + 2 * kWave;
      cf = cf < 0 ? 0 : (cf > last ? last : cf);
      ce = ce < 0 ? 0 : (ce > last ? last : ce);
      first[w] = x[cf];  // wave-uniform: scalar loads
      after[w] = x[ce];
    }
    // the FMAs run by run, in offset order
#pragma unroll
    for (int w = 0; w < CH / 3; ++w) {
      const T left = wave_shr1(m[w].y, first[w]);   // x[row + o - 1]
      const T right = wave_shl1(m[w].x, after[w]);  // x[row + o + 2]
      const T xa[3] = {left, m[w].x, m[w].y};
      const T xb[3] = {m[w].x, m[w].y, right};
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int k = 3 * w + j;
        const int64_t ca = row + off[k], cb = ca + 1;
        const bool oka = (o0 + k < cnt) & act_a & (v[k].x != T(0)) & (ca >= 0) & (ca < nrows);
        const bool okb = (o0 + k < cnt) & act_b & (v[k].y != T(0)) & (cb >= 0) & (cb < nrows);
        sa = masked_fma(oka, v[k].x, xa[j], sa);
        sb = masked_fma(okb, v[k].y, xb[j], sb);
      }
    }
  }
}

// One nine-offset chunk of the two-row runs-of-three sums, split into its
// loads and its FMAs so that the next chunk's loads can be issued first.
template <typename T>
struct Row2Chunk {
  using V2 = typename Vec2<T>::type;
  V2 v[kChunk];
  V2 m[kChunk / 3];
  T first[kChunk / 3], after[kChunk / 3];
};

template <typename T>
__device__ __forceinline__ void row2_chunk_load(Row2Chunk<T>& q, const T* __restrict__ vb, int vlane,
                                                const int* off, int o0, int cnt, int bw,
                                                const T* __restrict__ x, int64_t row0, int64_t row,
                                                int64_t last) {
  using V2 = typename Vec2<T>::type;
#pragma unroll
  for (int k = 0; k < kChunk; ++k)
    q.v[k] = __builtin_nontemporal_load(
        reinterpret_cast<const V2*>(vb + static_cast<uint32_t>(min(o0 + k, cnt - 1) * bw + vlane)));
#pragma unroll
  for (int w = 0; w < kChunk / 3; ++w) {
    // runs past the block's count read padding: clamped, their products unused
    const int64_t o = off[o0 + 3 * w + 1];
    if (row0 + o - 1 >= 0 && row0 + o + 2 * kWave <= last) {  // wave-uniform: the window lies inside x
      q.m[w] = *reinterpret_cast<const V2*>(x + (row + o));
    } else {
      int64_t c0 = row + o, c1 = row + 1 + o;
      c0 = c0 < 0 ? 0 : (c0 > last ? last : c0);
      c1 = c1 < 0 ? 0 : (c1 > last ? last : c1);
      q.m[w].x = x[c0];
      q.m[w].y = x[c1];
    }
  }
#pragma unroll
  for (int w = 0; w < kChunk / 3; ++w) {
    const int64_t o = off[o0 + 3 * w + 1];
    int64_t cf = row0 + o - 1, ce = row0 + o + 2 * kWave;
    cf = cf < 0 ? 0 : (cf > last ? last : cf);
    ce = ce < 0 ? 0 : (ce > last ? last : ce);
    q.first[w] = x[cf];  // wave-uniform: scalar loads
    q.after[w] = x[ce];
  }
}

// Row and column indices are below 2^31, so row + offset taken modulo 2^32
// lies in [0, nrows) exactly when the column does: one unsigned compare.
template <typename T>
__device__ __forceinline__ void row2_chunk_fma(const Row2Chunk<T>& q, const int* off, int o0, int cnt, int64_t row,
                                               int64_t nrows, bool act_a, bool act_b, T& sa, T& sb) {
  const uint32_t row32 = static_cast<uint32_t>(row), n32 = static_cast<uint32_t>(nrows);
#pragma unroll
  for (int w = 0; w < kChunk / 3; ++w) {
    const T left = wave_shr1(q.m[w].y, q.first[w]);   // x[row + o - 1]
    const T right = wave_shl1(q.m[w].x, q.after[w]);  // x[row + o + 2]
    const T xa[3] = {left, q.m[w].x, q.m[w].y};
    const T xb[3] = {q.m[w].x, q.m[w].y, right};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int k = 3 * w + j;
      const uint32_t ca = row32 + static_cast<uint32_t>(off[o0 + k]), cb = ca + 1u;
      const bool oka = (o0 + k < cnt) & act_a & (q.v[k].x != T(0)) & (ca < n32);
      const bool okb = (o0 + k < cnt) & act_b & (q.v[k].y != T(0)) & (cb < n32);
      sa = masked_fma(oka, q.v[k].x, xa[j], sa);
      sb = masked_fma(okb, q.v[k].y, xb[j], sb);
    }
  }
}

// bdia_sum_runs3_row2 for a block of at most 27 offsets: its three chunks
// unrolled, each chunk's loads issued before the previous chunk's FMAs, so a
// wave keeps the next nine value lines and x windows in flight while it
// computes.  Same arithmetic and order.
template <typename T>
__device__ __forceinline__ void bdia_sum_runs3_row2_27(const T* __restrict__ vb, int vlane,
                                                       const int* __restrict__ op, int cnt, int bw,
                                                       const T* __restrict__ x, int64_t row0, int64_t row,
                                                       int64_t nrows, bool act_a, bool act_b, T& sa,
                                                       T& sb) {
  const int64_t last = nrows - 1;
  int off[3 * kChunk];
#pragma unroll
  for (int k = 0; k < 3 * kChunk; ++k) off[k] = op[k];  // padded array
  sa = T(0);
  sb = T(0);
  Row2Chunk<T> c0, c1;
  row2_chunk_load(c0, vb, vlane, off, 0, cnt, bw, x, row0, row, last);
  if (cnt > kChunk) row2_chunk_load(c1, vb, vlane, off, kChunk, cnt, bw, x, row0, row, last);
  row2_chunk_fma(c0, off, 0, cnt, row, nrows, act_a, act_b, sa, sb);
  if (cnt > 2 * kChunk) row2_chunk_load(c0, vb, vlane, off, 2 * kChunk, cnt, bw, x, row0, row, last);
  if (cnt > kChunk) row2_chunk_fma(c1, off, kChunk, cnt, row, nrows, act_a, act_b, sa, sb);
  if (cnt > 2 * kChunk) row2_chunk_fma(c0, off, 2 * kChunk, cnt, row, nrows, act_a, act_b, sa, sb);
}

// The blocked part of one 64-row half (rows r0 + lane) of a block width that is
// a multiple of 64: k_hbdia<T, true, ...>'s arithmetic.
template <typename T, bool X32>
__device__ __forceinline__ T half_bdia(const T* __restrict__ vals, const int* __restrict__ offs,
                                            const T* __restrict__ x, int64_t r0, int64_t b, int start,
                                            int cnt, int bw, int lane, int64_t nrows, bool runs) {
  const int64_t row = r0 + lane;
  const bool active = row < nrows;
  const bool small = static_cast<int64_t>(cnt) * bw * static_cast<int64_t>(sizeof(T)) < (1ll << 32);
  if (cnt > 0 && small) {
    const T* vb = vals + static_cast<int64_t>(start) * bw + (r0 - b * bw);
    if (runs) return bdia_sum_runs3<T, X32>(vb, offs + start, cnt, bw, lane, x, r0, row, nrows, active);
    return bdia_sum_uniform<T, X32>(vb, offs + start, cnt, bw, lane, x, row, nrows, active);
  }
  if (cnt > 0 && active) {
    const T* vp = vals + static_cast<int64_t>(start) * bw + (row - b * bw);
    return bdia_sum_lane<T, X32>(vp, offs + start, cnt, bw, x, row, nrows);
  }
  return 0.0;
}

template <typename T, bool HAS_COO, bool X32>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(3))) void k_hbdia_row2(
    const T* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk,
    const T* __restrict__ x, T* __restrict__ y, int64_t nrows, int nb, int bw,
    const int* __restrict__ cr, const int* __restrict__ cc, const T* __restrict__ cv,
    const int* __restrict__ gptr, const uint32_t* __restrict__ runs3_blocks, int kshift) {
  __shared__ T s_v[kWavesPerWG][kWave];
  __shared__ T s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][2 * kWave];
  __shared__ int s_hi[kWavesPerWG][2 * kWave];
  __shared__ T s_y[kWavesPerWG][2 * kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned work = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const int64_t w2 = static_cast<int64_t>(work) * kWavesPerWG + wv;  // 128-row wave
  const int64_t row0 = w2 * 2 * kWave;
  if (row0 >= nrows) return;
  // the two 64-row halves: rows row0 + lane (A) and row0 + 64 + lane (B)
  const int64_t bA = row0 / bw, bB = (row0 + kWave) / bw;
  const bool hasB = row0 + kWave < nrows;
  int sA = 0, cA = 0, sB = 0, cB = 0;
  if (bA < nb) {
    sA = blk[bA];
    cA = blk[bA + 1] - sA;
  }
  if (hasB && bB < nb) {
    sB = blk[bB];
    cB = blk[bB + 1] - sB;
  }
  const bool runsA = bA < nb && ((runs3_blocks[bA >> 5] >> (bA & 31)) & 1u);
  const bool runsB = hasB && bB < nb && ((runs3_blocks[bB >> 5] >> (bB & 31)) & 1u);
  // both halves in one offset list of runs: one block, or two with equal offsets
  bool same = false;
  if (runsA && cA > 0 && cA <= kWave) {
    if (bA == bB) {
      same = true;
    } else if (hasB && cB == cA) {
      const int oa = lane < cA ? offs[sA + lane] : 0;
      const int ob = lane < cA ? offs[sB + lane] : 0;
      same = __ballot(oa != ob) == 0;
    }
  }
  T yA = T(0), yB = T(0);
  if (same) {
    const int64_t row = row0 + 2 * lane;
    const T* vb = vals + static_cast<int64_t>(sA) * bw + (row0 - bA * bw);
    // lanes of half B read block bB's lines, which follow block bA's
    const int vlane = 2 * lane + ((lane >= kWave / 2 && bB != bA) ? (cA - 1) * bw : 0);
    T sa, sb;
    if (cA <= 3 * kChunk)
      bdia_sum_runs3_row2_27(vb, vlane, offs + sA, cA, bw, x, row0, row, nrows, row < nrows, row + 1 < nrows, sa, sb);
    else
      bdia_sum_runs3_row2<T, kRow2Chunk>(vb, vlane, offs + sA, cA, bw, x, row0, row, nrows, row < nrows,
                                      row + 1 < nrows, sa, sb);
    // the fallback of both groups in the two-row layout when it fits one pass
    // (or there is none): each lane stores its two adjacent rows as one access
    {
      int q0 = 0, q1 = 0, q2 = 0;
      if constexpr (HAS_COO) {
        q0 = gptr[2 * w2];
        q1 = gptr[2 * w2 + 1];
        q2 = hasB ? gptr[2 * w2 + 2] : q1;
      }
      if (!HAS_COO || q2 - q0 <= kWave) {
        using V2 = typename Vec2<T>::type;
        T ta = sa, tb = sb;
        if constexpr (HAS_COO) {
          T ca = T(0), cb = T(0);
          if (q0 < q2)
            coo_pair_sum<T, true>(cr, cc, cv, x, q0, q1, q2, row0, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv], ca,
                                  cb);
          ta = sa + ca;  // y = bdia + coo for every row, as below
          tb = sb + cb;
        }
        if (row + 1 < nrows) {
          V2 t;
          t.x = ta;
          t.y = tb;
          __builtin_nontemporal_store(t, reinterpret_cast<V2*>(y + row));
        } else if (row < nrows) {
          __builtin_nontemporal_store(ta, y + row);
        }
        return;
      }
    }
    // back to the halves' row-per-lane order
    T* sy = s_y[wv];
    sy[2 * lane] = sa;
    sy[2 * lane + 1] = sb;
    wave_sync();
    yA = sy[lane];
    yB = sy[kWave + lane];
    wave_sync();
  } else {
    yA = half_bdia<T, X32>(vals, offs, x, row0, bA, sA, cA, bw, lane, nrows, runsA);
    if (hasB) yB = half_bdia<T, X32>(vals, offs, x, row0 + kWave, bB, sB, cB, bw, lane, nrows, runsB);
  }
  // fallback sums of both 64-row groups: one pass when their entries fit one
  T csA = T(0), csB = T(0);
  if constexpr (HAS_COO) {
    const int p0 = gptr[2 * w2], p1 = gptr[2 * w2 + 1];
    const int p2 = hasB ? gptr[2 * w2 + 2] : p1;
    bool any = false;
    if (p2 - p0 <= kWave) {
      if (p0 < p2) coo_pair_sum<T>(cr, cc, cv, x, p0, p1, p2, row0, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv],
                                        csA, csB);
    } else {
      if (p0 < p1)
        csA = coo_group_sum<T>(cr, cc, cv, x, p0, p1, row0, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv], &any);
      if (p1 < p2)
        csB = coo_group_sum<T>(cr, cc, cv, x, p1, p2, row0 + kWave, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv],
                                    &any);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t r0 = row0 + h * kWave;
    if (h == 1 && !hasB) break;
    const int64_t row = r0 + lane;
    const bool active = row < nrows;
    const int cnt = h ? cB : cA;
    const T sum = h ? yB : yA;
    if constexpr (HAS_COO) {
      const T cs = h ? csB : csA;
      if (active) {
        const T base = cnt > 0 ? sum : y[row];
        __builtin_nontemporal_store(base + cs, y + row);
      }
    } else {
      if (active && cnt > 0) __builtin_nontemporal_store(sum, y + row);
    }
  }
}

// FP64, block width 32 (the reference's default), runs-of-three hints: the
// two-row waves of k_hbdia_row2 over four consecutive width-32 blocks.  When
// the four blocks exist, hold the same number of offsets and the same offsets
// (nearly every wave of a stencil), lane l's rows row0 + 2l and row0 + 2l + 1
// lie in block b0 + l / 16, whose lines follow block b0's by (l / 16) * cnt
// lines: the runs-of-three two-row sums run with that per-lane value index.
// Otherwise each 64-row half takes k_hbdia's width-32 path.  Values, skips and
// FMA order per row are those of k_hbdia<T, false, ...>.
template <typename T, bool HAS_COO, bool X32>
__global__ __launch_bounds__(kWG) void k_hbdia_row2_w32(
    const T* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk,
    const T* __restrict__ x, T* __restrict__ y, int64_t nrows, int nb,
    const int* __restrict__ cr, const int* __restrict__ cc, const T* __restrict__ cv,
    const int* __restrict__ gptr, const uint32_t* __restrict__ runs3_blocks, int kshift) {
  __shared__ T s_v[kWavesPerWG][kWave];
  __shared__ T s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][2 * kWave];
  __shared__ int s_hi[kWavesPerWG][2 * kWave];
  __shared__ T s_y[kWavesPerWG][2 * kWave];
  constexpr int bw = 32;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned work = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const int64_t w2 = static_cast<int64_t>(work) * kWavesPerWG + wv;  // 128-row wave
  const int64_t row0 = w2 * 2 * kWave;
  if (row0 >= nrows) return;
  const bool hasB = row0 + kWave < nrows;
  const int64_t b0 = row0 / bw;
  bool quad = false;
  int s0 = 0, c = 0;
  if (b0 + 3 < nb) {
    s0 = blk[b0];
    c = blk[b0 + 1] - s0;
    if (c > 0 && c <= kWave && blk[b0 + 2] == s0 + 2 * c && blk[b0 + 3] == s0 + 3 * c &&
        blk[b0 + 4] == s0 + 4 * c && ((runs3_blocks[b0 >> 5] >> (b0 & 31)) & 1u)) {
      const int oa = lane < c ? offs[s0 + lane] : 0;
      bool diff = false;
#pragma unroll
      for (int j = 1; j < 4; ++j) diff = diff | (oa != (lane < c ? offs[s0 + j * c + lane] : 0));
      quad = __ballot(diff) == 0;
    }
  }
  T yA = T(0), yB = T(0);
  int cntA = 0, cntB = 0;  // the lane's block's offset count, per half
  if (quad) {
    const int64_t row = row0 + 2 * lane;
    const T* vb = vals + static_cast<int64_t>(s0) * bw;
    const int vlane = 2 * lane + (lane >> 4) * (c - 1) * bw;
    T sa, sb;
    if (c <= 3 * kChunk)
      bdia_sum_runs3_row2_27(vb, vlane, offs + s0, c, bw, x, row0, row, nrows, row < nrows, row + 1 < nrows, sa, sb);
    else
      bdia_sum_runs3_row2<T, kRow2Chunk>(vb, vlane, offs + s0, c, bw, x, row0, row, nrows, row < nrows,
                                      row + 1 < nrows, sa, sb);
    // the fallback of both groups in the two-row layout when it fits one pass
    // (or there is none): each lane stores its two adjacent rows as one access
    {
      int q0 = 0, q1 = 0, q2 = 0;
      if constexpr (HAS_COO) {
        q0 = gptr[2 * w2];
        q1 = gptr[2 * w2 + 1];
        q2 = hasB ? gptr[2 * w2 + 2] : q1;
      }
      if (!HAS_COO || q2 - q0 <= kWave) {
        using V2 = typename Vec2<T>::type;
        T ta = sa, tb = sb;
        if constexpr (HAS_COO) {
          T ca = T(0), cb = T(0);
          if (q0 < q2)
            coo_pair_sum<T, true>(cr, cc, cv, x, q0, q1, q2, row0, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv], ca,
                                  cb);
          ta = sa + ca;  // y = bdia + coo for every row, as below
          tb = sb + cb;
        }
        if (row + 1 < nrows) {
          V2 t;
          t.x = ta;
          t.y = tb;
          __builtin_nontemporal_store(t, reinterpret_cast<V2*>(y + row));
        } else if (row < nrows) {
          __builtin_nontemporal_store(ta, y + row);
        }
        return;
      }
    }
    T* sy = s_y[wv];
    sy[2 * lane] = sa;
    sy[2 * lane + 1] = sb;
    wave_sync();
    yA = sy[lane];
    yB = sy[kWave + lane];
    wave_sync();
    cntA = row0 + lane < nrows ? c : 0;
    cntB = row0 + kWave + lane < nrows ? c : 0;
  } else {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && !hasB) break;
      const int64_t r0 = row0 + h * kWave;
      const int64_t row = r0 + lane;
      const int64_t b = row / bw;
      int start = 0, cnt = 0;
      if (b < nb && row < nrows) {
        start = blk[b];
        cnt = blk[b + 1] - start;
      }
      const T sum =
          group_bdia_nonuniform<T, X32, true, false>(vals, offs, blk, x, r0, lane, nrows, nb, bw, b, start, cnt,
                                                          runs3_blocks, 0);
      if (h == 0) {
        yA = sum;
        cntA = cnt;
      } else {
        yB = sum;
        cntB = cnt;
      }
    }
  }
  // fallback sums of both 64-row groups: one pass when their entries fit one
  T csA = T(0), csB = T(0);
  if constexpr (HAS_COO) {
    const int p0 = gptr[2 * w2], p1 = gptr[2 * w2 + 1];
    const int p2 = hasB ? gptr[2 * w2 + 2] : p1;
    bool any = false;
    if (p2 - p0 <= kWave) {
      if (p0 < p2) coo_pair_sum<T>(cr, cc, cv, x, p0, p1, p2, row0, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv],
                                        csA, csB);
    } else {
      if (p0 < p1)
        csA = coo_group_sum<T>(cr, cc, cv, x, p0, p1, row0, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv], &any);
      if (p1 < p2)
        csB = coo_group_sum<T>(cr, cc, cv, x, p1, p2, row0 + kWave, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv],
                                    &any);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t r0 = row0 + h * kWave;
    if (h == 1 && !hasB) break;
    const int64_t row = r0 + lane;
    const bool active = row < nrows;
    const int cnt = h ? cntB : cntA;
    const T sum = h ? yB : yA;
    if constexpr (HAS_COO) {
      const T cs = h ? csB : csA;
      if (active) {
        const T base = cnt > 0 ? sum : y[row];
        __builtin_nontemporal_store(base + cs, y + row);
      }
    } else {
      if (active && cnt > 0) __builtin_nontemporal_store(sum, y + row);
    }
  }
}

// ---------------------------------------------------------------------------
// Multiple right-hand sides: Y[:, j] = A X[:, j] for j < nrhs, with X and Y
// column-major (column j at X + j*ldx, Y + j*ldy).  Values are loaded once per
// (block, offset) and applied to KB columns; each column's arithmetic is the
// single-vector kernel's (same FMA order, same skips), so every column equals
// a single-vector SpMV bit for bit.  Block width must be a multiple of 64.

// Fallback sums of the calling lane's row for KB columns (entry order, FMA).
template <typename T, int KB>
__device__ __forceinline__ void coo_group_sum_multi(const int* __restrict__ cr, const int* __restrict__ cc,
                                                    const T* __restrict__ cv, const T* __restrict__ X,
                                                    int64_t ldx, int kb, int p0, int p1, int64_t row0, int lane,
                                                    T* s_v, T* s_x, int* s_lo, int* s_hi, T* cs) {
#pragma unroll
  for (int j = 0; j < KB; ++j) cs[j] = T(0);
  for (int base = p0; base < p1; base += kWave) {
    const int i = base + lane;
    const bool ok = i < p1;
    int r = -1, col = 0;
    T vv = T(0);
    if (ok) {
      r = cr[i] - static_cast<int>(row0);
      col = cc[i];
      vv = cv[i];
      if (r < 0 || r >= kWave) r = -1;
    }
    const int rp = __shfl_up(r, 1);
    const int rn = __shfl_down(r, 1);
    s_lo[lane] = 0;
    s_hi[lane] = 0;
    s_v[lane] = vv;
    wave_sync();
    if (ok && r >= 0 && (lane == 0 || rp != r)) s_lo[r] = lane;
    if (ok && r >= 0 && (lane == kWave - 1 || i + 1 >= p1 || rn != r)) s_hi[r] = lane + 1;
    wave_sync();
    const int a = s_lo[lane], e = s_hi[lane];
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      if (j >= kb) break;
      s_x[lane] = ok ? X[static_cast<int64_t>(j) * ldx + col] : T(0);
      wave_sync();
      T acc = cs[j];
      for (int k = a; k < e; ++k) acc = fma_t<T>(s_v[k], s_x[k], acc);
      cs[j] = acc;
      wave_sync();
    }
  }
}

template <int KB> struct MrhsChunk { static constexpr int value = KB >= 16 ? 1 : (KB >= 8 ? 2 : (KB >= 4 ? 3 : 5)); };
// runs of three per chunk on the runs-of-three path
template <int KB> struct MrhsRuns { static constexpr int value = KB >= 4 ? 1 : 2; };

// One 64-row group (rows row0 + lane) of the multi-RHS product for columns
// [j0, j0 + kb): the blocked part and the fused fallback, stored into Y.
template <typename T, int KB, bool HAS_COO, bool X32, bool RUNS3>
__device__ __forceinline__ void multi_group(const T* __restrict__ vals, const int* __restrict__ offs,
                                            const int* __restrict__ blk, const T* __restrict__ X, int64_t ldx,
                                            T* __restrict__ Y, int64_t ldy, int64_t nrows, int nb, int bw,
                                            const int* __restrict__ cr, const int* __restrict__ cc,
                                            const T* __restrict__ cv, const int* __restrict__ gptr,
                                            const uint32_t* __restrict__ runs3_blocks, int64_t grp, int j0, int kb,
                                            int lane, T* s_v_w, T* s_x_w, int* s_lo_w, int* s_hi_w) {
  constexpr int CH = MrhsChunk<KB>::value;
  constexpr int RR = MrhsRuns<KB>::value;
  const int64_t row0 = grp * kWave;
  const T* Xj = X + static_cast<int64_t>(j0) * ldx;
  T* Yj = Y + static_cast<int64_t>(j0) * ldy;
  const int64_t row = row0 + lane;
  const bool active = row < nrows;
  const int64_t last = nrows - 1;
  const int64_t b = row0 / bw;
  int start = 0, cnt = 0;
  if (b < nb) {
    start = blk[b];
    cnt = blk[b + 1] - start;
  }
  T sum[KB];
#pragma unroll
  for (int j = 0; j < KB; ++j) sum[j] = T(0);
  bool runs = false;
  if constexpr (RUNS3) runs = cnt > 0 && ((runs3_blocks[b >> 5] >> (b & 31)) & 1u);
  if (runs) {
    // a multiple of three offsets, each triple consecutive: per column one vector
    // x load per run (runs3_x); same x values, skips and FMA order as below
    const T* vb = vals + static_cast<int64_t>(start) * bw + (row0 - b * bw);
    const int* op = offs + start;
    const uint32_t row32 = static_cast<uint32_t>(row), n32 = static_cast<uint32_t>(nrows);
    const bool small = static_cast<int64_t>(cnt) * bw * static_cast<int64_t>(sizeof(T)) < (1ll << 32);
    for (int o0 = 0; o0 < cnt; o0 += 3 * RR) {
      int off[3 * RR];
      T v[3 * RR];
#pragma unroll
      for (int k = 0; k < 3 * RR; ++k) off[k] = op[o0 + k];  // padded array
#pragma unroll
      for (int k = 0; k < 3 * RR; ++k) {
        const int kk = min(o0 + k, cnt - 1);
        v[k] = small ? __builtin_nontemporal_load(at_u32(vb, static_cast<uint32_t>(kk * bw + lane)))
                     : __builtin_nontemporal_load(vb + static_cast<int64_t>(kk) * bw + lane);
      }
      // every column's x windows for the chunk before any FMA
      T mid[KB][RR], first[KB][RR], after[KB][RR];
      uint32_t cm[RR];
#pragma unroll
      for (int w = 0; w < RR; ++w) {
        // rows are below 2^31: row + offset modulo 2^32 is below n32 exactly when
        // the column is in range; out of range, the offset's sign picks the clamp
        const int o = off[3 * w + 1];
        const uint32_t c = row32 + static_cast<uint32_t>(o);
        cm[w] = c < n32 ? c : (o < 0 ? 0u : n32 - 1u);
      }
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        if (j < kb) {
          const T* xj = Xj + static_cast<int64_t>(j) * ldx;
#pragma unroll
          for (int w = 0; w < RR; ++w) mid[j][w] = load_x<T, X32>(xj, cm[w]);
        }
      }
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        if (j < kb) {
          const T* xj = Xj + static_cast<int64_t>(j) * ldx;
#pragma unroll
          for (int w = 0; w < RR; ++w) {
            const int64_t o = off[3 * w + 1];
            int64_t cf = row0 + o - 1, ce = row0 + o + kWave;
            cf = cf < 0 ? 0 : (cf > last ? last : cf);
            ce = ce < 0 ? 0 : (ce > last ? last : ce);
            first[j][w] = load_x<T, X32>(xj, cf);
            after[j][w] = load_x<T, X32>(xj, ce);
          }
        }
      }
      // rows past the end compute sums that are never stored
      bool ok[3 * RR];
#pragma unroll
      for (int k = 0; k < 3 * RR; ++k)
        ok[k] = (o0 + k < cnt) & (v[k] != T(0)) & (row32 + static_cast<uint32_t>(off[k]) < n32);
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        if (j >= kb) break;
#pragma unroll
        for (int w = 0; w < RR; ++w) {
          const T xl[3] = {wave_shr1(mid[j][w], first[j][w]), mid[j][w], wave_shl1(mid[j][w], after[j][w])};
#pragma unroll
          for (int i = 0; i < 3; ++i) sum[j] = masked_fma(ok[3 * w + i], v[3 * w + i], xl[i], sum[j]);
        }
      }
    }
  } else if (cnt > 0) {
    const T* vb = vals + static_cast<int64_t>(start) * bw + (row0 - b * bw);
    const int* op = offs + start;
    const bool small = static_cast<int64_t>(cnt) * bw * static_cast<int64_t>(sizeof(T)) < (1ll << 32);
    for (int o0 = 0; o0 < cnt; o0 += CH) {
      int off[CH];
      T v[CH];
      int64_t cl[CH];
#pragma unroll
      for (int k = 0; k < CH; ++k) off[k] = op[o0 + k];
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        const int kk = min(o0 + k, cnt - 1);
        v[k] = small ? __builtin_nontemporal_load(at_u32(vb, static_cast<uint32_t>(kk * bw + lane)))
                     : __builtin_nontemporal_load(vb + static_cast<int64_t>(kk) * bw + lane);
        int64_t c = row + off[k];
        cl[k] = c < 0 ? 0 : (c > last ? last : c);
      }
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        if (j >= kb) break;
        const T* xj = Xj + static_cast<int64_t>(j) * ldx;
        T xl[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) xl[k] = load_x<T, X32>(xj, cl[k]);
#pragma unroll
        for (int k = 0; k < CH; ++k) {
          const int64_t c = row + off[k];
          const bool ok = (o0 + k < cnt) & active & (v[k] != T(0)) & (c >= 0) & (c < nrows);
          sum[j] = masked_fma(ok, v[k], xl[k], sum[j]);
        }
      }
    }
  }
  if constexpr (HAS_COO) {
    const int p0 = gptr[grp], p1 = gptr[grp + 1];
    T cs[KB];
    if (p0 < p1) {
      coo_group_sum_multi<T, KB>(cr, cc, cv, Xj, ldx, kb, p0, p1, row0, lane, s_v_w, s_x_w, s_lo_w, s_hi_w, cs);
    } else {
#pragma unroll
      for (int j = 0; j < KB; ++j) cs[j] = T(0);
    }
    if (active) {
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        if (j >= kb) break;
        T* yp = Yj + static_cast<int64_t>(j) * ldy + row;
        const T base = cnt > 0 ? sum[j] : *yp;
        __builtin_nontemporal_store(base + cs[j], yp);
      }
    }
  } else {
    if (active && cnt > 0) {
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        if (j >= kb) break;
        __builtin_nontemporal_store(sum[j], Yj + static_cast<int64_t>(j) * ldy + row);
      }
    }
  }
}

template <typename T, int KB, bool HAS_COO, bool X32, bool RUNS3>
__global__ __launch_bounds__(kWG) void k_hbdia_multi(
    const T* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk,
    const T* __restrict__ X, int64_t ldx, T* __restrict__ Y, int64_t ldy, int nrhs, int64_t nrows, int nb,
    int bw, const int* __restrict__ cr, const int* __restrict__ cc, const T* __restrict__ cv,
    const int* __restrict__ gptr, const uint32_t* __restrict__ runs3_blocks, int kshift) {
  __shared__ T s_v[kWavesPerWG][kWave];
  __shared__ T s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][kWave];
  __shared__ int s_hi[kWavesPerWG][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // one workgroup per (row block, column group): a row block's column groups are
  // consecutive work items, which the XCD runs keep on one XCD, so the later
  // groups read the block's values from that XCD's L2 instead of HBM
  const unsigned passes = static_cast<unsigned>((nrhs + KB - 1) / KB);
  const unsigned q = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const unsigned work = q / passes;
  const int64_t grp = static_cast<int64_t>(work) * kWavesPerWG + wv;
  if (grp * kWave >= nrows) return;
  const int j0 = static_cast<int>(q - work * passes) * KB;
  const int kb = min(KB, nrhs - j0);
  multi_group<T, KB, HAS_COO, X32, RUNS3>(vals, offs, blk, X, ldx, Y, ldy, nrows, nb, bw, cr, cc, cv, gptr,
                                          runs3_blocks, grp, j0, kb, lane, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv]);
}

// __builtin_amdgcn_class mask of the finite non-zero doubles (normal and
// subnormal, either sign).
constexpr int kFiniteNonzero = 0x198;

// Element i of one column of X: a 32-bit byte offset from the column's
// (uniform) base when the column fits 4 GiB.
template <bool X32>
__device__ __forceinline__ const double* x_at(const double* xj, int64_t i) {
  if constexpr (X32)
    return reinterpret_cast<const double*>(reinterpret_cast<const char*>(xj) +
                                           static_cast<uint64_t>(static_cast<uint32_t>(i) * 8u));
  return xj + i;
}

// Whole-wave shifts with zero fill: lane i gets lane i - 1's value (lane 0
// gets 0), or lane i + 1's (lane 63 gets 0).
__device__ __forceinline__ double wave_shr1_zero(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(b), 0x138, 0xF, 0xF, true));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), 0x138, 0xF, 0xF, true));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ double wave_shl1_zero(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(b), 0x130, 0xF, 0xF, true));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), 0x130, 0xF, 0xF, true));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// The blocked sums of a 128-row wave's two adjacent rows per lane (row, row +
// 1) for kb <= KB columns of X (column j at Xj + j ldx; columns past kb repeat
// column kb - 1 and are not stored), over an offset list of runs of three
// whose count is a multiple of 3.  The rows' values of an offset are one
// 16-byte load, shared by the columns.  Per column and run: the centre pair
// is one 16-byte load, the outer neighbours come by zero-fill DPP shifts from
// the adjacent lanes, and the two values beyond the wave's ends (left of lane
// 0's rows, right of lane 63's) are scalar loads, each added by one more FMA
// whose value is zero outside its lane.
//
// Every term that the exact sum skips (a zero value, a column outside the
// matrix) or that is added in place of another (a shifted-in zero, an edge
// FMA in another lane) is a zero value times an x, or a value times zero.
// With that x finite it adds a signed zero, which leaves a non-zero sum
// unchanged and a zero sum zero; an infinite or NaN x makes the sum NaN.  The
// remaining terms are the exact sum's FMAs in its order.  So a sum that ends
// finite and non-zero equals the exact sum bit for bit; the caller recomputes
// a wave in which any stored sum does not.  A run whose columns all lie inside
// the matrix for every lane of the wave needs no per-term test at all.
// One run's loads for multi_row2_fast.
template <int KB>
struct MultiRun {
  dbl2a8 v[3];   // the rows' values of the run's offsets
  dbl2a8 m[KB];  // per column: x at the run's centre for both rows
  double el[KB], er[KB];  // per column: x left of lane 0's rows, right of lane 63's
  int off[3];
  bool inside;   // every column of the run inside the matrix for every lane
};

template <int KB, bool X32, bool NT>
__device__ __forceinline__ void multi_run_load(MultiRun<KB>& q, const double* __restrict__ vb, int64_t vlane,
                                               const int* __restrict__ op, int o0, int bw,
                                               const double* const* xc, int64_t row0, int64_t row,
                                               int64_t last) {
#pragma unroll
  for (int k = 0; k < 3; ++k) q.off[k] = op[o0 + k];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const dbl2a8* p = reinterpret_cast<const dbl2a8*>(vb + static_cast<int64_t>(o0 + k) * bw + vlane);
    q.v[k] = NT ? __builtin_nontemporal_load(p) : *p;
  }
  const int64_t o = q.off[1];
  const int64_t lo = row0 + o - 1, hi = row0 + o + 2 * kWave;  // wave-uniform
  q.inside = lo >= 0 && hi <= last;
  const int64_t cl = lo < 0 ? 0 : (lo > last ? last : lo), ch = hi < 0 ? 0 : (hi > last ? last : hi);
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    q.el[j] = *x_at<X32>(xc[j], cl);  // scalar loads
    q.er[j] = *x_at<X32>(xc[j], ch);
  }
  if (q.inside) {
#pragma unroll
    for (int j = 0; j < KB; ++j) q.m[j] = *reinterpret_cast<const dbl2a8*>(x_at<X32>(xc[j], row + o));
  } else {
    int64_t c0 = row + o, c1 = row + 1 + o;
    c0 = c0 < 0 ? 0 : (c0 > last ? last : c0);
    c1 = c1 < 0 ? 0 : (c1 > last ? last : c1);
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      q.m[j].x = *x_at<X32>(xc[j], c0);
      q.m[j].y = *x_at<X32>(xc[j], c1);
    }
  }
}

// The FMAs of one run: va / vbb are the rows' values, vl / vr those of the
// edge terms (x[lo] for lane 0's first row, x[hi] for lane 63's second).
template <int KB>
__device__ __forceinline__ void multi_run_fma(const MultiRun<KB>& q, const double* va, const double* vbb, int lane,
                                              double* sa, double* sb) {
  const double vl = lane == 0 ? va[0] : 0.0;
  const double vr = lane == kWave - 1 ? vbb[2] : 0.0;
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    const double left = wave_shr1_zero(q.m[j].y), right = wave_shl1_zero(q.m[j].x);
    sa[j] = __builtin_fma(va[0], left, sa[j]);
    sa[j] = __builtin_fma(vl, q.el[j], sa[j]);
    sa[j] = __builtin_fma(va[1], q.m[j].x, sa[j]);
    sa[j] = __builtin_fma(va[2], q.m[j].y, sa[j]);
    sb[j] = __builtin_fma(vbb[0], q.m[j].x, sb[j]);
    sb[j] = __builtin_fma(vbb[1], q.m[j].y, sb[j]);
    sb[j] = __builtin_fma(vbb[2], right, sb[j]);
    sb[j] = __builtin_fma(vr, q.er[j], sb[j]);
  }
}

template <int KB>
__device__ __forceinline__ void multi_run_sum(const MultiRun<KB>& q, uint32_t row32, uint32_t n32, int lane,
                                              double* sa, double* sb) {
  if (q.inside) {
    const double va[3] = {q.v[0].x, q.v[1].x, q.v[2].x}, vbb[3] = {q.v[0].y, q.v[1].y, q.v[2].y};
    multi_run_fma<KB>(q, va, vbb, lane, sa, sb);
  } else {
    // columns outside the matrix: their values as zeros
    double va[3], vbb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t ca = row32 + static_cast<uint32_t>(q.off[k]);
      va[k] = ca < n32 ? q.v[k].x : 0.0;
      vbb[k] = ca + 1u < n32 ? q.v[k].y : 0.0;
    }
    multi_run_fma<KB>(q, va, vbb, lane, sa, sb);
  }
}

// The blocked sums of a 128-row wave's two adjacent rows per lane (row, row +
// 1) for kb <= KB columns of X (column j at Xj + j ldx; columns past kb repeat
// column kb - 1 and are not stored), over an offset list of runs of three
// whose count is a multiple of 3.  The rows' values of an offset are one
// 16-byte load, shared by the columns (non-temporal when NT: no later column
// pass reads them).  Per column and run: the centre pair is one 16-byte load,
// the outer neighbours come by zero-fill DPP shifts from the adjacent lanes,
// and the two values beyond the wave's ends (left of lane 0's rows, right of
// lane 63's) are scalar loads, each added by one more FMA whose value is zero
// outside its lane.  (Issuing each run's loads before the previous run's FMAs
// measured no faster: the launch is bound by L2 requests, not latency.)
//
// Every term that the exact sum skips (a zero value, a column outside the
// matrix) or that is added in place of another (a shifted-in zero, an edge
// FMA in another lane) is a zero value times an x, or a value times zero.
// With that x finite it adds a signed zero, which leaves a non-zero sum
// unchanged and a zero sum zero; an infinite or NaN x makes the sum NaN.  The
// remaining terms are the exact sum's FMAs in its order.  So a sum that ends
// finite and non-zero equals the exact sum bit for bit; the caller recomputes
// a wave in which any stored sum does not.  A run whose columns all lie inside
// the matrix for every lane of the wave needs no per-term test at all.
template <int KB, bool X32, bool NT>
__device__ __forceinline__ void multi_row2_fast(const double* __restrict__ vb, int64_t vlane,
                                                const int* __restrict__ op, int cnt, int bw,
                                                const double* __restrict__ Xj, int64_t ldx, int kb, int64_t row0,
                                                int64_t row, int64_t nrows, int lane, double* sa, double* sb) {
  const int64_t last = nrows - 1;
  const uint32_t row32 = static_cast<uint32_t>(row), n32 = static_cast<uint32_t>(nrows);
  const double* xc[KB];
#pragma unroll
  for (int j = 0; j < KB; ++j) xc[j] = Xj + static_cast<int64_t>(min(j, kb - 1)) * ldx;
#pragma unroll
  for (int j = 0; j < KB; ++j) sa[j] = sb[j] = 0.0;
  for (int o0 = 0; o0 < cnt; o0 += 3) {  // one run of three per step
    MultiRun<KB> q;
    multi_run_load<KB, X32, NT>(q, vb, vlane, op, o0, bw, xc, row0, row, last);
    multi_run_sum<KB>(q, row32, n32, lane, sa, sb);
  }
}

// multi_row2_fast's sums with every skip of the exact order applied per term.
template <int KB>
__device__ __forceinline__ void multi_row2_exact(const double* __restrict__ vb, int64_t vlane,
                                                 const int* __restrict__ op, int cnt, int bw,
                                                 const double* __restrict__ Xj, int64_t ldx, int kb, int64_t row0,
                                                 int64_t row, int64_t nrows, double* sa, double* sb) {
  const int64_t last = nrows - 1;
  const uint32_t row32 = static_cast<uint32_t>(row), n32 = static_cast<uint32_t>(nrows);
#pragma unroll
  for (int j = 0; j < KB; ++j) sa[j] = sb[j] = 0.0;
  for (int o0 = 0; o0 < cnt; o0 += 3) {  // one run of three per step
    int off[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) off[k] = op[o0 + k];  // padded array
    dbl2a8 v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      v[k] = __builtin_nontemporal_load(
          reinterpret_cast<const dbl2a8*>(vb + static_cast<int64_t>(min(o0 + k, cnt - 1)) * bw + vlane));
    // every column's centre pair and edge values before any FMA
    const int64_t o = off[1];
    const bool inside = row0 + o - 1 >= 0 && row0 + o + 2 * kWave <= last;  // wave-uniform
    int64_t cf = row0 + o - 1, ce = row0 + o + 2 * kWave;
    cf = cf < 0 ? 0 : (cf > last ? last : cf);
    ce = ce < 0 ? 0 : (ce > last ? last : ce);
    int64_t c0 = row + o, c1 = row + 1 + o;
    c0 = c0 < 0 ? 0 : (c0 > last ? last : c0);
    c1 = c1 < 0 ? 0 : (c1 > last ? last : c1);
    dbl2a8 m[KB];
    double first[KB], after[KB];
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      if (j < kb) {
        const double* xj = Xj + static_cast<int64_t>(j) * ldx;
        if (inside) {
          m[j] = *reinterpret_cast<const dbl2a8*>(xj + (row + o));
        } else {
          m[j].x = xj[c0];
          m[j].y = xj[c1];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      if (j < kb) {
        const double* xj = Xj + static_cast<int64_t>(j) * ldx;
        first[j] = xj[cf];  // wave-uniform: scalar loads
        after[j] = xj[ce];
      }
    }
    // rows past the end compute sums that are never stored
    bool oka[3], okb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t ca = row32 + static_cast<uint32_t>(off[k]), cb = ca + 1u;
      oka[k] = (o0 + k < cnt) & (v[k].x != 0.0) & (ca < n32);
      okb[k] = (o0 + k < cnt) & (v[k].y != 0.0) & (cb < n32);
    }
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      if (j >= kb) break;
      const double xa[3] = {wave_shr1(m[j].y, first[j]), m[j].x, m[j].y};
      const double xb[3] = {m[j].x, m[j].y, wave_shl1(m[j].x, after[j])};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sa[j] = masked_fma(oka[k], v[k].x, xa[k], sa[j]);
        sb[j] = masked_fma(okb[k], v[k].y, xb[k], sb[j]);
      }
    }
  }
}

// FP64 multi-RHS, block width a multiple of 64, runs-of-three hints: 128-row
// waves with two adjacent rows per lane, as k_hbdia_row2.  A wave whose two
// halves share one offset list sums with multi_row2_fast (the skips folded
// into the shared values, no per-term test per column) and recomputes with
// every skip applied only when a stored sum is zero or not finite (NT: one
// column pass, the values read non-temporally).  Other
// waves run the 64-row code per half.  Each column's values, skips and FMA
// order per row are the single-vector kernel's, so every column equals a
// single-vector SpMV bit for bit.
template <int KB, bool HAS_COO, bool X32, bool NT>
__global__ __launch_bounds__(kWG) void k_hbdia_multi_row2(
    const double* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk,
    const double* __restrict__ X, int64_t ldx, double* __restrict__ Y, int64_t ldy, int nrhs, int64_t nrows,
    int nb, int bw, const int* __restrict__ cr, const int* __restrict__ cc, const double* __restrict__ cv,
    const int* __restrict__ gptr, const uint32_t* __restrict__ runs3_blocks, int kshift) {
  __shared__ double s_v[kWavesPerWG][kWave];
  __shared__ double s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][kWave];
  __shared__ int s_hi[kWavesPerWG][kWave];
  __shared__ double s_y[kWavesPerWG][2 * kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned passes = static_cast<unsigned>((nrhs + KB - 1) / KB);
  const unsigned q = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const unsigned work = q / passes;
  const int64_t w2 = static_cast<int64_t>(work) * kWavesPerWG + wv;  // 128-row wave
  const int64_t row0 = w2 * 2 * kWave;
  if (row0 >= nrows) return;
  const int j0 = static_cast<int>(q - work * passes) * KB;
  const int kb = min(KB, nrhs - j0);
  const int64_t bA = row0 / bw, bB = (row0 + kWave) / bw;
  const bool hasB = row0 + kWave < nrows;
  int sA = 0, cA = 0, sB = 0, cB = 0;
  if (bA < nb) {
    sA = blk[bA];
    cA = blk[bA + 1] - sA;
  }
  if (hasB && bB < nb) {
    sB = blk[bB];
    cB = blk[bB + 1] - sB;
  }
  const bool runsA = bA < nb && ((runs3_blocks[bA >> 5] >> (bA & 31)) & 1u);
  bool same = false;
  if (runsA && cA > 0 && cA <= kWave) {
    if (bA == bB) {
      same = true;
    } else if (hasB && cB == cA) {
      const int oa = lane < cA ? offs[sA + lane] : 0;
      const int ob = lane < cA ? offs[sB + lane] : 0;
      same = __ballot(oa != ob) == 0;
    }
  }
  if (!same) {
    multi_group<double, KB, HAS_COO, X32, true>(vals, offs, blk, X, ldx, Y, ldy, nrows, nb, bw, cr, cc, cv, gptr,
                                                runs3_blocks, 2 * w2, j0, kb, lane, s_v[wv], s_x[wv], s_lo[wv],
                                                s_hi[wv]);
    if (hasB)
      multi_group<double, KB, HAS_COO, X32, true>(vals, offs, blk, X, ldx, Y, ldy, nrows, nb, bw, cr, cc, cv, gptr,
                                                  runs3_blocks, 2 * w2 + 1, j0, kb, lane, s_v[wv], s_x[wv], s_lo[wv],
                                                  s_hi[wv]);
    return;
  }
  const double* Xj = X + static_cast<int64_t>(j0) * ldx;
  double* Yj = Y + static_cast<int64_t>(j0) * ldy;
  const int64_t row = row0 + 2 * lane;
  const bool act_a = row < nrows, act_b = row + 1 < nrows;
  const double* vb = vals + static_cast<int64_t>(sA) * bw + (row0 - bA * bw);
  // lanes of half B read block bB's lines, which follow block bA's
  const int64_t vlane = 2 * lane + ((lane >= kWave / 2 && bB != bA) ? static_cast<int64_t>(cA - 1) * bw : 0);
  const int* op = offs + sA;
  double sa[KB], sb[KB];
  multi_row2_fast<KB, X32, NT>(vb, vlane, op, cA, bw, Xj, ldx, kb, row0, row, nrows, lane, sa, sb);
  // a finite non-zero sum is the exact sum bit for bit; a wave with another
  // (a zero, an infinity or NaN) in a row it stores recomputes every skip
  bool redo = false;
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    if (j < kb) {
      redo = redo | (act_a & !__builtin_amdgcn_class(sa[j], kFiniteNonzero));
      redo = redo | (act_b & !__builtin_amdgcn_class(sb[j], kFiniteNonzero));
    }
  }
  if (__ballot(redo) != 0) multi_row2_exact<KB>(vb, vlane, op, cA, bw, Xj, ldx, kb, row0, row, nrows, sa, sb);
  if constexpr (!HAS_COO) {
    // both rows of a lane are adjacent in every column: one 16-byte store
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      if (j >= kb) break;
      double* yp = Yj + static_cast<int64_t>(j) * ldy + row;
      if (act_b) {
        dbl2a8 t;
        t.x = sa[j];
        t.y = sb[j];
        __builtin_nontemporal_store(t, reinterpret_cast<dbl2a8*>(yp));
      } else if (act_a) {
        __builtin_nontemporal_store(sa[j], yp);
      }
    }
  } else {
    // back to the halves' row-per-lane order, then each half's fallback sums
    double yA[KB], yB[KB];
    double* sy = s_y[wv];
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      if (j >= kb) break;
      sy[2 * lane] = sa[j];
      sy[2 * lane + 1] = sb[j];
      wave_sync();
      yA[j] = sy[lane];
      yB[j] = sy[kWave + lane];
      wave_sync();
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && !hasB) break;
      const int64_t g = 2 * w2 + h;
      const int64_t r0 = row0 + h * kWave;
      const int p0 = gptr[g], p1 = gptr[g + 1];
      double cs[KB];
      if (p0 < p1) {
        coo_group_sum_multi<double, KB>(cr, cc, cv, Xj, ldx, kb, p0, p1, r0, lane, s_v[wv], s_x[wv], s_lo[wv],
                                        s_hi[wv], cs);
      } else {
#pragma unroll
        for (int j = 0; j < KB; ++j) cs[j] = 0.0;
      }
      if (r0 + lane < nrows) {
#pragma unroll
        for (int j = 0; j < KB; ++j) {
          if (j >= kb) break;
          double* yp = Yj + static_cast<int64_t>(j) * ldy + r0 + lane;
          __builtin_nontemporal_store((h ? yB[j] : yA[j]) + cs[j], yp);
        }
      }
    }
  }
}

// The columns of k_hbdia_multi_res for a wave whose 27 offsets (nine runs of
// three, centres cen) are held as values v (columns outside the matrix as
// zero values).  INSIDE: every run's window, lane 0's left and lane 63's right
// neighbour included, lies inside x, so each centre pair is one 16-byte load;
// otherwise each element is a clamped load.  The loads of a column are issued
// with no branch between them, and the next column's go out before this
// column's store, which would otherwise hold them back on the in-order memory
// counter.
template <bool X32, bool INSIDE>
__device__ __forceinline__ void multi_res_columns(const dbl2a8* v, const int* cen, const double* __restrict__ X,
                                                  int64_t ldx, double* __restrict__ Y, int64_t ldy, int nrhs,
                                                  int64_t row0, int64_t row, int64_t nrows, int lane) {
  const int64_t last = nrows - 1;
  const bool act_a = row < nrows, act_b = row + 1 < nrows;
  const bool lane0 = lane == 0, lane63 = lane == kWave - 1;
  dbl2a8 m[kChunk];       // per run: x at its centre for both rows
  double el[kChunk], er[kChunk];  // per run: x left of lane 0's rows, right of lane 63's
  auto load_x = [&](int j) {
    const double* xj = X + static_cast<int64_t>(j) * ldx;
    // an opaque copy of the row: the x addresses are formed per column rather
    // than hoisted out of the loop and held live beside the values
    int64_t rj = row;
    asm volatile("" : "+v"(rj));
#pragma unroll
    for (int w = 0; w < kChunk; ++w) {
      const int64_t o = cen[w];
      if constexpr (INSIDE) {
        m[w] = *reinterpret_cast<const dbl2a8*>(x_at<X32>(xj, rj + o));
      } else {
        int64_t c0 = rj + o, c1 = rj + 1 + o;
        c0 = c0 < 0 ? 0 : (c0 > last ? last : c0);
        c1 = c1 < 0 ? 0 : (c1 > last ? last : c1);
        m[w].x = *x_at<X32>(xj, c0);
        m[w].y = *x_at<X32>(xj, c1);
      }
    }
#pragma unroll
    for (int w = 0; w < kChunk; ++w) {
      const int64_t o = cen[w];
      int64_t l = row0 + o - 1, h = row0 + o + 2 * kWave;  // wave-uniform: scalar loads
      if constexpr (!INSIDE) {
        l = l < 0 ? 0 : (l > last ? last : l);
        h = h < 0 ? 0 : (h > last ? last : h);
      }
      el[w] = *x_at<X32>(xj, l);
      er[w] = *x_at<X32>(xj, h);
    }
  };
  load_x(0);
#pragma unroll 1
  for (int j = 0; j < nrhs; ++j) {
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int w = 0; w < kChunk; ++w) {
      double left = wave_shr1_zero(m[w].y), right = wave_shl1_zero(m[w].x);
      left = lane0 ? el[w] : left;     // x[row + o - 1]
      right = lane63 ? er[w] : right;  // x[row + o + 2]
      sa = __builtin_fma(v[3 * w].x, left, sa);
      sa = __builtin_fma(v[3 * w + 1].x, m[w].x, sa);
      sa = __builtin_fma(v[3 * w + 2].x, m[w].y, sa);
      sb = __builtin_fma(v[3 * w].y, m[w].x, sb);
      sb = __builtin_fma(v[3 * w + 1].y, m[w].y, sb);
      sb = __builtin_fma(v[3 * w + 2].y, right, sb);
    }
    // a finite non-zero sum is the exact sum bit for bit; a column with
    // another (a zero, an infinity or NaN) in a row it stores is summed again
    // from the same registers with every skip applied per term (a zero value,
    // which includes every column outside the matrix)
    const bool redo = (act_a & !__builtin_amdgcn_class(sa, kFiniteNonzero)) |
                      (act_b & !__builtin_amdgcn_class(sb, kFiniteNonzero));
    if (__ballot(redo) != 0) {
      sa = sb = 0.0;
#pragma unroll
      for (int w = 0; w < kChunk; ++w) {
        double left = wave_shr1_zero(m[w].y), right = wave_shl1_zero(m[w].x);
        left = lane0 ? el[w] : left;
        right = lane63 ? er[w] : right;
        const double xa[3] = {left, m[w].x, m[w].y}, xb[3] = {m[w].x, m[w].y, right};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          // opaque copies: the skip masks are formed here, per column,
          // instead of being hoisted out of the column loop and held live
          double a = v[3 * w + k].x, b = v[3 * w + k].y;
          asm volatile("" : "+v"(a), "+v"(b));
          sa = masked_fma(a != 0.0, a, xa[k], sa);
          sb = masked_fma(b != 0.0, b, xb[k], sb);
        }
      }
    }
    if (j + 1 < nrhs) load_x(j + 1);
    double* yp = Y + static_cast<int64_t>(j) * ldy + row;
    if (act_b) {
      dbl2a8 t;
      t.x = sa;
      t.y = sb;
      __builtin_nontemporal_store(t, reinterpret_cast<dbl2a8*>(yp));
    } else if (act_a) {
      __builtin_nontemporal_store(sa, yp);
    }
  }
}

// FP64 multi-RHS with more columns than one pass of k_hbdia_multi_row2 holds,
// no fallback, runs-of-three hints: 128-row waves, two adjacent rows per lane,
// every column in one wave.  A wave whose halves share one offset list of 27
// offsets (the interior of a 27-point stencil) loads its rows' values once
// (non-temporal, 108 VGPRs) and keeps them in registers while it walks the
// columns (multi_res_columns), so the value stream is read once per product
// instead of once per column pass.  Per column and run: the centre pair is one
// 16-byte load, the outer neighbours come by zero-fill DPP shifts, and lane 0's
// left / lane 63's right neighbour is selected from a scalar load of the value
// beyond the wave's end.  The exactness argument is multi_row2_fast's: a
// skipped term adds a signed zero, so a stored sum that is finite and non-zero
// is the exact sum bit for bit, and a column whose sums are not all so is
// summed again with every skip applied.  Other waves run the 64-row code per
// half, eight columns at a time.
template <bool X32>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(3))) void k_hbdia_multi_res(
    const double* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk,
    const double* __restrict__ X, int64_t ldx, double* __restrict__ Y, int64_t ldy, int nrhs, int64_t nrows,
    int nb, int bw, const uint32_t* __restrict__ runs3_blocks, int kshift) {
  constexpr int kOff = 3 * kChunk;  // 27
  __shared__ double s_v[kWavesPerWG][kWave];
  __shared__ double s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][kWave];
  __shared__ int s_hi[kWavesPerWG][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned work = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const int64_t w2 = static_cast<int64_t>(work) * kWavesPerWG + wv;  // 128-row wave
  const int64_t row0 = w2 * 2 * kWave;
  if (row0 >= nrows) return;
  const int64_t bA = row0 / bw, bB = (row0 + kWave) / bw;
  const bool hasB = row0 + kWave < nrows;
  int sA = 0, cA = 0, sB = 0, cB = 0;
  if (bA < nb) {
    sA = blk[bA];
    cA = blk[bA + 1] - sA;
  }
  if (hasB && bB < nb) {
    sB = blk[bB];
    cB = blk[bB + 1] - sB;
  }
  const bool runsA = bA < nb && ((runs3_blocks[bA >> 5] >> (bA & 31)) & 1u);
  bool same = false;
  if (runsA && cA == kOff) {
    if (bA == bB) {
      same = true;
    } else if (hasB && cB == cA) {
      const int oa = lane < cA ? offs[sA + lane] : 0;
      const int ob = lane < cA ? offs[sB + lane] : 0;
      same = __ballot(oa != ob) == 0;
    }
  }
  if (!same) {
    constexpr int KB = 8;
    for (int j0 = 0; j0 < nrhs; j0 += KB) {
      const int kb = min(KB, nrhs - j0);
      multi_group<double, KB, false, X32, true>(vals, offs, blk, X, ldx, Y, ldy, nrows, nb, bw, nullptr, nullptr,
                                                nullptr, nullptr, runs3_blocks, 2 * w2, j0, kb, lane, s_v[wv],
                                                s_x[wv], s_lo[wv], s_hi[wv]);
      if (hasB)
        multi_group<double, KB, false, X32, true>(vals, offs, blk, X, ldx, Y, ldy, nrows, nb, bw, nullptr, nullptr,
                                                  nullptr, nullptr, runs3_blocks, 2 * w2 + 1, j0, kb, lane, s_v[wv],
                                                  s_x[wv], s_lo[wv], s_hi[wv]);
    }
    return;
  }
  const int64_t last = nrows - 1;
  const int64_t row = row0 + 2 * lane;
  const uint32_t row32 = static_cast<uint32_t>(row), n32 = static_cast<uint32_t>(nrows);
  const double* vb = vals + static_cast<int64_t>(sA) * bw + (row0 - bA * bw);
  // lanes of half B read block bB's lines, which follow block bA's
  const int64_t vlane = 2 * lane + ((lane >= kWave / 2 && bB != bA) ? static_cast<int64_t>(cA - 1) * bw : 0);
  const int* op = offs + sA;
  // the wave's values, once; columns outside the matrix as zero values
  dbl2a8 v[kOff];
#pragma unroll
  for (int k = 0; k < kOff; ++k) {
    const dbl2a8 t = __builtin_nontemporal_load(
        reinterpret_cast<const dbl2a8*>(vb + static_cast<int64_t>(k) * bw + vlane));
    const uint32_t ca = row32 + static_cast<uint32_t>(op[k]);
    v[k].x = ca < n32 ? t.x : 0.0;
    v[k].y = ca + 1u < n32 ? t.y : 0.0;
  }
  // the runs' centres, and whether every run's window lies inside x
  int cen[kChunk];
  bool inside = true;
#pragma unroll
  for (int w = 0; w < kChunk; ++w) {
    const int64_t o = op[3 * w + 1];
    cen[w] = static_cast<int>(o);
    inside = inside && row0 + o - 1 >= 0 && row0 + o + 2 * kWave <= last;  // wave-uniform
  }
  if (inside)
    multi_res_columns<X32, true>(v, cen, X, ldx, Y, ldy, nrhs, row0, row, nrows, lane);
  else
    multi_res_columns<X32, false>(v, cen, X, ldx, Y, ldy, nrhs, row0, row, nrows, lane);
}

// ---------------------------------------------------------------------------
// Many matrices in one persistent launch.  A tile is 64 consecutive rows of one
// matrix; it is processed with exactly the arithmetic of k_hbdia<T, true, ...>
// on that matrix alone (blocked part: ascending-offset FMA chain with the zero
// and column-range skips; fallback fused when the matrix has one), so every
// matrix's rows equal the single-matrix results bit for bit.
//
// A tile is a short chain of dependent loads (record -> offsets -> values and
// x -> fallback x gather), so a wave overlaps tiles: the next tile's record and
// first offsets are loaded while the current tile's vector loads are in
// flight, and the current tile's fallback entries are loaded before its
// blocked part is summed.

// Two tiles of one matrix, rows row0 .. row0 + 127, whose blocks share the
// offset list off[0 .. cnt), cnt <= CH <= kChunk.  The kernel instantiates CH
// per count, so a pair issues cnt value and x loads rather than kChunk (the
// 2048-matrix batch: 0.642 -> 0.603 ms).  Lane l holds rows row0 + 2 l and
// row0 + 2 l + 1, its two values of an offset one 16-byte load from tile A's
// lines (l < 32) or tile B's.  With a fallback (fused) its entries [p0, p2)
// (tile A's [p0, p1)) number at most 64 and are summed in the two-row layout.
// Per row the arithmetic is the single tile's: ascending-offset FMA chain with
// the zero and column-range skips, then y = (blocked sum) + fallback.
template <typename T, bool HAS_COO, bool X32, int CH = kChunk>
__device__ __forceinline__ void tile_pair(int value_line_a, int value_line_b, int rows, int row0, int num_rows, int bw,
                                          int cnt, const int* off, int p0, int p1, int p2, int fused,
                                          const T* __restrict__ vals, const int* __restrict__ cr,
                                          const int* __restrict__ cc, const T* __restrict__ cv,
                                          const T* __restrict__ xp, T* __restrict__ yp, int lane, T* sv, T* sx,
                                          int* slo, int* shi) {
  using V2 = typename Vec2<T>::type;
  const T* x = xp + rows;
  T* y = yp + rows;
  const int64_t nrows = num_rows;
  const int64_t row = static_cast<int64_t>(row0) + 2 * lane;
  const bool act_a = row < nrows, act_b = row + 1 < nrows;
  const int64_t vline = lane < kWave / 2 ? static_cast<int64_t>(value_line_a) * kWave + 2 * lane
                                         : static_cast<int64_t>(value_line_b) * kWave + 2 * (lane - kWave / 2);
  const T* vb = vals + vline;
  const int lastk = cnt - 1;
  V2 v[CH], xv[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k)
    v[k] = __builtin_nontemporal_load(reinterpret_cast<const V2*>(vb + static_cast<int64_t>(min(k, lastk)) * bw));
  const int64_t last = nrows - 1;
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const int64_t o = off[k];
    if (row0 + o >= 0 && row0 + o + 2 * kWave <= last + 1) {  // wave-uniform: every column inside x
      xv[k] = *reinterpret_cast<const V2*>(x + (row + o));
    } else {
      int64_t c0 = row + o, c1 = row + 1 + o;
      c0 = c0 < 0 ? 0 : (c0 > last ? last : c0);
      c1 = c1 < 0 ? 0 : (c1 > last ? last : c1);
      xv[k].x = x[c0];
      xv[k].y = x[c1];
    }
  }
  T ca = T(0), cb = T(0);
  if constexpr (HAS_COO) {
    if (fused && p0 < p2)
      coo_pair_sum<T, true>(cr, cc, cv, x, p0, p1, p2, row0, lane, sv, sx, slo, shi, ca, cb);
  }
  const uint32_t n32 = static_cast<uint32_t>(num_rows);
  T sa = T(0), sb = T(0);
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const uint32_t c = static_cast<uint32_t>(row) + static_cast<uint32_t>(off[k]);
    const bool oka = (k < cnt) & act_a & (v[k].x != T(0)) & (c < n32);
    const bool okb = (k < cnt) & act_b & (v[k].y != T(0)) & (c + 1u < n32);
    sa = masked_fma(oka, v[k].x, xv[k].x, sa);
    sb = masked_fma(okb, v[k].y, xv[k].y, sb);
  }
  if (HAS_COO && fused) {
    sa = sa + ca;
    sb = sb + cb;
  }
  if (act_b) {
    V2 t;
    t.x = sa;
    t.y = sb;
    __builtin_nontemporal_store(t, reinterpret_cast<V2*>(y + row));
  } else if (act_a) {
    __builtin_nontemporal_store(sa, y + row);
  }
}

// tile_pair for two tiles whose offset list is longer than one chunk (the
// tile records name the same list: hbdia_batch_make_tiles gives equal lists of
// consecutive tiles one offsets index).  The list is walked in chunks of
// kChunk in ascending order, the FMA chain running on across chunks as in the
// single-tile path; the next chunk's offsets (wave-uniform loads) are loaded
// before the current chunk is summed.  A full chunk of three runs of
// consecutive offsets inside x loads each run's centre pair once and shifts
// the outer neighbours in across lanes.
template <typename T, bool HAS_COO, bool X32>
__device__ __forceinline__ void tile_pair_long(int value_line_a, int value_line_b, int row0, int num_rows, int bw,
                                               int cnt, const int* first_off, const int* __restrict__ op, int p0,
                                               int p1, int p2, int fused, const T* __restrict__ vals,
                                               const int* __restrict__ cr, const int* __restrict__ cc,
                                               const T* __restrict__ cv, const T* __restrict__ x, T* __restrict__ y,
                                               int lane, T* sv, T* sx, int* slo, int* shi) {
  using V2 = typename Vec2<T>::type;
  const int64_t nrows = num_rows;
  const int64_t row = static_cast<int64_t>(row0) + 2 * lane;
  const bool act_a = row < nrows, act_b = row + 1 < nrows;
  const int64_t vline = lane < kWave / 2 ? static_cast<int64_t>(value_line_a) * kWave + 2 * lane
                                         : static_cast<int64_t>(value_line_b) * kWave + 2 * (lane - kWave / 2);
  const T* vb = vals + vline;
  const int64_t last = nrows - 1;
  const uint32_t n32 = static_cast<uint32_t>(num_rows);
  T ca = T(0), cb = T(0);
  if constexpr (HAS_COO) {
    if (fused && p0 < p2)
      coo_pair_sum<T, true>(cr, cc, cv, x, p0, p1, p2, row0, lane, sv, sx, slo, shi, ca, cb);
  }
  int off[kChunk];
#pragma unroll
  for (int k = 0; k < kChunk; ++k) off[k] = first_off[k];
  T sa = T(0), sb = T(0);
  for (int o0 = 0; o0 < cnt; o0 += kChunk) {
    const int lastk = min(kChunk, cnt - o0) - 1;
    V2 v[kChunk], xv[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      v[k] = __builtin_nontemporal_load(
          reinterpret_cast<const V2*>(vb + static_cast<int64_t>(o0 + min(k, lastk)) * bw));
    // a full chunk of three runs of consecutive offsets whose windows lie
    // inside x: one 16-byte load per run, the outer neighbours by lane shifts
    // (the same x values as the per-offset loads)
    bool runs = lastk == kChunk - 1;
#pragma unroll
    for (int w = 0; w < kChunk / 3; ++w) {
      const int64_t o = off[3 * w + 1];
      runs = runs && off[3 * w] == o - 1 && off[3 * w + 2] == o + 1 && row0 + o - 1 >= 0 &&
             row0 + o + 2 * kWave <= last;
    }
    if (runs) {
#pragma unroll
      for (int w = 0; w < kChunk / 3; ++w) {
        const int64_t o = off[3 * w + 1];
        const V2 m = *reinterpret_cast<const V2*>(x + (row + o));
        const T el = x[row0 + o - 1], er = x[row0 + o + 2 * kWave];  // wave-uniform: scalar loads
        xv[3 * w].x = wave_shr1(m.y, el);  // x[row + o - 1]
        xv[3 * w].y = m.x;
        xv[3 * w + 1] = m;
        xv[3 * w + 2].x = m.y;
        xv[3 * w + 2].y = wave_shl1(m.x, er);  // x[row + o + 2]
      }
    } else {
#pragma unroll
      for (int k = 0; k < kChunk; ++k) {
        const int64_t o = off[k];
        if (row0 + o >= 0 && row0 + o + 2 * kWave <= last + 1) {  // wave-uniform: every column inside x
          xv[k] = *reinterpret_cast<const V2*>(x + (row + o));
        } else {
          int64_t c0 = row + o, c1 = row + 1 + o;
          c0 = c0 < 0 ? 0 : (c0 > last ? last : c0);
          c1 = c1 < 0 ? 0 : (c1 > last ? last : c1);
          xv[k].x = x[c0];
          xv[k].y = x[c1];
        }
      }
    }
    int cur[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) cur[k] = off[k];
    if (o0 + kChunk < cnt) {  // the next chunk's offsets (the pool is padded past every list)
#pragma unroll
      for (int k = 0; k < kChunk; ++k) off[k] = op[o0 + kChunk + k];
    }
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const uint32_t c = static_cast<uint32_t>(row) + static_cast<uint32_t>(cur[k]);
      const bool ink = k <= lastk;
      const bool oka = ink & act_a & (v[k].x != T(0)) & (c < n32);
      const bool okb = ink & act_b & (v[k].y != T(0)) & (c + 1u < n32);
      sa = masked_fma(oka, v[k].x, xv[k].x, sa);
      sb = masked_fma(okb, v[k].y, xv[k].y, sb);
    }
  }
  if (HAS_COO && fused) {
    sa = sa + ca;
    sb = sb + cb;
  }
  if (act_b) {
    V2 t;
    t.x = sa;
    t.y = sb;
    __builtin_nontemporal_store(t, reinterpret_cast<V2*>(y + row));
  } else if (act_a) {
    __builtin_nontemporal_store(sa, y + row);
  }
}

constexpr int kTileGroup = 16;  // tile records fetched together by one wave
constexpr int kJobsInLds = 1024;  // serving: job starts held in LDS (8 KB)
constexpr int kTileDwords = 20;  // 80-byte tile record
static_assert(sizeof(hbdia_batch_tile_t) == kTileDwords * 4, "tile record layout");
static_assert(HBDIA_BATCH_INLINE_OFFSETS >= kChunk, "a tile record carries the first chunk's offsets");

// JOBS: the work is a list of jobs (serving), each the tiles of one pooled
// matrix applied to the job's own x and y (device addresses carried by the job,
// element 0 at the matrix's row 0); the wave's work indices are mapped to (job,
// tile) by a binary search over the jobs' work_begin.  A job's y is taken to
// start at zero: every row is written, rows without blocks as 0 (+ fallback),
// so the caller clears nothing.
template <typename T, bool HAS_COO, bool X32, bool CONTIG, bool JOBS, bool LONG>
__global__ __launch_bounds__(kWG) void k_hbdia_batch(const int32_t* __restrict__ tiles,
                                                     const int32_t* __restrict__ wave_tiles,
                                                     const T* __restrict__ vals, const int* __restrict__ offs,
                                                     const int* __restrict__ cr, const int* __restrict__ cc,
                                                     const T* __restrict__ cv, int64_t num_tiles, int tiles_per_step,
                                                     const hbdia_serve_job_t* __restrict__ jobs, int num_jobs,
                                                     const T* __restrict__ xp, T* __restrict__ yp) {
  static_assert(!(JOBS && CONTIG), "jobs run on the interleaved schedule");
  static_assert(kChunk == HBDIA_BATCH_CHUNK, "chunk of the batch kernel");
  __shared__ T s_v[kWavesPerWG][kWave];
  __shared__ T s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][2 * kWave];  // 128 rows for a pair of tiles
  __shared__ int s_hi[kWavesPerWG][2 * kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int64_t wave = static_cast<int64_t>(blockIdx.x) * kWavesPerWG + wv;
  int64_t step = static_cast<int64_t>(gridDim.x) * kWavesPerWG * tiles_per_step;
  int64_t first = 0, limit = num_tiles;
  if constexpr (CONTIG) {
    first = wave_tiles[wave];
    limit = wave_tiles[wave + 1];
  } else if (!JOBS && wave_tiles != nullptr) {
    // XCD split: the workgroups of XCD k (blockIdx % 8 == k) interleave over
    // the tiles [wave_tiles[k], wave_tiles[k + 1])
    const unsigned xcd = blockIdx.x & 7u;
    wave = static_cast<int64_t>(blockIdx.x >> 3) * kWavesPerWG + wv;
    step = static_cast<int64_t>(gridDim.x >> 3) * kWavesPerWG * tiles_per_step;
    first = wave_tiles[xcd];
    limit = wave_tiles[xcd + 1];
  }
  // JOBS: the jobs' first work indices in LDS, so the search that maps each
  // tile group to its job is a few LDS reads instead of a chain of dependent
  // global loads (up to kJobsInLds jobs; more are searched in memory)
  __shared__ int64_t s_job[JOBS ? kJobsInLds : 1];
  const bool jobs_lds = JOBS && num_jobs <= kJobsInLds;  // uniform
  if constexpr (JOBS) {
    if (jobs_lds) {
      for (int i = threadIdx.x; i < num_jobs; i += blockDim.x) s_job[i] = jobs[i].work_begin;
      __syncthreads();
    }
  }
  T* sv = s_v[wv];
  T* sx = s_x[wv];
  int* slo = s_lo[wv];
  int* shi = s_hi[wv];
  // The wave's tiles form a sequence s = 0, 1, ...: contiguous, tile first + s;
  // interleaved, runs of tiles_per_step tiles `step` apart.  Valid tiles are a
  // prefix of the sequence.  Lane 4j + q holds dwords [5q, 5q + 5) of the
  // record of sequence entry s0 + j.
  const int jl = lane >> 2, ql = lane & 3;
  for (int64_t s0 = 0;; s0 += kTileGroup) {
    const int64_t sj = s0 + jl;
    int64_t tj;
    if constexpr (CONTIG) tj = first + sj;
    else tj = first + wave * tiles_per_step + (sj / tiles_per_step) * step + sj % tiles_per_step;
    const bool vj = tj < limit;
    int64_t rec = vj ? tj : 0;
    uint64_t jx = 0, jy = 0;  // JOBS: the job's x and y addresses
    if constexpr (JOBS) {
      // the last job whose work starts at or before tj
      int lo = 0, hi = num_jobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((jobs_lds ? s_job[mid] : jobs[mid].work_begin) <= rec) lo = mid;
        else hi = mid - 1;
      }
      rec = vj ? jobs[lo].tile_begin + (tj - jobs[lo].work_begin) : 0;
      jx = vj ? jobs[lo].x : 0;
      jy = vj ? jobs[lo].y : 0;
    }
    const int32_t* rp = tiles + rec * kTileDwords + ql * 5;
    int rv[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) rv[r] = rp[r];
    const int ng = __popcll(__ballot(vj && ql == 0));
    if (ng == 0) break;
    // fields of tile jj of the group, read from the lanes holding its record;
    // the first tile's are read before the loop, each next one at the end of
    // an iteration, so the loop never waits on the record load
    struct Fields {
      int value_line, rows, row0, num_rows, bw, cnt, offs_at, p0, p1, fused;
      int off[kChunk];
      uint64_t xb, yb;  // JOBS: the tile's job's x and y
    };
    auto read_fields = [&](int jj) {
      Fields f;
#define TILE_FIELD(i) __builtin_amdgcn_readlane(rv[(i) % 5], 4 * jj + (i) / 5)
      f.value_line = TILE_FIELD(0);
      f.rows = TILE_FIELD(1);
      f.row0 = TILE_FIELD(2);
      f.num_rows = TILE_FIELD(3);
      f.bw = TILE_FIELD(4);
      f.cnt = TILE_FIELD(5);
      f.offs_at = TILE_FIELD(6);
      f.p0 = TILE_FIELD(7);
      f.p1 = TILE_FIELD(8);
      f.fused = TILE_FIELD(9);
      f.off[0] = TILE_FIELD(10);
      f.off[1] = TILE_FIELD(11);
      f.off[2] = TILE_FIELD(12);
      f.off[3] = TILE_FIELD(13);
      f.off[4] = TILE_FIELD(14);
      f.off[5] = TILE_FIELD(15);
      f.off[6] = TILE_FIELD(16);
      f.off[7] = TILE_FIELD(17);
      f.off[8] = TILE_FIELD(18);
#undef TILE_FIELD
      f.xb = f.yb = 0;
      if constexpr (JOBS) {
        auto lane64 = [&](uint64_t v) {
          const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(v), 4 * jj);
          const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(v >> 32), 4 * jj);
          return (static_cast<uint64_t>(hi) << 32) | lo;
        };
        f.xb = lane64(jx);
        f.yb = lane64(jy);
      }
      return f;
    };
    Fields nextf = read_fields(0);
    for (int jj = 0; jj < ng; ++jj) {
      const Fields cf = nextf;
      if (jj + 1 < ng) nextf = read_fields(jj + 1);
      // ---- two consecutive tiles of one matrix whose blocks hold the same
      // offsets (one chunk at most) and whose fallback fits one pass: 128 rows,
      // two adjacent rows per lane, 16-byte value loads; the arithmetic per row
      // is the single tile's below
      bool pair = jj + 1 < ng && nextf.rows == cf.rows && nextf.row0 == cf.row0 + kWave && nextf.bw == cf.bw &&
                  nextf.cnt == cf.cnt && cf.cnt > 0 && nextf.fused == cf.fused &&
                  (!HAS_COO || !cf.fused || (nextf.p0 == cf.p1 && nextf.p1 - cf.p0 <= kWave));
      if (cf.cnt <= kChunk) {
#pragma unroll
        for (int k = 0; k < kChunk; ++k) pair = pair && (k >= cf.cnt || nextf.off[k] == cf.off[k]);
      } else {
        pair = pair && nextf.offs_at == cf.offs_at;  // one list (equal lists share an index)
      }
      if (LONG && pair && cf.cnt > kChunk) {
        const Fields cg = nextf;
        ++jj;
        if (jj + 1 < ng) nextf = read_fields(jj + 1);
        const T* xq = JOBS ? reinterpret_cast<const T*>(cf.xb) : xp + cf.rows;
        T* yq = JOBS ? reinterpret_cast<T*>(cf.yb) : yp + cf.rows;
        tile_pair_long<T, HAS_COO, X32>(cf.value_line, cg.value_line, cf.row0, cf.num_rows, cf.bw, cf.cnt, cf.off,
                                        offs + cf.offs_at, cf.p0, cf.p1, cg.p1, cf.fused, vals, cr, cc, cv, xq, yq,
                                        lane, sv, sx, slo, shi);
        continue;
      }
      if (pair && cf.cnt <= kChunk) {
        const Fields cg = nextf;
        ++jj;
        if (jj + 1 < ng) nextf = read_fields(jj + 1);
        // a job's x and y start at its matrix's row 0; the batch's at the pool's
        const T* xq = JOBS ? reinterpret_cast<const T*>(cf.xb) : xp + cf.rows;
        T* yq = JOBS ? reinterpret_cast<T*>(cf.yb) : yp + cf.rows;
        // the pair's offset count bounds its loads and terms
#define HBDIA_PAIR(CH)                                                                                         \
  tile_pair<T, HAS_COO, X32, CH>(cf.value_line, cg.value_line, 0, cf.row0, cf.num_rows, cf.bw, cf.cnt, cf.off, \
                                 cf.p0, cf.p1, cg.p1, cf.fused, vals, cr, cc, cv, xq, yq, lane, sv, sx, slo, shi)
        switch (cf.cnt) {
          case 1: case 2: case 3: HBDIA_PAIR(3); break;
          case 4: HBDIA_PAIR(4); break;
          case 5: HBDIA_PAIR(5); break;
          case 6: HBDIA_PAIR(6); break;
          case 7: HBDIA_PAIR(7); break;
          case 8: HBDIA_PAIR(8); break;
          default: HBDIA_PAIR(kChunk); break;
        }
#undef HBDIA_PAIR
        continue;
      }
      const int value_line = cf.value_line, rows = cf.rows, row0 = cf.row0, num_rows = cf.num_rows, bw = cf.bw,
                cnt = cf.cnt, offs_at = cf.offs_at, p0 = cf.p0, p1 = cf.p1, fused = cf.fused;
      const int* off = cf.off;
      const int64_t nrows = num_rows;
      const int64_t row = static_cast<int64_t>(row0) + lane;  // row within the matrix
      const bool active = row < nrows;
      const T* x = JOBS ? reinterpret_cast<const T*>(cf.xb) : xp + rows;
      T* y = JOBS ? reinterpret_cast<T*>(cf.yb) : yp + rows;
      // A matrix's rows are below 2^31: row + offset modulo 2^32 is below n32
      // exactly when the column lies in [0, nrows), and an out-of-range column
      // lies below 0 for a negative offset and past the last row otherwise.
      // Rows past the end compute unused sums (they are not stored).
      const uint32_t row32 = static_cast<uint32_t>(row), n32 = static_cast<uint32_t>(num_rows);
      const uint32_t last32 = n32 - 1u;
      auto col_of = [&](int o, bool& in) -> uint32_t {
        const uint32_t c = row32 + static_cast<uint32_t>(o);
        in = c < n32;
        return in ? c : (o < 0 ? 0u : last32);
      };

      // ---- issue the tile's loads: fallback entries, first chunk of values and x
      int r_ = 0, c_ = 0;
      T ev = T(0);
      if constexpr (HAS_COO) {
        const int i = p0 + lane;  // inside the padded pool: loads need no guard
        r_ = cr[i];
        c_ = cc[i];
        ev = cv[i];
      }
      const T* vb = vals + static_cast<int64_t>(value_line) * kWave;
      const int lastk = max(cnt - 1, 0);
      T v[kChunk], xl[kChunk];
#pragma unroll
      for (int k = 0; k < kChunk; ++k)
        v[k] = __builtin_nontemporal_load(at_u32(vb, static_cast<uint32_t>(min(k, lastk) * bw + lane)));
      bool in1[kChunk];
#pragma unroll
      for (int k = 0; k < kChunk; ++k) xl[k] = load_x<T, X32>(x, col_of(off[k], in1[k]));
      // every load above is issued before anything waits for one of them
      __builtin_amdgcn_sched_barrier(0);
      T ex = T(0);
      int er = -1;
      if constexpr (HAS_COO) {
        const int okm = -static_cast<int>(p0 + lane < p1);  // all ones or zero
        ex = load_x<T, X32>(x, c_ & okm);
        er = ((r_ - row0) & okm) | ~okm;  // -1 past the tile's entries
        if (er < 0 || er >= kWave) er = -1;  // unsorted input: drop rather than write outside LDS
      }

      // ---- blocked part: the FMA chain over the first chunk, then any further chunks
      T sum = T(0);
#pragma unroll
      for (int k = 0; k < kChunk; ++k) sum = masked_fma((k < cnt) & (v[k] != T(0)) & in1[k], v[k], xl[k], sum);
      if (cnt > kChunk) {
        const int* op = offs + offs_at;
        for (int o0 = kChunk; o0 < cnt; o0 += kChunk) {
          int of2[kChunk];
          T v2[kChunk], x2[kChunk];
#pragma unroll
          for (int k = 0; k < kChunk; ++k) of2[k] = op[o0 + k];  // padded pool
#pragma unroll
          for (int k = 0; k < kChunk; ++k)
            v2[k] = __builtin_nontemporal_load(at_u32(vb, static_cast<uint32_t>(min(o0 + k, cnt - 1) * bw + lane)));
          bool in2[kChunk];
#pragma unroll
          for (int k = 0; k < kChunk; ++k) x2[k] = load_x<T, X32>(x, col_of(of2[k], in2[k]));
#pragma unroll
          for (int k = 0; k < kChunk; ++k)
            sum = masked_fma((o0 + k < cnt) & (v2[k] != T(0)) & in2[k], v2[k], x2[k], sum);
        }
      }

      // ---- fallback: per-row sums in entry order, fused into y
      if constexpr (HAS_COO) {
        // stored on every path: a fallback load left unconsumed would make the
        // next tile wait for it (and for this tile's store) before reusing its register
        slo[lane] = 0;
        shi[lane] = 0;
        sv[lane] = ev;
        sx[lane] = ex;
      }
      if (HAS_COO && fused) {
        T cs = T(0);
        if (p0 < p1) {
          const int erp = __shfl_up(er, 1);
          const int ern = __shfl_down(er, 1);
          wave_sync();
          if (er >= 0 && (lane == 0 || erp != er)) slo[er] = lane;
          if (er >= 0 && (lane == kWave - 1 || p0 + lane + 1 >= p1 || ern != er)) shi[er] = lane + 1;
          wave_sync();
          const int a = slo[lane], e = shi[lane];
          for (int k = a; k < e; ++k) cs = fma_t<T>(sv[k], sx[k], cs);
          wave_sync();
          if (p1 - p0 > kWave) {  // more than 64 entries (rare): the generic loop
            bool any = false;
            cs = coo_group_sum_from<T>(cr, cc, cv, x, p0 + kWave, p1, row0, lane, sv, sx, slo, shi, cs, &any);
          }
        }
        if (active) {
          // a job's y starts at zero: rows without blocks are written too
          const T base = cnt > 0 ? sum : (JOBS ? T(0) : y[row]);
          __builtin_nontemporal_store(base + cs, y + row);
        }
      } else {
        if (active && (cnt > 0 || JOBS)) __builtin_nontemporal_store(cnt > 0 ? sum : T(0), y + row);
      }
    }
    if (ng < kTileGroup) break;
  }
}




// ---------------------------------------------------------------------------
// Reduced-precision storage.  Values (blocked part and fallback) are held as
// FP32 or BF16 copies of the FP64 matrix; x and y stay FP64.  Each row is
// accumulated in FP32 (products of the stored value and x rounded to FP32),
// together with the FP32 sum of |a x|.  A row whose cancellation ratio
// sum|a x| / |sum a x| exceeds `ratio` (or whose sum is zero or not finite) is
// appended to a device list and recomputed by k_fix_rows from the FP64 arrays
// with the FP64 kernel's exact arithmetic.

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// FP32 -> BF16, round to nearest even (NaN stays NaN).
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return static_cast<uint16_t>((u >> 16) | 0x40u);
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(r >> 16);
}

template <typename TV> __device__ __forceinline__ float lp_to_f32(TV v);
template <> __device__ __forceinline__ float lp_to_f32<float>(float v) { return v; }
template <> __device__ __forceinline__ float lp_to_f32<uint16_t>(uint16_t v) { return bf16_to_f32(v); }

__global__ void k_f64_to_f32(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = static_cast<float>(src[i]);
}

__global__ void k_f64_to_bf16(const double* __restrict__ src, uint16_t* __restrict__ dst, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = f32_to_bf16(static_cast<float>(src[i]));
}

__global__ void k_zero_i32(int* __restrict__ p) { *p = 0; }

// FP64 recomputation of one row with the FP64 kernel's arithmetic: blocked
// part in ascending offset order (FMA, zero and range skips), fallback entries
// of the row in entry order (FMA), y = (bdia or old y) + coo.  The blocked
// part takes CH offsets per round of loads: nine in the fix-up pass, three in
// kernels that recompute their own flagged rows (few registers there).
template <int CH>
__device__ __forceinline__ void fix_row(int64_t row, const double* __restrict__ vals, const int* __restrict__ offs,
                                        const int* __restrict__ blk, const double* __restrict__ x,
                                        double* __restrict__ y, int64_t nrows, int nb, int bw,
                                        const int* __restrict__ cr, const int* __restrict__ cc,
                                        const double* __restrict__ cv, const int* __restrict__ gptr) {
  const int64_t b = row / bw;
  int start = 0, cnt = 0;
  if (b < nb) {
    start = blk[b];
    cnt = blk[b + 1] - start;
  }
  double sum = 0.0;
  if (cnt > 0) {
    const double* vp = vals + static_cast<int64_t>(start) * bw + (row - b * bw);
    sum = bdia_sum_lane<double, false, CH>(vp, offs + start, cnt, bw, x, row, nrows);
  }
  if (gptr) {
    const int64_t g = row / kWave;
    int lo = gptr[g], hi = gptr[g + 1];
    const int end = hi;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cr[mid] < row) lo = mid + 1;
      else hi = mid;
    }
    double cs = 0.0;
    for (int i = lo; i < end && cr[i] == row; ++i) cs = __builtin_fma(cv[i], x[cc[i]], cs);
    y[row] = (cnt > 0 ? sum : y[row]) + cs;
  } else if (cnt > 0) {
    y[row] = sum;
  }
}

// FP64 arrays for recomputing flagged rows inside the reduced-precision
// kernels (FIX_INLINE): with them a flagged row is recomputed by its own lane
// before the kernel ends, and no fix-up pass follows.
struct MixedFp64 {
  const double* vals;
  const double* cv;
};

template <typename TV, bool HAS_COO, bool RUNS3, bool FIX_INLINE>
__global__ __launch_bounds__(kWG) void k_hbdia_mixed(
    const TV* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk,
    const double* __restrict__ x, double* __restrict__ y, int64_t nrows, int nb, int bw,
    const int* __restrict__ cr, const int* __restrict__ cc, const TV* __restrict__ cv,
    const int* __restrict__ gptr, float ratio, float u_store, uint64_t* __restrict__ flag_mask,
    const uint32_t* __restrict__ runs3_blocks, int kshift, MixedFp64 f64) {
  __shared__ float s_v[kWavesPerWG][kWave];
  __shared__ float s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][kWave];
  __shared__ int s_hi[kWavesPerWG][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned work = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const int64_t grp = static_cast<int64_t>(work) * kWavesPerWG + wv;
  const int64_t row0 = grp * kWave;
  if (row0 >= nrows) return;
  const int64_t row = row0 + lane;
  const bool active = row < nrows;
  const int64_t last = nrows - 1;
  const int64_t b = row0 / bw;
  int start = 0, cnt = 0;
  if (b < nb) {
    start = blk[b];
    cnt = blk[b + 1] - start;
  }
  float sum = 0.0f, sabs = 0.0f;
  bool runs = false;
  if constexpr (RUNS3) runs = cnt > 0 && ((runs3_blocks[b >> 5] >> (b & 31)) & 1u);
  if (cnt > 0) {
    const TV* vb = vals + static_cast<int64_t>(start) * bw + (row0 - b * bw);
    const int* op = offs + start;
    for (int o0 = 0; o0 < cnt; o0 += kChunk) {
      int off[kChunk];
      float v[kChunk];
      double xl[kChunk];
#pragma unroll
      for (int k = 0; k < kChunk; ++k) off[k] = op[o0 + k];  // padded array
#pragma unroll
      for (int k = 0; k < kChunk; ++k)
        v[k] = lp_to_f32<TV>(__builtin_nontemporal_load(vb + static_cast<int64_t>(min(o0 + k, cnt - 1)) * bw + lane));
      if (runs) {
        runs3_x<double, false, kChunk / 3>(x, off, row0, row, last, xl);
      } else {
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
          int64_t c = row + off[k];
          c = c < 0 ? 0 : (c > last ? last : c);
          xl[k] = x[c];
        }
      }
#pragma unroll
      for (int k = 0; k < kChunk; ++k) {
        const int64_t c = row + off[k];
        const bool ok = (o0 + k < cnt) & active & (v[k] != 0.0f) & (c >= 0) & (c < nrows);
        const float xf = static_cast<float>(xl[k]);
        sum = ok ? __builtin_fmaf(v[k], xf, sum) : sum;
        sabs = ok ? __builtin_fmaf(__builtin_fabsf(v[k]), __builtin_fabsf(xf), sabs) : sabs;
      }
    }
  }
  float cs = 0.0f, cabs = 0.0f;
  int nterms = cnt;
  if constexpr (HAS_COO) {
    const int p0 = gptr[grp], p1 = gptr[grp + 1];
    for (int base = p0; base < p1; base += kWave) {
      const int i = base + lane;
      const bool ok = i < p1;
      int r = -1;
      float vv = 0.0f, xx = 0.0f;
      if (ok) {
        r = cr[i] - static_cast<int>(row0);
        vv = lp_to_f32<TV>(cv[i]);
        xx = static_cast<float>(x[cc[i]]);
        if (r < 0 || r >= kWave) r = -1;
      }
      const int rp = __shfl_up(r, 1);
      const int rn = __shfl_down(r, 1);
      s_lo[wv][lane] = 0;
      s_hi[wv][lane] = 0;
      s_v[wv][lane] = vv;
      s_x[wv][lane] = xx;
      wave_sync();
      if (ok && r >= 0 && (lane == 0 || rp != r)) s_lo[wv][r] = lane;
      if (ok && r >= 0 && (lane == kWave - 1 || i + 1 >= p1 || rn != r)) s_hi[wv][r] = lane + 1;
      wave_sync();
      const int a = s_lo[wv][lane], e = s_hi[wv][lane];
      nterms += e - a;
      for (int k = a; k < e; ++k) {
        cs = __builtin_fmaf(s_v[wv][k], s_x[wv][k], cs);
        cabs = __builtin_fmaf(__builtin_fabsf(s_v[wv][k]), __builtin_fabsf(s_x[wv][k]), cabs);
      }
      wave_sync();
    }
  }
  // Cancellation test against the smallest |y| consistent with the FP32
  // result: err bounds |tot - y| by (u_store + u_x + gamma_n) sum|a x| (FP32
  // value or BF16 storage, x rounded to FP32, one rounding per FMA, plus the
  // final addition), so a row is kept only if (|tot| - err) * ratio >= sum|a x|.
  const float tot = sum + cs;
  const float mag = sabs + cabs;
  const float u32 = 5.9604645e-08f;  // 2^-24
  const float err = (u_store + u32 * static_cast<float>(nterms + 2)) * mag * 1.0001f;
  const bool fix = active && (!((__builtin_fabsf(tot) - err) * ratio >= mag) || !__builtin_isfinite(tot)) &&
                   (cnt > 0 || cs != 0.0f || cabs != 0.0f);
  // every group stores its mask (zero included): no reset between calls
  const uint64_t m = __ballot(fix);
  if (lane == 0) flag_mask[grp] = m;
  if (FIX_INLINE && fix)
    fix_row<3>(row, f64.vals, offs, blk, x, y, nrows, nb, bw, cr, cc, f64.cv, HAS_COO ? gptr : nullptr);
  if (!active || fix) return;
  if constexpr (HAS_COO) {
    const double base = cnt > 0 ? static_cast<double>(sum) : y[row];
    __builtin_nontemporal_store(base + static_cast<double>(cs), y + row);
  } else {
    if (cnt > 0) __builtin_nontemporal_store(static_cast<double>(sum), y + row);
  }
}

// FP32 sum and sum of |a x| of one row, the blocked part (rows r0 + lane of a
// 64-row half), exactly k_hbdia_mixed's arithmetic
template <typename TV>
__device__ __forceinline__ void mixed_half(const TV* __restrict__ vals, const int* __restrict__ offs,
                                           const double* __restrict__ x, int64_t r0, int64_t b, int start, int cnt,
                                           int bw, int lane, int64_t nrows, bool runs, float& sum, float& sabs) {
  sum = 0.0f;
  sabs = 0.0f;
  if (cnt <= 0) return;
  const int64_t row = r0 + lane;
  const bool active = row < nrows;
  const int64_t last = nrows - 1;
  const TV* vb = vals + static_cast<int64_t>(start) * bw + (r0 - b * bw);
  const int* op = offs + start;
  for (int o0 = 0; o0 < cnt; o0 += kChunk) {
    int off[kChunk];
    float v[kChunk];
    double xl[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) off[k] = op[o0 + k];  // padded array
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      v[k] = lp_to_f32<TV>(__builtin_nontemporal_load(vb + static_cast<int64_t>(min(o0 + k, cnt - 1)) * bw + lane));
    if (runs) {
      runs3_x<double, false, kChunk / 3>(x, off, r0, row, last, xl);
    } else {
#pragma unroll
      for (int k = 0; k < kChunk; ++k) {
        int64_t c = row + off[k];
        c = c < 0 ? 0 : (c > last ? last : c);
        xl[k] = x[c];
      }
    }
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const int64_t c = row + off[k];
      const bool ok = (o0 + k < cnt) & active & (v[k] != 0.0f) & (c >= 0) & (c < nrows);
      const float xf = static_cast<float>(xl[k]);
      sum = ok ? __builtin_fmaf(v[k], xf, sum) : sum;
      sabs = ok ? __builtin_fmaf(__builtin_fabsf(v[k]), __builtin_fabsf(xf), sabs) : sabs;
    }
  }
}

// The fallback sums of one 64-row group (rows r0 + lane) for the calling
// lane's row, in entry order: FP32 sum, FP32 sum of |a x| and the number of
// terms added to nterms; exactly k_hbdia_mixed's.
template <typename TV>
__device__ __forceinline__ void mixed_coo_group(int p0, int p1, int64_t r0, int lane, const double* __restrict__ x,
                                                const int* __restrict__ cr, const int* __restrict__ cc,
                                                const TV* __restrict__ cv, float* s_v, float* s_x, int* s_lo,
                                                int* s_hi, float& cs, float& cabs, int& nterms) {
  for (int base = p0; base < p1; base += kWave) {
    const int i = base + lane;
    const bool ok = i < p1;
    int r = -1;
    float vv = 0.0f, xx = 0.0f;
    if (ok) {
      r = cr[i] - static_cast<int>(r0);
      vv = lp_to_f32<TV>(cv[i]);
      xx = static_cast<float>(x[cc[i]]);
      if (r < 0 || r >= kWave) r = -1;
    }
    const int rp = __shfl_up(r, 1);
    const int rn = __shfl_down(r, 1);
    s_lo[lane] = 0;
    s_hi[lane] = 0;
    s_v[lane] = vv;
    s_x[lane] = xx;
    wave_sync();
    if (ok && r >= 0 && (lane == 0 || rp != r)) s_lo[r] = lane;
    if (ok && r >= 0 && (lane == kWave - 1 || i + 1 >= p1 || rn != r)) s_hi[r] = lane + 1;
    wave_sync();
    const int a = s_lo[lane], e = s_hi[lane];
    nterms += e - a;
    for (int k = a; k < e; ++k) {
      cs = __builtin_fmaf(s_v[k], s_x[k], cs);
      cabs = __builtin_fmaf(__builtin_fabsf(s_v[k]), __builtin_fabsf(s_x[k]), cabs);
    }
    wave_sync();
  }
}

// The cancellation flag and the store of one 64-row group (rows r0 + lane)
// from its blocked and fallback sums, exactly k_hbdia_mixed's
// Returns the group's flag mask (the rows left for the FP64 recomputation).
template <bool HAS_COO>
__device__ __forceinline__ uint64_t mixed_store(int64_t grp, int64_t r0, int lane, float sum, float sabs, int cnt,
                                                float cs, float cabs, int nterms, double* __restrict__ y,
                                                int64_t nrows, float ratio, float u_store,
                                                uint64_t* __restrict__ flag_mask) {
  const int64_t row = r0 + lane;
  const bool active = row < nrows;
  const float tot = sum + cs;
  const float mag = sabs + cabs;
  const float u32 = 5.9604645e-08f;  // 2^-24
  const float err = (u_store + u32 * static_cast<float>(nterms + 2)) * mag * 1.0001f;
  const bool fix = active && (!((__builtin_fabsf(tot) - err) * ratio >= mag) || !__builtin_isfinite(tot)) &&
                   (cnt > 0 || cs != 0.0f || cabs != 0.0f);
  const uint64_t m = __ballot(fix);
  if (lane == 0) flag_mask[grp] = m;
  if (!active || fix) return m;
  if constexpr (HAS_COO) {
    const double base = cnt > 0 ? static_cast<double>(sum) : y[row];
    __builtin_nontemporal_store(base + static_cast<double>(cs), y + row);
  } else {
    if (cnt > 0) __builtin_nontemporal_store(static_cast<double>(sum), y + row);
  }
  return m;
}

// Four adjacent rows per lane for the reduced-precision kernel (block width a
// multiple of 64, runs-of-three hints): a 256-row wave whose four 64-row
// quarters share one offset list loads the four rows' stored values of an
// offset as one access (16 bytes FP32, 8 bytes BF16) and the four rows' x of a
// run's centre as two 16-byte loads; the run's outer neighbours come by DPP
// (converted to FP32 first) and the two values beyond the wave's ends are
// scalar loads.  Per run the four rows' FP32 sums take every offset of the run
// as an FMA: a zero value adds a signed zero, which leaves the sums and the
// bound unchanged, and an infinite or NaN x makes the row's sum NaN, which
// flags it for FP64.  So a run whose columns lie inside the matrix for every
// lane needs no per-term test; other runs zero the values of columns outside
// the matrix.  The flags, fallback sums and stores are k_hbdia_mixed's per
// 64-row group, after the sums are moved back to row-per-lane order in LDS.

typedef float flt4a4m __attribute__((ext_vector_type(4), aligned(4)));
typedef unsigned u32x2a4 __attribute__((ext_vector_type(2), aligned(4)));

// the four rows' stored values of one offset (p at a multiple of 4 elements)
// A lane's four stored values of an offset: the raw load (issued first, so the
// x loads of the same run go out before anything waits on it) and its FP32
// values.
template <typename TV> struct LpRaw;
template <> struct LpRaw<float> { using type = flt4a4m; };
template <> struct LpRaw<uint16_t> { using type = u32x2a4; };
template <typename TV>
__device__ __forceinline__ typename LpRaw<TV>::type lp_quad_raw(const TV* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const typename LpRaw<TV>::type*>(p));
}
__device__ __forceinline__ void lp_quad_cvt(const flt4a4m& t, float* v) {
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}
__device__ __forceinline__ void lp_quad_cvt(const u32x2a4& t, float* v) {
  v[0] = __uint_as_float(t.x << 16);
  v[1] = __uint_as_float(t.x & 0xffff0000u);
  v[2] = __uint_as_float(t.y << 16);
  v[3] = __uint_as_float(t.y & 0xffff0000u);
}
template <typename TV>
__device__ __forceinline__ void lp_quad(const TV* p, float* v) {
  lp_quad_cvt(lp_quad_raw<TV>(p), v);
}

// mixed_coo_group for the four 64-row groups of a 256-row wave whose entries
// [p[0], p[4]) fit one 64-entry pass (group g holds [p[g], p[g + 1]), rows
// row0 + 64 g ...): one load for all, then the same per-row sums in entry
// order.  s_lo and s_hi hold 256 entries.
template <typename TV>
__device__ __forceinline__ void mixed_coo_quad(const int* p, int64_t row0, int lane, const double* __restrict__ x,
                                               const int* __restrict__ cr, const int* __restrict__ cc,
                                               const TV* __restrict__ cv, float* s_v, float* s_x, int* s_lo,
                                               int* s_hi, float* cs, float* cabs, int* nterms) {
  const int i = p[0] + lane;
  const bool ok = i < p[4];
  int r = -1;
  float vv = 0.0f, xx = 0.0f;
  if (ok) {
    r = cr[i] - static_cast<int>(row0);
    vv = lp_to_f32<TV>(cv[i]);
    xx = static_cast<float>(x[cc[i]]);
    // an entry outside its own group's rows (unsorted input) is dropped
    if (r < 0 || r >= 4 * kWave || i < p[r >> 6] || i >= p[(r >> 6) + 1]) r = -1;
  }
  const int rp = __shfl_up(r, 1);
  const int rn = __shfl_down(r, 1);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    s_lo[g * kWave + lane] = 0;
    s_hi[g * kWave + lane] = 0;
  }
  s_v[lane] = vv;
  s_x[lane] = xx;
  wave_sync();
  if (ok && r >= 0 && (lane == 0 || rp != r)) s_lo[r] = lane;
  if (ok && r >= 0 && (lane == kWave - 1 || i + 1 >= p[4] || rn != r)) s_hi[r] = lane + 1;
  wave_sync();
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int a = s_lo[h * kWave + lane], e = s_hi[h * kWave + lane];
    nterms[h] += e - a;
    float c = 0.0f, ca = 0.0f;
    for (int k = a; k < e; ++k) {
      c = __builtin_fmaf(s_v[k], s_x[k], c);
      ca = __builtin_fmaf(__builtin_fabsf(s_v[k]), __builtin_fabsf(s_x[k]), ca);
    }
    cs[h] = c;
    cabs[h] = ca;
  }
  wave_sync();
}

template <typename TV, bool HAS_COO, bool FIX_INLINE>
__global__ __launch_bounds__(kWG) void k_hbdia_mixed_row4(
    const TV* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk,
    const double* __restrict__ x, double* __restrict__ y, int64_t nrows, int nb, int bw,
    const int* __restrict__ cr, const int* __restrict__ cc, const TV* __restrict__ cv,
    const int* __restrict__ gptr, float ratio, float u_store, uint64_t* __restrict__ flag_mask,
    const uint32_t* __restrict__ runs3_blocks, int kshift, MixedFp64 f64) {
  __shared__ float s_v[kWavesPerWG][kWave];
  __shared__ float s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][4 * kWave];
  __shared__ int s_hi[kWavesPerWG][4 * kWave];
  __shared__ __attribute__((aligned(16))) float s_y[kWavesPerWG][8 * kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned work = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const int64_t w4 = static_cast<int64_t>(work) * kWavesPerWG + wv;  // 256-row wave
  const int64_t row0 = w4 * 4 * kWave;
  if (row0 >= nrows) return;
  const int64_t ngroups = (nrows + kWave - 1) / kWave;
  const int nq = static_cast<int>(min<int64_t>(4, ngroups - 4 * w4));  // 64-row groups present
  int64_t bq[4];
  int sq[4], cq[4];
  bool rq[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    bq[q] = (row0 + q * kWave) / bw;
    const bool have = q < nq && bq[q] < nb;
    sq[q] = have ? blk[bq[q]] : 0;
    cq[q] = have ? blk[bq[q] + 1] - sq[q] : 0;
    rq[q] = have && ((runs3_blocks[bq[q] >> 5] >> (bq[q] & 31)) & 1u);
  }
  // one offset list of runs for the whole wave: its blocks (consecutive) hold
  // the same count and offsets
  const int c0 = cq[0];
  bool same = nq == 4 && rq[0] && c0 > 0 && c0 <= kWave && c0 % 3 == 0;
  if (same) {
    const int oa = lane < c0 ? offs[sq[0] + lane] : 0;
    bool diff = false;
#pragma unroll
    for (int q = 1; q < 4; ++q) {
      if (bq[q] != bq[0]) {
        same = same && cq[q] == c0;
        if (cq[q] == c0) diff = diff | (oa != (lane < c0 ? offs[sq[q] + lane] : 0));
      }
    }
    same = same && __ballot(diff) == 0;
  }
  float sum[4], sab[4];
  if (same) {
    const int64_t row = row0 + 4 * lane;
    const int64_t last = nrows - 1;
    const uint32_t row32 = static_cast<uint32_t>(row), n32 = static_cast<uint32_t>(nrows);
    const int64_t bl = row / bw;  // the lane's block: its values follow the wave's first block's
    const TV* vb = vals + static_cast<int64_t>(sq[0]) * bw +
                   ((bl - bq[0]) * c0 * bw + (row - bl * bw));
    const int* op = offs + sq[0];
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f}, a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // one run of three per step; every run but the last loads the next run's
    // offsets and values while it waits for its own x (the last run, peeled,
    // loads nothing ahead, so no step keeps a value it did not load)
    int offn[3];
    typename LpRaw<TV>::type rvn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) offn[k] = op[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) rvn[k] = lp_quad_raw<TV>(vb + static_cast<int64_t>(k) * bw);
    auto run = [&](auto ahead, int o0) {
      int off[3];
      typename LpRaw<TV>::type rv[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        off[k] = offn[k];
        rv[k] = rvn[k];
      }
      const int64_t o = off[1];
      const int64_t lo = row0 + o - 1, hi = row0 + o + 4 * kWave;  // wave-uniform
      const bool inside = lo >= 0 && hi <= last;
      const int64_t cl = lo < 0 ? 0 : (lo > last ? last : lo), ch = hi < 0 ? 0 : (hi > last ? last : hi);
      double m[4];
      if (inside) {
        const dbl2a8 p0 = *reinterpret_cast<const dbl2a8*>(x + (row + o));
        const dbl2a8 p1 = *reinterpret_cast<const dbl2a8*>(x + (row + o + 2));
        m[0] = p0.x;
        m[1] = p0.y;
        m[2] = p1.x;
        m[3] = p1.y;
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          int64_t c = row + i + o;
          c = c < 0 ? 0 : (c > last ? last : c);
          m[i] = x[c];
        }
      }
      const double eld = x[cl], erd = x[ch];  // scalar loads
      if constexpr (decltype(ahead)::value) {
#pragma unroll
        for (int k = 0; k < 3; ++k) offn[k] = op[o0 + 3 + k];
#pragma unroll
        for (int k = 0; k < 3; ++k) rvn[k] = lp_quad_raw<TV>(vb + static_cast<int64_t>(o0 + 3 + k) * bw);
      }
      // every load of the run is out; now the values
      float v[3][4];
#pragma unroll
      for (int k = 0; k < 3; ++k) lp_quad_cvt(rv[k], v[k]);
      if (!inside) {
        // columns outside the matrix: their values as zeros
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const uint32_t c = row32 + static_cast<uint32_t>(off[k]);
#pragma unroll
          for (int i = 0; i < 4; ++i) v[k][i] = c + static_cast<uint32_t>(i) < n32 ? v[k][i] : 0.0f;
        }
      }
      const float el = static_cast<float>(eld), er = static_cast<float>(erd);
      float xf[6];
#pragma unroll
      for (int i = 0; i < 4; ++i) xf[i + 1] = static_cast<float>(m[i]);
      xf[0] = wave_shr1(xf[4], el);  // x[row + o - 1]
      xf[5] = wave_shl1(xf[1], er);  // x[row + o + 4]
#pragma unroll
      for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          s[i] = __builtin_fmaf(v[k][i], xf[i + k], s[i]);
          a[i] = __builtin_fmaf(__builtin_fabsf(v[k][i]), __builtin_fabsf(xf[i + k]), a[i]);
        }
      }
    };
    for (int o0 = 0; o0 < c0 - 3; o0 += 3) run(std::true_type{}, o0);
    run(std::false_type{}, c0 - 3);
    // back to row-per-lane order: group q's row r0 + lane at [64 q + lane]
    float* sy = s_y[wv];
    *reinterpret_cast<float4*>(sy + 4 * lane) = make_float4(s[0], s[1], s[2], s[3]);
    *reinterpret_cast<float4*>(sy + 4 * kWave + 4 * lane) = make_float4(a[0], a[1], a[2], a[3]);
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      sum[q] = sy[q * kWave + lane];
      sab[q] = sy[4 * kWave + q * kWave + lane];
    }
    wave_sync();
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      sum[q] = sab[q] = 0.0f;
      if (q < nq)
        mixed_half<TV>(vals, offs, x, row0 + q * kWave, bq[q], sq[q], cq[q], bw, lane, nrows, rq[q], sum[q], sab[q]);
    }
  }
  float cs[4] = {0.0f, 0.0f, 0.0f, 0.0f}, cabs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int nterms[4] = {cq[0], cq[1], cq[2], cq[3]};
  if constexpr (HAS_COO) {
    int p[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) p[q] = gptr[4 * w4 + min(q, nq)];
    if (p[4] - p[0] <= kWave) {
      if (p[0] < p[4])
        mixed_coo_quad<TV>(p, row0, lane, x, cr, cc, cv, s_v[wv], s_x[wv], s_lo[wv], s_hi[wv], cs, cabs, nterms);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        mixed_coo_group<TV>(p[q], p[q + 1], row0 + q * kWave, lane, x, cr, cc, cv, s_v[wv], s_x[wv], s_lo[wv],
                            s_hi[wv], cs[q], cabs[q], nterms[q]);
    }
  }
  uint64_t fm[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q < nq)
      fm[q] = mixed_store<HAS_COO>(4 * w4 + q, row0 + q * kWave, lane, sum[q], sab[q], cq[q], cs[q], cabs[q],
                                   nterms[q], y, nrows, ratio, u_store, flag_mask);
  }
  if constexpr (FIX_INLINE) {
    // flagged rows, recomputed in FP64 by their own lanes once every sum is stored
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (fm[q] != 0 && ((fm[q] >> lane) & 1u))
        fix_row<3>(row0 + q * kWave + lane, f64.vals, offs, blk, x, y, nrows, nb, bw, cr, cc, f64.cv,
                       HAS_COO ? gptr : nullptr);
    }
  }
}


// Recomputes every row whose bit is set in flag_mask (one 64-bit mask per
// 64-row group).  Each wave scans the masks of 64 consecutive groups
// (grid-stride), appends the flagged rows to a wave-private LDS queue and
// recomputes them 64 at a time, one row per lane, so lanes stay busy whether
// the flags are sparse or dense.
__global__ __launch_bounds__(256) void k_fix_groups(const double* __restrict__ vals, const int* __restrict__ offs,
                                                    const int* __restrict__ blk, const double* __restrict__ x,
                                                    double* __restrict__ y, int64_t nrows, int nb, int bw,
                                                    const int* __restrict__ cr, const int* __restrict__ cc,
                                                    const double* __restrict__ cv, const int* __restrict__ gptr,
                                                    const uint64_t* __restrict__ flag_mask) {
  __shared__ int64_t s_q[4][2 * kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x >> 6;
  int64_t* q = s_q[wv];
  const uint64_t below = (1ull << lane) - 1ull;  // lanes under this one
  const int64_t ngroups = (nrows + kWave - 1) / kWave;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + wv;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
  int qn = 0;  // wave-uniform queue length, < 64 between groups
  for (int64_t gbase = wave * kWave; gbase < ngroups; gbase += nwaves * kWave) {
    const uint64_t mine = gbase + lane < ngroups ? flag_mask[gbase + lane] : 0ull;
    uint64_t nz = __ballot(mine != 0ull);
    while (nz) {
      const int j = __ffsll(static_cast<unsigned long long>(nz)) - 1;
      nz &= nz - 1;
      const uint64_t m = __shfl(mine, j);
      if ((m >> lane) & 1ull) q[qn + __popcll(m & below)] = (gbase + j) * kWave + lane;
      qn += __popcll(m);
      if (qn >= kWave) {  // a full batch: recompute q[0..63], move the rest down
        wave_sync();
        fix_row<kChunk>(q[lane], vals, offs, blk, x, y, nrows, nb, bw, cr, cc, cv, gptr);
        const int64_t rest = lane + kWave < qn ? q[lane + kWave] : 0;
        wave_sync();
        if (lane + kWave < qn) q[lane] = rest;
        qn -= kWave;
        wave_sync();
      }
    }
  }
  wave_sync();
  if (lane < qn) fix_row<kChunk>(q[lane], vals, offs, blk, x, y, nrows, nb, bw, cr, cc, cv, gptr);
}

// COO fallback alone: res[r] = row sum for rows with entries (others keep res).
template <typename T>
__global__ __launch_bounds__(kWG) void k_coo(const int* __restrict__ cr, const int* __restrict__ cc,
                                             const T* __restrict__ cv, const int* __restrict__ gptr,
                                             const T* __restrict__ x, T* __restrict__ res,
                                             int64_t nrows) {
  __shared__ T s_v[kWavesPerWG][kWave];
  __shared__ T s_x[kWavesPerWG][kWave];
  __shared__ int s_lo[kWavesPerWG][kWave];
  __shared__ int s_hi[kWavesPerWG][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t grp = static_cast<int64_t>(blockIdx.x) * kWavesPerWG + wv;
  const int64_t row0 = grp * kWave;
  if (row0 >= nrows) return;
  const int p0 = gptr[grp], p1 = gptr[grp + 1];
  if (p0 >= p1) return;
  bool any = false;
  const T cs = coo_group_sum<T>(cr, cc, cv, x, p0, p1, row0, lane, s_v[wv], s_x[wv], s_lo[wv],
                                s_hi[wv], &any);
  const int64_t row = row0 + lane;
  if (any && row < nrows) res[row] = cs;
}

// group_ptr[g] = lower_bound(rows, 64 g) for g in [0, G].
__global__ void k_group_ptr(const int* __restrict__ rows, int64_t nnz, int64_t ngroups,
                            int* __restrict__ gptr) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g > ngroups) return;
  const int64_t key = g * kWave;
  int64_t lo = 0, hi = nnz;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (static_cast<int64_t>(rows[mid]) < key) lo = mid + 1;
    else hi = mid;
  }
  gptr[g] = static_cast<int>(lo);
}

template <typename T>
__global__ void k_vector_add(T* __restrict__ dst, const T* __restrict__ src, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] += src[i];
}

int64_t num_groups(int64_t nrows) { return (nrows + kWave - 1) / kWave; }

template <typename T>
int launch_hbdia(const T* vals, const int32_t* offs, const int32_t* blk, int32_t nb, int32_t bw,
                 const int32_t* cr, const int32_t* cc, const T* cv, const int32_t* gptr,
                 const T* x, T* y, int64_t nrows, hbdia_stream_t stream, bool with_coo,
                 const char* name, const uint32_t* runs3_blocks = nullptr) {
  if (nrows < 0 || nb < 0 || bw <= 0)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": negative size or block_width <= 0");
  if (nrows == 0) return HBDIA_SUCCESS;
  if (!x || !y || !blk || (nb > 0 && !offs))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null pointer");
  if (with_coo && (!cr || !cc || !cv || !gptr))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null COO pointer");
  const int64_t groups = num_groups(nrows);
  const int64_t wgs = (groups + kWavesPerWG - 1) / kWavesPerWG;
  if (wgs > 0x7fffffffLL) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": too many rows");
  const dim3 grid(static_cast<unsigned>(wgs)), block(kWG);
  const bool uniform = (bw % kWave) == 0;
  // 32-bit byte offsets for x when x fits, and for a block's values always
  // (checked by the caller through the largest block).
  const bool x32 = static_cast<uint64_t>(nrows) * sizeof(T) < (1ull << 32);
  hipStream_t s = hs(stream);
  const bool runs3 = (uniform || bw == 32) && runs3_blocks != nullptr;
  const int kshift = xcd_chunk_shift(wgs);
  {
    if (runs3 && bw == 32 && nrows <= 0x7fffffffLL) {
      // 128-row waves over four width-32 blocks, two rows per lane
      const int64_t wgs2 = (nrows + 2 * kWG - 1) / (2 * kWG);
      const dim3 grid2(static_cast<unsigned>(wgs2));
      const int ks2 = xcd_chunk_shift(wgs2);
#define HBDIA_LAUNCH_W32(C, X)                                                                                 \
  hipLaunchKernelGGL((k_hbdia_row2_w32<T, C, X>), grid2, block, 0, s, vals, offs, blk, x, y, nrows, nb, cr, cc, cv, \
                     gptr, runs3_blocks, ks2)
      if (with_coo) { if (x32) HBDIA_LAUNCH_W32(true, true); else HBDIA_LAUNCH_W32(true, false); }
      else { if (x32) HBDIA_LAUNCH_W32(false, true); else HBDIA_LAUNCH_W32(false, false); }
#undef HBDIA_LAUNCH_W32
      return check_hip(hipGetLastError(), name);
    }
    if (runs3 && uniform && nrows <= 0x7fffffffLL) {
      // 128-row waves, two rows per lane (32-bit column bound checks)
      const int64_t wgs2 = (nrows + 2 * kWG - 1) / (2 * kWG);
      const dim3 grid2(static_cast<unsigned>(wgs2));
      const int ks2 = xcd_chunk_shift(wgs2);
#define HBDIA_LAUNCH_R2(C, X)                                                                                  \
  hipLaunchKernelGGL((k_hbdia_row2<T, C, X>), grid2, block, 0, s, vals, offs, blk, x, y, nrows, nb, bw, cr, cc, cv, \
                     gptr, runs3_blocks, ks2)
      if (with_coo) { if (x32) HBDIA_LAUNCH_R2(true, true); else HBDIA_LAUNCH_R2(true, false); }
      else { if (x32) HBDIA_LAUNCH_R2(false, true); else HBDIA_LAUNCH_R2(false, false); }
#undef HBDIA_LAUNCH_R2
      return check_hip(hipGetLastError(), name);
    }
  }
#define HBDIA_LAUNCH(U, C, X, R)                                                                             \
  hipLaunchKernelGGL((k_hbdia<T, U, C, X, R>), grid, block, 0, s, vals, offs, blk, x, y, nrows, nb, bw, cr, cc, \
                     cv, gptr, runs3_blocks, kshift)
  if (runs3 && uniform) {
    if (with_coo) { if (x32) HBDIA_LAUNCH(true, true, true, true); else HBDIA_LAUNCH(true, true, false, true); }
    else { if (x32) HBDIA_LAUNCH(true, false, true, true); else HBDIA_LAUNCH(true, false, false, true); }
  } else if (runs3) {
    if (with_coo) { if (x32) HBDIA_LAUNCH(false, true, true, true); else HBDIA_LAUNCH(false, true, false, true); }
    else { if (x32) HBDIA_LAUNCH(false, false, true, true); else HBDIA_LAUNCH(false, false, false, true); }
  } else if (uniform) {
    if (with_coo) { if (x32) HBDIA_LAUNCH(true, true, true, false); else HBDIA_LAUNCH(true, true, false, false); }
    else { if (x32) HBDIA_LAUNCH(true, false, true, false); else HBDIA_LAUNCH(true, false, false, false); }
  } else {
    if (with_coo) { if (x32) HBDIA_LAUNCH(false, true, true, false); else HBDIA_LAUNCH(false, true, false, false); }
    else { if (x32) HBDIA_LAUNCH(false, false, true, false); else HBDIA_LAUNCH(false, false, false, false); }
  }
#undef HBDIA_LAUNCH
  return check_hip(hipGetLastError(), name);
}

template <typename T, int KB>
void launch_multi_kb(const T* vals, const int32_t* offs, const int32_t* blk, int32_t nb, int32_t bw,
                     const int32_t* cr, const int32_t* cc, const T* cv, const int32_t* gptr, const T* X,
                     int64_t ldx, T* Y, int64_t ldy, int nrhs, int64_t nrows, bool with_coo, bool x32,
                     const uint32_t* runs3, int kshift, dim3 grid, hipStream_t s) {
#define HBDIA_LAUNCH_M(C, XX, R)                                                                              \
  hipLaunchKernelGGL((k_hbdia_multi<T, KB, C, XX, R>), grid, dim3(kWG), 0, s, vals, offs, blk, X, ldx, Y, ldy, \
                     nrhs, nrows, nb, bw, cr, cc, cv, gptr, runs3, kshift)
  if (runs3) {
    if (with_coo) { if (x32) HBDIA_LAUNCH_M(true, true, true); else HBDIA_LAUNCH_M(true, false, true); }
    else { if (x32) HBDIA_LAUNCH_M(false, true, true); else HBDIA_LAUNCH_M(false, false, true); }
  } else {
    if (with_coo) { if (x32) HBDIA_LAUNCH_M(true, true, false); else HBDIA_LAUNCH_M(true, false, false); }
    else { if (x32) HBDIA_LAUNCH_M(false, true, false); else HBDIA_LAUNCH_M(false, false, false); }
  }
#undef HBDIA_LAUNCH_M
}

template <int KB>
void launch_multi_row2_kb(const double* vals, const int32_t* offs, const int32_t* blk, int32_t nb, int32_t bw,
                          const int32_t* cr, const int32_t* cc, const double* cv, const int32_t* gptr,
                          const double* X, int64_t ldx, double* Y, int64_t ldy, int nrhs, int64_t nrows,
                          bool with_coo, bool x32, const uint32_t* runs3, int kshift, dim3 grid, hipStream_t s) {
#define HBDIA_LAUNCH_M2(C, XX, N)                                                                                \
  hipLaunchKernelGGL((k_hbdia_multi_row2<KB, C, XX, N>), grid, dim3(kWG), 0, s, vals, offs, blk, X, ldx, Y, ldy, \
                     nrhs, nrows, nb, bw, cr, cc, cv, gptr, runs3, kshift)
  // one column pass streams the values; with more, the later passes find them in L2
  if (nrhs <= KB) {
    if (with_coo) { if (x32) HBDIA_LAUNCH_M2(true, true, true); else HBDIA_LAUNCH_M2(true, false, true); }
    else { if (x32) HBDIA_LAUNCH_M2(false, true, true); else HBDIA_LAUNCH_M2(false, false, true); }
  } else {
    if (with_coo) { if (x32) HBDIA_LAUNCH_M2(true, true, false); else HBDIA_LAUNCH_M2(true, false, false); }
    else { if (x32) HBDIA_LAUNCH_M2(false, true, false); else HBDIA_LAUNCH_M2(false, false, false); }
  }
#undef HBDIA_LAUNCH_M2
}

template <typename T>
int launch_multi(const T* vals, const int32_t* offs, const int32_t* blk, int32_t nb, int32_t bw,
                 const int32_t* cr, const int32_t* cc, const T* cv, const int32_t* gptr, const T* X, int64_t ldx,
                 T* Y, int64_t ldy, int32_t nrhs, int64_t nrows, hbdia_stream_t stream, const char* name,
                 const uint32_t* runs3 = nullptr) {
  if (nrows < 0 || nb < 0 || bw <= 0 || nrhs < 0)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": negative size or block_width <= 0");
  if (nrows == 0 || nrhs == 0) return HBDIA_SUCCESS;
  if (bw % kWave != 0)
    return fail(HBDIA_ERROR_NOT_SUPPORTED, std::string(name) + ": block_width must be a multiple of 64");
  if (!X || !Y || !blk || (nb > 0 && !offs))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null pointer");
  if (ldx < nrows || ldy < nrows)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": leading dimension smaller than num_rows");
  const bool with_coo = gptr != nullptr;
  if (with_coo && (!cr || !cc || !cv))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null COO pointer");
  if (nrows > 0x7fffffffLL) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": too many rows");
  const int64_t groups = num_groups(nrows);
  const int64_t wgs = (groups + kWavesPerWG - 1) / kWavesPerWG;
  const bool x32 = static_cast<uint64_t>(nrows) * sizeof(T) < (1ull << 32);
  // 8 columns per pass; more columns take more passes, a row block's passes
  // adjacent.  Passes of 16 columns measured slower (k = 16: 0.679 against
  // 0.417 ms for two passes of 8), and 32 spill.
  const int kb = nrhs >= 8 ? 8 : (nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1));
  const int64_t items = wgs * ((nrhs + kb - 1) / kb);
  if (items > 0x7fffffffLL) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": too many rows x columns");
  hipStream_t s = hs(stream);
  if constexpr (std::is_same<T, double>::value) {
    if (runs3) {
      // 128-row waves, two rows per lane, up to 8 columns per pass; a row
      // block's column passes are adjacent work items, so the later ones read
      // its values from L2.  The launch is bound by L2 requests (the values
      // once per pass, x nine times per column): 8 columns per pass measured
      // k = 32 at 0.617 ms against 0.668 ms with 4.
      if (!with_coo && nrhs > 8) {
        // more columns than one pass: every column in one wave, the values
        // held in registers and read once
        const int64_t wgs2 = (nrows + 2 * kWG - 1) / (2 * kWG);
        const dim3 grid2(static_cast<unsigned>(wgs2));
        const int ks2 = xcd_chunk_shift(wgs2);
        if (x32)
          hipLaunchKernelGGL((k_hbdia_multi_res<true>), grid2, dim3(kWG), 0, s, vals, offs, blk, X, ldx, Y, ldy, nrhs,
                             nrows, nb, bw, runs3, ks2);
        else
          hipLaunchKernelGGL((k_hbdia_multi_res<false>), grid2, dim3(kWG), 0, s, vals, offs, blk, X, ldx, Y, ldy,
                             nrhs, nrows, nb, bw, runs3, ks2);
        return check_hip(hipGetLastError(), name);
      }
      const int kb2 = kb;
      const int64_t items2 = (nrows + 2 * kWG - 1) / (2 * kWG) * ((nrhs + kb2 - 1) / kb2);
      const dim3 grid2(static_cast<unsigned>(items2));
      const int ks2 = xcd_chunk_shift(items2);
      switch (kb2) {
        case 8: launch_multi_row2_kb<8>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, ldx, Y, ldy, nrhs, nrows, with_coo, x32, runs3, ks2, grid2, s); break;
        case 4: launch_multi_row2_kb<4>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, ldx, Y, ldy, nrhs, nrows, with_coo, x32, runs3, ks2, grid2, s); break;
        case 2: launch_multi_row2_kb<2>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, ldx, Y, ldy, nrhs, nrows, with_coo, x32, runs3, ks2, grid2, s); break;
        default: launch_multi_row2_kb<1>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, ldx, Y, ldy, nrhs, nrows, with_coo, x32, runs3, ks2, grid2, s); break;
      }
      return check_hip(hipGetLastError(), name);
    }
  }
  const dim3 grid(static_cast<unsigned>(items));
  const int ks = xcd_chunk_shift(items);
  switch (kb) {
    case 8: launch_multi_kb<T, 8>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, ldx, Y, ldy, nrhs, nrows, with_coo, x32, runs3, ks, grid, s); break;
    case 4: launch_multi_kb<T, 4>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, ldx, Y, ldy, nrhs, nrows, with_coo, x32, runs3, ks, grid, s); break;
    case 2: launch_multi_kb<T, 2>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, ldx, Y, ldy, nrhs, nrows, with_coo, x32, runs3, ks, grid, s); break;
    default: launch_multi_kb<T, 1>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, ldx, Y, ldy, nrhs, nrows, with_coo, x32, runs3, ks, grid, s); break;
  }
  return check_hip(hipGetLastError(), name);
}

int launch_mixed(int vdtype, const void* vals_lp, const double* vals64, const int32_t* offs, const int32_t* blk,
                 int32_t nb, int32_t bw, const int32_t* cr, const int32_t* cc, const void* cv_lp, const double* cv64,
                 const int32_t* gptr, const double* x, double* y, int64_t nrows, double ratio, void* workspace,
                 hbdia_stream_t stream, const uint32_t* runs3 = nullptr) {
  const char* name = "hbdia_hbdia_spmv_mixed";
  if (nrows < 0 || nb < 0 || bw <= 0) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": bad size");
  if (nrows == 0) return HBDIA_SUCCESS;
  if (bw % kWave != 0) return fail(HBDIA_ERROR_NOT_SUPPORTED, std::string(name) + ": block_width must be a multiple of 64");
  if (vdtype != HBDIA_VALUES_F32 && vdtype != HBDIA_VALUES_BF16)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": vdtype must be HBDIA_VALUES_F32 or HBDIA_VALUES_BF16");
  if (!vals_lp || !vals64 || !blk || !x || !y || !workspace || (nb > 0 && !offs))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null pointer");
  if (nrows > 0x7fffffffLL) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": too many rows");
  const bool with_coo = gptr != nullptr;
  if (with_coo && (!cr || !cc || !cv_lp || !cv64))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null COO pointer");
  if (!(ratio > 0.0)) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": ratio must be positive");
  uint64_t* mask = reinterpret_cast<uint64_t*>(workspace);
  hipStream_t s = hs(stream);
  const int64_t groups = num_groups(nrows);
  const dim3 grid(static_cast<unsigned>((groups + kWavesPerWG - 1) / kWavesPerWG)), block(kWG);
  const float r = static_cast<float>(ratio);
  const float u_store = vdtype == HBDIA_VALUES_F32 ? 5.9604645e-08f : 3.9062500e-03f + 5.9604645e-08f;
  const int ks = xcd_chunk_shift(grid.x);
  // FP32 storage flags few rows: each is recomputed in FP64 by its own lane
  // at the end of the kernel (a short stall of that wave, hidden by the
  // others), and no fix-up pass follows (192^3: 0.1775 -> 0.1735 ms per
  // SpMV).  BF16 storage flags many more (a third of the waves at ratio 1e4 at
  // 192^3): in-kernel they measured 0.142 -> 0.185 ms, so they go to the
  // fix-up pass, which recomputes them 64 at a time.
  const bool inl = vdtype == HBDIA_VALUES_F32;
  const MixedFp64 f64{vals64, cv64};
  if (runs3) {
    // 256-row waves, four rows per lane
    const int64_t wgs4 = (nrows + 4 * kWG - 1) / (4 * kWG);
    const dim3 grid2(static_cast<unsigned>(wgs4));
    const int ks2 = xcd_chunk_shift(wgs4);
#define HBDIA_LAUNCH_MX2(TV, C, I)                                                                                \
  hipLaunchKernelGGL((k_hbdia_mixed_row4<TV, C, I>), grid2, block, 0, s, static_cast<const TV*>(vals_lp), offs, blk, \
                     x, y, nrows, nb, bw, cr, cc, static_cast<const TV*>(cv_lp), gptr, r, u_store, mask, runs3, ks2, f64)
    if (vdtype == HBDIA_VALUES_F32) {
      if (with_coo) { if (inl) HBDIA_LAUNCH_MX2(float, true, true); else HBDIA_LAUNCH_MX2(float, true, false); }
      else { if (inl) HBDIA_LAUNCH_MX2(float, false, true); else HBDIA_LAUNCH_MX2(float, false, false); }
    } else {
      if (with_coo) { if (inl) HBDIA_LAUNCH_MX2(uint16_t, true, true); else HBDIA_LAUNCH_MX2(uint16_t, true, false); }
      else { if (inl) HBDIA_LAUNCH_MX2(uint16_t, false, true); else HBDIA_LAUNCH_MX2(uint16_t, false, false); }
    }
#undef HBDIA_LAUNCH_MX2
  } else {
#define HBDIA_LAUNCH_MX(TV, C, I)                                                                             \
  hipLaunchKernelGGL((k_hbdia_mixed<TV, C, false, I>), grid, block, 0, s, static_cast<const TV*>(vals_lp), offs, blk, \
                     x, y, nrows, nb, bw, cr, cc, static_cast<const TV*>(cv_lp), gptr, r, u_store, mask, runs3, ks, f64)
    if (vdtype == HBDIA_VALUES_F32) {
      if (with_coo) { if (inl) HBDIA_LAUNCH_MX(float, true, true); else HBDIA_LAUNCH_MX(float, true, false); }
      else { if (inl) HBDIA_LAUNCH_MX(float, false, true); else HBDIA_LAUNCH_MX(float, false, false); }
    } else {
      if (with_coo) { if (inl) HBDIA_LAUNCH_MX(uint16_t, true, true); else HBDIA_LAUNCH_MX(uint16_t, true, false); }
      else { if (inl) HBDIA_LAUNCH_MX(uint16_t, false, true); else HBDIA_LAUNCH_MX(uint16_t, false, false); }
    }
#undef HBDIA_LAUNCH_MX
  }
  if (!inl) {
    const int64_t scan_waves = (groups + kWave - 1) / kWave;
    const int64_t fix_blocks = std::min<int64_t>((scan_waves + 3) / 4, 2048);
    hipLaunchKernelGGL(k_fix_groups, dim3(static_cast<unsigned>(fix_blocks)), dim3(256), 0, s, vals64, offs, blk, x,
                       y, nrows, nb, bw, cr, cc, cv64, gptr, mask);
  }
  return check_hip(hipGetLastError(), name);
}

template <typename T>
int launch_batch(const hbdia_batch_t* bt, const hbdia_batch_schedule_t* sc, const T* x, T* y, hbdia_stream_t stream,
                 const char* name) {
  if (!bt || !sc) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null batch or schedule");
  if (bt->num_tiles < 0) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": negative sizes");
  if (sc->num_workgroups < 1 || sc->tiles_per_step < 1 ||
      (sc->kind != HBDIA_BATCH_INTERLEAVED && sc->kind != HBDIA_BATCH_CONTIGUOUS && sc->kind != HBDIA_BATCH_XCD) ||
      (sc->kind == HBDIA_BATCH_XCD && sc->num_workgroups % 8 != 0))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid schedule");
  if (bt->num_tiles == 0) return HBDIA_SUCCESS;
  if (!x || !y || !bt->tiles || !bt->values || !bt->offsets)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null pointer");
  if (bt->has_coo && (!bt->coo_rows || !bt->coo_cols || !bt->coo_values))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null fallback pool");
  if (sc->kind != HBDIA_BATCH_INTERLEAVED && !bt->wave_tiles)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": schedule without wave_tiles");
  const dim3 grid(static_cast<unsigned>(sc->num_workgroups)), block(kWG);
  hipStream_t s = hs(stream);
  const bool coo = bt->has_coo != 0, x32 = bt->x32 != 0, contig = sc->kind == HBDIA_BATCH_CONTIGUOUS;
  const bool lng = bt->long_lists != 0;  // the contiguous schedule runs long lists unpaired
  // interleaved: no split array (the kernel reads wave_tiles as the XCD split when given one)
  const int32_t* split = sc->kind == HBDIA_BATCH_INTERLEAVED ? nullptr : bt->wave_tiles;
#define HBDIA_BATCH_LAUNCH(C, X, G)                                                                       \
  do {                                                                                                    \
    if (lng && !G)                                                                                        \
      hipLaunchKernelGGL((k_hbdia_batch<T, C, X, false, false, true>), grid, block, 0, s,                 \
                         reinterpret_cast<const int32_t*>(bt->tiles), split,                              \
                         static_cast<const T*>(bt->values), bt->offsets, bt->coo_rows, bt->coo_cols,      \
                         static_cast<const T*>(bt->coo_values), bt->num_tiles, sc->tiles_per_step,        \
                         nullptr, 0, x, y);                                                               \
    else                                                                                                  \
      hipLaunchKernelGGL((k_hbdia_batch<T, C, X, G, false, false>), grid, block, 0, s,                    \
                         reinterpret_cast<const int32_t*>(bt->tiles), split,                              \
                         static_cast<const T*>(bt->values), bt->offsets, bt->coo_rows, bt->coo_cols,      \
                         static_cast<const T*>(bt->coo_values), bt->num_tiles, sc->tiles_per_step,        \
                         nullptr, 0, x, y);                                                               \
  } while (0)
  if (coo) {
    if (x32) {
      if (contig) HBDIA_BATCH_LAUNCH(true, true, true); else HBDIA_BATCH_LAUNCH(true, true, false);
    } else {
      if (contig) HBDIA_BATCH_LAUNCH(true, false, true); else HBDIA_BATCH_LAUNCH(true, false, false);
    }
  } else {
    if (x32) {
      if (contig) HBDIA_BATCH_LAUNCH(false, true, true); else HBDIA_BATCH_LAUNCH(false, true, false);
    } else {
      if (contig) HBDIA_BATCH_LAUNCH(false, false, true); else HBDIA_BATCH_LAUNCH(false, false, false);
    }
  }
#undef HBDIA_BATCH_LAUNCH
  return check_hip(hipGetLastError(), name);
}

template <typename T>
int launch_serve(const hbdia_batch_t* bt, const hbdia_serve_job_t* jobs, int32_t num_jobs, int64_t total_tiles,
                 int32_t num_workgroups, int32_t tiles_per_step, hbdia_stream_t stream, const char* name) {
  if (!bt || num_jobs < 0 || total_tiles < 0 || num_workgroups < 1 || tiles_per_step < 1)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid argument");
  if (num_jobs == 0 || total_tiles == 0) return HBDIA_SUCCESS;
  if (!jobs || !bt->tiles || !bt->values || !bt->offsets)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null pointer");
  if (bt->has_coo && (!bt->coo_rows || !bt->coo_cols || !bt->coo_values))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null fallback pool");
  const dim3 grid(static_cast<unsigned>(num_workgroups)), block(kWG);
  const T* no_x = nullptr;  // each job carries its own x and y
  T* no_y = nullptr;
  hipStream_t s = hs(stream);
#define HBDIA_SERVE_LAUNCH(C, X)                                                                            \
  do {                                                                                                      \
    if (bt->long_lists)                                                                                     \
      hipLaunchKernelGGL((k_hbdia_batch<T, C, X, false, true, true>), grid, block, 0, s,                    \
                         reinterpret_cast<const int32_t*>(bt->tiles), nullptr,                              \
                         static_cast<const T*>(bt->values), bt->offsets, bt->coo_rows, bt->coo_cols,        \
                         static_cast<const T*>(bt->coo_values), total_tiles, tiles_per_step, jobs,          \
                         num_jobs, no_x, no_y);                                                             \
    else                                                                                                    \
      hipLaunchKernelGGL((k_hbdia_batch<T, C, X, false, true, false>), grid, block, 0, s,                   \
                         reinterpret_cast<const int32_t*>(bt->tiles), nullptr,                              \
                         static_cast<const T*>(bt->values), bt->offsets, bt->coo_rows, bt->coo_cols,        \
                         static_cast<const T*>(bt->coo_values), total_tiles, tiles_per_step, jobs,          \
                         num_jobs, no_x, no_y);                                                             \
  } while (0)
  if (bt->has_coo) {
    if (bt->x32) HBDIA_SERVE_LAUNCH(true, true); else HBDIA_SERVE_LAUNCH(true, false);
  } else {
    if (bt->x32) HBDIA_SERVE_LAUNCH(false, true); else HBDIA_SERVE_LAUNCH(false, false);
  }
#undef HBDIA_SERVE_LAUNCH
  return check_hip(hipGetLastError(), name);
}

template <typename T>
int batch_occupancy(bool coo, bool x32, bool contig, bool lng, int* out) {
  hipError_t e;
#define HBDIA_BATCH_OCC(C, X, G)                                                                                \
  do {                                                                                                          \
    if (lng && !G)                                                                                              \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_hbdia_batch<T, C, X, false, false, true>, kWG, 0); \
    else                                                                                                        \
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_hbdia_batch<T, C, X, G, false, false>, kWG, 0);    \
  } while (0)
  if (coo) {
    if (x32) {
      if (contig) HBDIA_BATCH_OCC(true, true, true); else HBDIA_BATCH_OCC(true, true, false);
    } else {
      if (contig) HBDIA_BATCH_OCC(true, false, true); else HBDIA_BATCH_OCC(true, false, false);
    }
  } else {
    if (x32) {
      if (contig) HBDIA_BATCH_OCC(false, true, true); else HBDIA_BATCH_OCC(false, true, false);
    } else {
      if (contig) HBDIA_BATCH_OCC(false, false, true); else HBDIA_BATCH_OCC(false, false, false);
    }
  }
#undef HBDIA_BATCH_OCC
  return check_hip(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
}

template <typename T>
int launch_coo(const int32_t* cr, const int32_t* cc, const T* cv, const int32_t* gptr, const T* x,
               T* res, int64_t nrows, hbdia_stream_t stream, const char* name) {
  if (nrows < 0) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": negative size");
  if (nrows == 0) return HBDIA_SUCCESS;
  if (!cr || !cc || !cv || !gptr || !x || !res)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null pointer");
  const int64_t groups = num_groups(nrows);
  const int64_t wgs = (groups + kWavesPerWG - 1) / kWavesPerWG;
  if (wgs > 0x7fffffffLL) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": too many rows");
  hipLaunchKernelGGL((k_coo<T>), dim3(static_cast<unsigned>(wgs)), dim3(kWG), 0, hs(stream), cr, cc, cv, gptr, x, res, nrows);
  return check_hip(hipGetLastError(), name);
}

template <typename T>
int launch_vector_add(T* dst, const T* src, int64_t n, hbdia_stream_t stream, const char* name) {
  if (n < 0) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": negative size");
  if (n == 0) return HBDIA_SUCCESS;
  if (!dst || !src) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null pointer");
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((k_vector_add<T>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, hs(stream), dst, src, n);
  return check_hip(hipGetLastError(), name);
}

}  // namespace

extern "C" {

const char* hbdia_last_error(void) { return g_last_error.c_str(); }

int hbdia_get_device_count(int* count) {
  if (!count) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_get_device_count: null");
  return check_hip(hipGetDeviceCount(count), "hipGetDeviceCount");
}
int hbdia_set_device(int device) { return check_hip(hipSetDevice(device), "hipSetDevice"); }
int hbdia_get_device(int* device) {
  if (!device) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_get_device: null");
  return check_hip(hipGetDevice(device), "hipGetDevice");
}
int hbdia_device_synchronize(void) { return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"); }
int hbdia_mem_info(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_mem_info: null");
  return check_hip(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}
int hbdia_device_name(char* buf, size_t len) {
  if (!buf || len == 0) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_device_name: null");
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  hipDeviceProp_t p;
  rc = check_hip(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
  if (rc) return rc;
  std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
  return HBDIA_SUCCESS;
}

int hbdia_device_compute_units(int* count) {
  if (!count) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_device_compute_units: null");
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  return check_hip(hipDeviceGetAttribute(count, hipDeviceAttributeMultiprocessorCount, dev),
                   "hipDeviceGetAttribute");
}

int hbdia_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_malloc: null");
  *ptr = nullptr;
  if (bytes == 0) return HBDIA_SUCCESS;
  return check_hip(hipMalloc(ptr, bytes), "hipMalloc");
}
int hbdia_free(void* ptr) { return ptr ? check_hip(hipFree(ptr), "hipFree") : HBDIA_SUCCESS; }
int hbdia_malloc_host(void** ptr, size_t bytes) {
  if (!ptr) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_malloc_host: null");
  *ptr = nullptr;
  if (bytes == 0) return HBDIA_SUCCESS;
  return check_hip(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
}
int hbdia_free_host(void* ptr) { return ptr ? check_hip(hipHostFree(ptr), "hipHostFree") : HBDIA_SUCCESS; }
int hbdia_malloc_managed(void** ptr, size_t bytes) {
  if (!ptr) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_malloc_managed: null");
  *ptr = nullptr;
  if (bytes == 0) return HBDIA_SUCCESS;
  return check_hip(hipMallocManaged(ptr, bytes, hipMemAttachGlobal), "hipMallocManaged");
}
int hbdia_memcpy_h2d(void* dst, const void* src, size_t bytes, hbdia_stream_t stream) {
  if (bytes == 0) return HBDIA_SUCCESS;
  if (!dst || !src) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_memcpy_h2d: null");
  return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, hs(stream)), "hipMemcpyAsync(H2D)");
}
int hbdia_memcpy_d2h(void* dst, const void* src, size_t bytes, hbdia_stream_t stream) {
  if (bytes == 0) return HBDIA_SUCCESS;
  if (!dst || !src) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_memcpy_d2h: null");
  return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, hs(stream)), "hipMemcpyAsync(D2H)");
}
int hbdia_memcpy_d2d(void* dst, const void* src, size_t bytes, hbdia_stream_t stream) {
  if (bytes == 0) return HBDIA_SUCCESS;
  if (!dst || !src) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_memcpy_d2d: null");
  return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, hs(stream)), "hipMemcpyAsync(D2D)");
}
int hbdia_memcpy(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return HBDIA_SUCCESS;
  if (!dst || !src) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_memcpy: null");
  return check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "hipMemcpy");
}
int hbdia_memset(void* dst, int value, size_t bytes, hbdia_stream_t stream) {
  if (bytes == 0) return HBDIA_SUCCESS;
  if (!dst) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_memset: null");
  return check_hip(hipMemsetAsync(dst, value, bytes, hs(stream)), "hipMemsetAsync");
}

int hbdia_stream_create(hbdia_stream_t* stream) {
  if (!stream) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_stream_create: null");
  hipStream_t s = nullptr;
  int rc = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  *stream = reinterpret_cast<hbdia_stream_t>(s);
  return rc;
}
int hbdia_stream_destroy(hbdia_stream_t stream) {
  return stream ? check_hip(hipStreamDestroy(hs(stream)), "hipStreamDestroy") : HBDIA_SUCCESS;
}
int hbdia_stream_synchronize(hbdia_stream_t stream) {
  return check_hip(hipStreamSynchronize(hs(stream)), "hipStreamSynchronize");
}
int hbdia_event_create(hbdia_event_t* event) {
  if (!event) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_event_create: null");
  hipEvent_t e = nullptr;
  int rc = check_hip(hipEventCreate(&e), "hipEventCreate");
  *event = reinterpret_cast<hbdia_event_t>(e);
  return rc;
}
int hbdia_event_destroy(hbdia_event_t event) {
  return event ? check_hip(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)), "hipEventDestroy")
               : HBDIA_SUCCESS;
}
int hbdia_event_record(hbdia_event_t event, hbdia_stream_t stream) {
  if (!event) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_event_record: null");
  return check_hip(hipEventRecord(reinterpret_cast<hipEvent_t>(event), hs(stream)), "hipEventRecord");
}
int hbdia_event_synchronize(hbdia_event_t event) {
  if (!event) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_event_synchronize: null");
  return check_hip(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)), "hipEventSynchronize");
}

int hbdia_event_query(hbdia_event_t event, int* done) {
  if (!event || !done) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_event_query: null");
  const hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(event));
  if (e == hipSuccess) {
    *done = 1;
    return HBDIA_SUCCESS;
  }
  if (e == hipErrorNotReady) {
    *done = 0;
    return HBDIA_SUCCESS;
  }
  return check_hip(e, "hipEventQuery");
}
int hbdia_stream_wait_event(hbdia_stream_t stream, hbdia_event_t event) {
  if (!event) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_stream_wait_event: null");
  return check_hip(hipStreamWaitEvent(hs(stream), reinterpret_cast<hipEvent_t>(event), 0),
                   "hipStreamWaitEvent");
}
int hbdia_event_elapsed_ms(float* ms, hbdia_event_t start, hbdia_event_t stop) {
  if (!ms || !start || !stop) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_event_elapsed_ms: null");
  return check_hip(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                       reinterpret_cast<hipEvent_t>(stop)),
                   "hipEventElapsedTime");
}
int hbdia_graph_begin_capture(hbdia_stream_t stream) {
  if (!stream) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_graph_begin_capture: null stream");
  return check_hip(hipStreamBeginCapture(hs(stream), hipStreamCaptureModeThreadLocal),
                   "hipStreamBeginCapture");
}
int hbdia_graph_end_capture(hbdia_stream_t stream, hbdia_graph_t* graph) {
  if (!stream || !graph) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_graph_end_capture: null");
  hipGraph_t g = nullptr;
  int rc = check_hip(hipStreamEndCapture(hs(stream), &g), "hipStreamEndCapture");
  if (rc) return rc;
  hipGraphExec_t ge = nullptr;
  rc = check_hip(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  *graph = reinterpret_cast<hbdia_graph_t>(ge);
  return rc;
}
int hbdia_graph_launch(hbdia_graph_t graph, hbdia_stream_t stream) {
  if (!graph) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_graph_launch: null");
  return check_hip(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph), hs(stream)), "hipGraphLaunch");
}
int hbdia_graph_destroy(hbdia_graph_t graph) {
  return graph ? check_hip(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)), "hipGraphExecDestroy")
               : HBDIA_SUCCESS;
}

int hbdia_bdia_spmv_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                        const double* x, double* y, int64_t num_rows, int32_t num_blocks,
                        int32_t block_width, hbdia_stream_t stream) {
  return launch_hbdia<double>(values, offsets, block_start, num_blocks, block_width, nullptr, nullptr,
                              nullptr, nullptr, x, y, num_rows, stream, false, "hbdia_bdia_spmv_f64");
}
int hbdia_bdia_spmv_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                        const float* x, float* y, int64_t num_rows, int32_t num_blocks,
                        int32_t block_width, hbdia_stream_t stream) {
  return launch_hbdia<float>(values, offsets, block_start, num_blocks, block_width, nullptr, nullptr,
                             nullptr, nullptr, x, y, num_rows, stream, false, "hbdia_bdia_spmv_f32");
}
int hbdia_hbdia_spmv_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                         int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                         const int32_t* coo_cols, const double* coo_values,
                         const int32_t* coo_group_ptr, const double* x, double* y,
                         int64_t num_rows, hbdia_stream_t stream) {
  return launch_hbdia<double>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols,
                              coo_values, coo_group_ptr, x, y, num_rows, stream, true,
                              "hbdia_hbdia_spmv_f64");
}
int hbdia_hbdia_spmv_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                         int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                         const int32_t* coo_cols, const float* coo_values,
                         const int32_t* coo_group_ptr, const float* x, float* y,
                         int64_t num_rows, hbdia_stream_t stream) {
  return launch_hbdia<float>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols,
                             coo_values, coo_group_ptr, x, y, num_rows, stream, true,
                             "hbdia_hbdia_spmv_f32");
}

int hbdia_bdia_spmv_ex_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                           const double* x, double* y, int64_t num_rows, int32_t num_blocks, int32_t block_width,
                           const uint32_t* runs3_blocks, hbdia_stream_t stream) {
  return launch_hbdia<double>(values, offsets, block_start, num_blocks, block_width, nullptr, nullptr, nullptr,
                              nullptr, x, y, num_rows, stream, false, "hbdia_bdia_spmv_ex_f64", runs3_blocks);
}
int hbdia_bdia_spmv_ex_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                           const float* x, float* y, int64_t num_rows, int32_t num_blocks, int32_t block_width,
                           const uint32_t* runs3_blocks, hbdia_stream_t stream) {
  return launch_hbdia<float>(values, offsets, block_start, num_blocks, block_width, nullptr, nullptr, nullptr,
                             nullptr, x, y, num_rows, stream, false, "hbdia_bdia_spmv_ex_f32", runs3_blocks);
}
int hbdia_hbdia_spmv_ex_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const double* coo_values, const int32_t* coo_group_ptr,
                            const double* x, double* y, int64_t num_rows, const uint32_t* runs3_blocks,
                            hbdia_stream_t stream) {
  return launch_hbdia<double>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols, coo_values,
                              coo_group_ptr, x, y, num_rows, stream, true, "hbdia_hbdia_spmv_ex_f64", runs3_blocks);
}
int hbdia_hbdia_spmv_ex_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const float* coo_values, const int32_t* coo_group_ptr,
                            const float* x, float* y, int64_t num_rows, const uint32_t* runs3_blocks,
                            hbdia_stream_t stream) {
  return launch_hbdia<float>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols, coo_values,
                             coo_group_ptr, x, y, num_rows, stream, true, "hbdia_hbdia_spmv_ex_f32", runs3_blocks);
}
int hbdia_hbdia_spmm_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                         int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                         const int32_t* coo_cols, const double* coo_values, const int32_t* coo_group_ptr,
                         const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t nrhs, int64_t num_rows,
                         hbdia_stream_t stream) {
  return launch_multi<double>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols, coo_values,
                              coo_group_ptr, X, ldx, Y, ldy, nrhs, num_rows, stream, "hbdia_hbdia_spmm_f64");
}
int hbdia_hbdia_spmm_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                         int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                         const int32_t* coo_cols, const float* coo_values, const int32_t* coo_group_ptr,
                         const float* X, int64_t ldx, float* Y, int64_t ldy, int32_t nrhs, int64_t num_rows,
                         hbdia_stream_t stream) {
  return launch_multi<float>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols, coo_values,
                             coo_group_ptr, X, ldx, Y, ldy, nrhs, num_rows, stream, "hbdia_hbdia_spmm_f32");
}
int hbdia_hbdia_spmm_ex_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const double* coo_values, const int32_t* coo_group_ptr,
                            const double* X, int64_t ldx, double* Y, int64_t ldy, int32_t nrhs, int64_t num_rows,
                            const uint32_t* runs3_blocks, hbdia_stream_t stream) {
  return launch_multi<double>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols, coo_values,
                              coo_group_ptr, X, ldx, Y, ldy, nrhs, num_rows, stream, "hbdia_hbdia_spmm_ex_f64",
                              runs3_blocks);
}
int hbdia_hbdia_spmm_ex_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const float* coo_values, const int32_t* coo_group_ptr,
                            const float* X, int64_t ldx, float* Y, int64_t ldy, int32_t nrhs, int64_t num_rows,
                            const uint32_t* runs3_blocks, hbdia_stream_t stream) {
  return launch_multi<float>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols, coo_values,
                             coo_group_ptr, X, ldx, Y, ldy, nrhs, num_rows, stream, "hbdia_hbdia_spmm_ex_f32",
                             runs3_blocks);
}

int hbdia_batch_make_tiles(const hbdia_batch_desc_t* desc, int32_t num_matrices, const int32_t* offsets,
                           const int32_t* block_start, const int32_t* group_ptr, int32_t value_size,
                           hbdia_batch_tile_t* tiles, int64_t num_tiles) {
  if (num_matrices < 0 || (num_matrices > 0 && (!desc || !offsets || !block_start || !tiles)) ||
      (value_size != 4 && value_size != 8))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_batch_make_tiles: invalid argument");
  constexpr int64_t kMax32 = 0x7fffffff;
  for (int32_t m = 0; m < num_matrices; ++m) {
    const hbdia_batch_desc_t& d = desc[m];
    if (d.num_rows < 0 || d.block_width <= 0 || d.block_width % kWave != 0)
      return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_batch_make_tiles: block width not a multiple of 64");
    if (d.coo_nnz > 0 && !group_ptr)
      return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_batch_make_tiles: fallback without group pointers");
    if (d.values % kWave != 0 || d.rows + d.num_rows > kMax32)
      return fail(HBDIA_ERROR_NOT_SUPPORTED, "hbdia_batch_make_tiles: pools beyond 32-bit tile fields");
    const int64_t groups = num_groups(d.num_rows);
    if (d.first_tile < 0 || d.first_tile + groups > num_tiles)
      return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_batch_make_tiles: tiles outside the tile array");
    const int32_t* bs = block_start + d.block_start;
    int32_t prev_at = -1, prev_cnt = -1;  // the previous tile's list (canonical index)
    for (int64_t g = 0; g < groups; ++g) {
      hbdia_batch_tile_t& tl = tiles[d.first_tile + g];
      const int64_t row0 = g * kWave;
      const int64_t b = row0 / d.block_width;
      int32_t start = 0, cnt = 0;
      if (b < d.num_blocks) {
        start = bs[b];
        cnt = bs[b + 1] - start;
      }
      if (static_cast<int64_t>(cnt) * d.block_width * value_size >= (int64_t{1} << 32))
        return fail(HBDIA_ERROR_NOT_SUPPORTED, "hbdia_batch_make_tiles: a block's values exceed 2^32 bytes");
      const int64_t v = cnt > 0 ? d.values + static_cast<int64_t>(start) * d.block_width + (row0 - b * d.block_width)
                                : 0;
      if (v / kWave > kMax32)
        return fail(HBDIA_ERROR_NOT_SUPPORTED, "hbdia_batch_make_tiles: values pool beyond 2^37 entries");
      tl.value_line = static_cast<int32_t>(v / kWave);
      tl.rows = static_cast<int32_t>(d.rows);
      tl.row0 = static_cast<int32_t>(row0);
      tl.num_rows = d.num_rows;
      tl.stride = d.block_width;
      tl.count = cnt;
      // equal lists of consecutive tiles get one index, so the kernel can tell
      // that two tiles share their offsets from the records alone
      int32_t at = d.offsets + start;
      if (cnt > 0 && cnt == prev_cnt && at != prev_at &&
          std::equal(offsets + at, offsets + at + cnt, offsets + prev_at))
        at = prev_at;
      tl.offsets = at;
      prev_at = at;
      prev_cnt = cnt;
      for (int k = 0; k < HBDIA_BATCH_INLINE_OFFSETS; ++k)
        tl.first_offsets[k] = k < cnt ? offsets[d.offsets + start + k] : 0;
      if (d.coo_nnz > 0) {
        const int32_t* gp = group_ptr + d.group_ptr;
        tl.coo_begin = d.coo + gp[g];
        tl.coo_end = d.coo + gp[g + 1];
        tl.fused = 1;
      } else {
        tl.coo_begin = tl.coo_end = 0;
        tl.fused = 0;
      }
    }
  }
  return HBDIA_SUCCESS;
}

int hbdia_batch_spmv_f64(const hbdia_batch_t* batch, const hbdia_batch_schedule_t* schedule, const double* x,
                         double* y, hbdia_stream_t stream) {
  return launch_batch<double>(batch, schedule, x, y, stream, "hbdia_batch_spmv_f64");
}

int hbdia_batch_spmv_f32(const hbdia_batch_t* batch, const hbdia_batch_schedule_t* schedule, const float* x,
                         float* y, hbdia_stream_t stream) {
  return launch_batch<float>(batch, schedule, x, y, stream, "hbdia_batch_spmv_f32");
}

int hbdia_serve_spmv_f64(const hbdia_batch_t* pool, const hbdia_serve_job_t* jobs, int32_t num_jobs,
                         int64_t total_tiles, int32_t num_workgroups, int32_t tiles_per_step, hbdia_stream_t stream) {
  return launch_serve<double>(pool, jobs, num_jobs, total_tiles, num_workgroups, tiles_per_step, stream,
                              "hbdia_serve_spmv_f64");
}

int hbdia_serve_spmv_f32(const hbdia_batch_t* pool, const hbdia_serve_job_t* jobs, int32_t num_jobs,
                         int64_t total_tiles, int32_t num_workgroups, int32_t tiles_per_step, hbdia_stream_t stream) {
  return launch_serve<float>(pool, jobs, num_jobs, total_tiles, num_workgroups, tiles_per_step, stream,
                             "hbdia_serve_spmv_f32");
}

int hbdia_batch_occupancy(const hbdia_batch_t* batch, int32_t kind, int32_t value_size, int* workgroups_per_cu) {
  if (!batch || !workgroups_per_cu || (value_size != 4 && value_size != 8) ||
      (kind != HBDIA_BATCH_INTERLEAVED && kind != HBDIA_BATCH_CONTIGUOUS && kind != HBDIA_BATCH_XCD))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_batch_occupancy: invalid argument");
  const bool coo = batch->has_coo != 0, x32 = batch->x32 != 0, contig = kind == HBDIA_BATCH_CONTIGUOUS;
  const bool lng = batch->long_lists != 0;
  return value_size == 8 ? batch_occupancy<double>(coo, x32, contig, lng, workgroups_per_cu)
                         : batch_occupancy<float>(coo, x32, contig, lng, workgroups_per_cu);
}
int hbdia_convert_values(const double* src, void* dst, int64_t n, int vdtype, hbdia_stream_t stream) {
  if (n < 0) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_convert_values: negative size");
  if (n == 0) return HBDIA_SUCCESS;
  if (!src || !dst) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_convert_values: null pointer");
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (vdtype == HBDIA_VALUES_F32)
    hipLaunchKernelGGL(k_f64_to_f32, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, hs(stream), src, static_cast<float*>(dst), n);
  else if (vdtype == HBDIA_VALUES_BF16)
    hipLaunchKernelGGL(k_f64_to_bf16, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, hs(stream), src, static_cast<uint16_t*>(dst), n);
  else
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_convert_values: vdtype must be HBDIA_VALUES_F32 or HBDIA_VALUES_BF16");
  return check_hip(hipGetLastError(), "hbdia_convert_values");
}
size_t hbdia_mixed_workspace_bytes(int64_t num_rows) {
  return static_cast<size_t>(num_groups(num_rows > 0 ? num_rows : 0) + 1) * sizeof(uint64_t);
}
int hbdia_hbdia_spmv_mixed(int vdtype, const void* values_lp, const double* values_f64, const int32_t* offsets,
                           const int32_t* block_start, int32_t num_blocks, int32_t block_width,
                           const int32_t* coo_rows, const int32_t* coo_cols, const void* coo_values_lp,
                           const double* coo_values_f64, const int32_t* coo_group_ptr, const double* x, double* y,
                           int64_t num_rows, double ratio, void* workspace, hbdia_stream_t stream) {
  return launch_mixed(vdtype, values_lp, values_f64, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols,
                      coo_values_lp, coo_values_f64, coo_group_ptr, x, y, num_rows, ratio, workspace, stream);
}
int hbdia_hbdia_spmv_mixed_ex(int vdtype, const void* values_lp, const double* values_f64, const int32_t* offsets,
                              const int32_t* block_start, int32_t num_blocks, int32_t block_width,
                              const int32_t* coo_rows, const int32_t* coo_cols, const void* coo_values_lp,
                              const double* coo_values_f64, const int32_t* coo_group_ptr, const double* x, double* y,
                              int64_t num_rows, double ratio, void* workspace, const uint32_t* runs3_blocks,
                              hbdia_stream_t stream) {
  return launch_mixed(vdtype, values_lp, values_f64, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols,
                      coo_values_lp, coo_values_f64, coo_group_ptr, x, y, num_rows, ratio, workspace, stream,
                      runs3_blocks);
}
}  // extern "C"

// Scratch of the allocation-free reduced-precision product: one 64-bit FP64
// mask per 64-row group (the workspace of hbdia_hbdia_spmv_mixed_ex), 2 MiB
// for up to HBDIA_NOALLOC_MAX_ROWS rows.  Each device holds its own copy for
// the life of the process.  In host code `scratch` names a host-side stand-in:
// its device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) uint64_t scratch[HBDIA_NOALLOC_MAX_ROWS / HBDIA_ROW_GROUP + 1];

namespace {

constexpr int kMaxScratchDevices = 256;
std::atomic<uint64_t*> g_scratch_dev[kMaxScratchDevices];
std::mutex g_scratch_mu;

// The current device's copy of `scratch`, resolved on the first call on that device.
int scratch_ptr(uint64_t** out, const char* name) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices)
    return fail(HBDIA_ERROR_DEVICE, std::string(name) + ": no current device");
  uint64_t* p = g_scratch_dev[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    p = g_scratch_dev[dev].load(std::memory_order_relaxed);
    if (!p) {
      void* addr = nullptr;
      const int rc = check_hip(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)), name);
      if (rc != HBDIA_SUCCESS) return rc;
      if (!addr) return fail(HBDIA_ERROR_DEVICE, std::string(name) + ": no device address for the scratch array");
      p = static_cast<uint64_t*>(addr);
      g_scratch_dev[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return HBDIA_SUCCESS;
}

}  // namespace

extern "C" {

int hbdia_hbdia_spmv_mixed_noalloc(int vdtype, const void* values_lp, const double* values_f64,
                                   const int32_t* offsets, const int32_t* block_start, int32_t num_blocks,
                                   int32_t block_width, const int32_t* coo_rows, const int32_t* coo_cols,
                                   const void* coo_values_lp, const double* coo_values_f64,
                                   const int32_t* coo_group_ptr, const double* x, double* y, int64_t num_rows,
                                   double ratio, const uint32_t* runs3_blocks, hbdia_stream_t stream) {
  const char* name = "hbdia_hbdia_spmv_mixed_noalloc";
  if (num_rows > HBDIA_NOALLOC_MAX_ROWS)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT,
                std::string(name) + ": num_rows exceeds HBDIA_NOALLOC_MAX_ROWS (" +
                    std::to_string(HBDIA_NOALLOC_MAX_ROWS) + "); use hbdia_hbdia_spmv_mixed_ex with a workspace");
  uint64_t* masks = nullptr;
  const int rc = scratch_ptr(&masks, name);
  if (rc != HBDIA_SUCCESS) return rc;
  // every launch rewrites the mask of every group it covers: nothing to clear
  return launch_mixed(vdtype, values_lp, values_f64, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols,
                      coo_values_lp, coo_values_f64, coo_group_ptr, x, y, num_rows, ratio, masks, stream,
                      runs3_blocks);
}

int hbdia_coo_spmv_f64(const int32_t* coo_rows, const int32_t* coo_cols, const double* coo_values,
                       const int32_t* coo_group_ptr, const double* x, double* res,
                       int64_t num_rows, hbdia_stream_t stream) {
  return launch_coo<double>(coo_rows, coo_cols, coo_values, coo_group_ptr, x, res, num_rows, stream,
                            "hbdia_coo_spmv_f64");
}
int hbdia_coo_spmv_f32(const int32_t* coo_rows, const int32_t* coo_cols, const float* coo_values,
                       const int32_t* coo_group_ptr, const float* x, float* res,
                       int64_t num_rows, hbdia_stream_t stream) {
  return launch_coo<float>(coo_rows, coo_cols, coo_values, coo_group_ptr, x, res, num_rows, stream,
                           "hbdia_coo_spmv_f32");
}
int hbdia_coo_group_ptr(const int32_t* coo_rows, int64_t coo_nnz, int64_t num_rows,
                        int32_t* group_ptr, hbdia_stream_t stream) {
  if (coo_nnz < 0 || num_rows < 0 || coo_nnz > 0x7fffffffLL)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_coo_group_ptr: bad size");
  if (!group_ptr || (coo_nnz > 0 && !coo_rows))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_coo_group_ptr: null pointer");
  const int64_t groups = num_groups(num_rows);
  const int64_t threads = groups + 1;
  const int64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL(k_group_ptr, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, hs(stream),
                     coo_rows, coo_nnz, groups, group_ptr);
  return check_hip(hipGetLastError(), "hbdia_coo_group_ptr");
}
int hbdia_vector_add_f64(double* dst, const double* src, int64_t n, hbdia_stream_t stream) {
  return launch_vector_add<double>(dst, src, n, stream, "hbdia_vector_add_f64");
}
int hbdia_vector_add_f32(float* dst, const float* src, int64_t n, hbdia_stream_t stream) {
  return launch_vector_add<float>(dst, src, n, stream, "hbdia_vector_add_f32");
}

}  // extern "C"
