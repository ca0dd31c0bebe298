// hbdia_spmm_rm.hip -- multiple right-hand sides with X and Y row-major
// (interleaved): the k values of a row are adjacent, X[row * k + j].
//
// No reference counterpart (the reference multiplies one vector per call);
// every column equals the single-vector hbdiaSpMV result bit for bit.
//
// Design (gfx950, HBM-bound):
//   * one workgroup per 64-row group (one block: block width % 64 == 0);
//     its waves split the row's 16-byte chunks of columns (2 FP64 / 4 FP32
//     columns each), lane = row, so a value a[row][o] is used for all of the
//     wave's columns from one register;
//   * the block's offsets are walked in runs of up to three consecutive
//     values (a 27-point stencil's runs); for a run o .. o + m - 1 the rows
//     [row0 + o, row0 + o + 63 + m) of X are one contiguous region (k values
//     per row), staged into LDS with direct global-to-LDS 16-byte loads, one
//     padding chunk per LDS row so that lane = row reads are bank-conflict
//     free; the run's values (64 rows per offset) are staged beside them;
//   * two LDS buffers: run s + 1 is staged while run s is summed, with one
//     workgroup barrier per run;
//   * per (row, column) the FMA chain runs over the offsets in ascending
//     order, skipping zero values and columns outside [0, rows), as the
//     single-vector kernels do; a run in which no lane skips takes plain
//     FMAs, others the masked form;
//   * the fallback (COO) entries of a row are summed per column in entry
//     order and added as y = (blocked sum or old y) + fallback, exactly the
//     single-vector fused kernel's rule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "hbdia_hip.h"
#include "hbdia_internal.hpp"

namespace {

using hbdia_detail::check_hip;
using hbdia_detail::fail;
using hbdia_detail::fma_t;
using hbdia_detail::masked_fma;
using hbdia_detail::xcd_chunk_shift;
using hbdia_detail::xcd_chunk_work;

constexpr int kWave = 64;
constexpr int kMaxRun = 3;                      // offsets staged together
constexpr int kWinRows = kWave + kMaxRun - 1;   // X rows of one staged window
constexpr int kMaxWaves = 4;

// LDS byte pointers in the LDS address space (what the direct loads take)
typedef __attribute__((address_space(3))) unsigned char lds_byte;

template <typename T> struct Chunk16;
template <> struct Chunk16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};
template <> struct Chunk16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};

// One 16-byte direct global-to-LDS load: lane l of the wave writes its 16
// bytes at dst + 16 l (dst wave-uniform).
__device__ __forceinline__ void load_lds16(const void* src, lds_byte* dst) {
  __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
}

// Bytes of one LDS buffer for C 16-byte chunks of X per row: the window
// (kWinRows rows of C + 1 chunks) and the run's values (kMaxRun x 64).
template <typename T>
__host__ __device__ constexpr int buffer_bytes(int C) {
  return kWinRows * (C + 1) * 16 + kMaxRun * kWave * static_cast<int>(sizeof(T));
}

// Where a thread stages: the window is copied rpi rows at a time (rpi =
// nthreads / (C + 1)), thread t taking row t / (C + 1) of the step and chunk
// t % (C + 1) of that row (the last chunk of an LDS row is padding: it is
// given a copy of the row's last chunk, never read); value chunk t (line
// t / VC, part t % VC) of the run's value lines.  Fixed per thread, so the
// staging loops do no division.
struct StageMap {
  int rpi;   // window rows per step
  int tr;    // the thread's row within a step
  int tcc;   // its chunk within the row (C: the padding chunk)
  int tc;    // the X chunk it copies
  bool on;   // the thread takes part in the window steps
  int nthreads;
};

// Stages one run: X rows [g0, g0 + 64 + m - 1) (clamped into [0, nrows): the
// products of clamped rows are masked) into buf as rows of C + 1 chunks, then
// the run's m value lines at a fixed place after the window, by direct
// global-to-LDS loads (no VGPRs): a wave's 64 lanes write 64 consecutive
// chunks, which the map makes contiguous.  Completion is this wave's vmcnt.
template <typename T>
__device__ __forceinline__ void stage_run(lds_byte* buf, const StageMap& sm, const T* __restrict__ X, int nrhs,
                                          int C, int64_t g0, int m, int64_t nrows, const T* __restrict__ vline,
                                          int bw, int tid) {
  constexpr int E = Chunk16<T>::n;
  constexpr int VC = kWave * static_cast<int>(sizeof(T)) / 16;  // chunks of one value line
  const int rows = kWave + m - 1;
  const int64_t last = nrows - 1;
  const int wbase = tid & ~(kWave - 1);  // the wave's first thread
  for (int w0 = 0; w0 < rows; w0 += sm.rpi) {
    const int w = w0 + sm.tr;
    int64_t g = g0 + w;
    g = g < 0 ? 0 : (g > last ? last : g);
    if (sm.on && w < rows) load_lds16(X + g * nrhs + sm.tc * E, buf + (w0 * (C + 1) + wbase) * 16);
  }
  lds_byte* vbuf = buf + kWinRows * (C + 1) * 16;
  for (int u0 = 0; u0 < m * VC; u0 += sm.nthreads) {
    const int u = u0 + tid;
    if (u < m * VC) {
      const int d = u / VC;
      load_lds16(vline + static_cast<int64_t>(d) * bw + (u - d * VC) * E, vbuf + (u0 + wbase) * 16);
    }
  }
}

template <typename T, int KPW, bool HAS_COO>
__global__ __launch_bounds__(kWave * kMaxWaves) void k_spmm_rm(
    const T* __restrict__ vals, const int* __restrict__ offs, const int* __restrict__ blk, const T* __restrict__ X,
    T* __restrict__ Y, int nrhs, int64_t nrows, int nb, int bw, const int* __restrict__ cr,
    const int* __restrict__ cc, const T* __restrict__ cv, const int* __restrict__ gptr, int kshift) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  lds_byte* lds = (lds_byte*)lds_raw;  // an address-space cast: the array is in LDS
  using V = typename Chunk16<T>::type;
  constexpr int E = Chunk16<T>::n;
  const int C = nrhs / E;           // 16-byte chunks per row
  const int RS = (C + 1) * 16;      // LDS row stride
  const int bufb = buffer_bytes<T>(C);
  const int tid = threadIdx.x, nthreads = blockDim.x;
  const int lane = tid & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const unsigned work = xcd_chunk_work(blockIdx.x, gridDim.x, kshift);
  const int64_t row0 = static_cast<int64_t>(work) * kWave;  // the same in every wave: uniform exits
  if (row0 >= nrows) return;
  const int64_t row = row0 + lane;
  const bool active = row < nrows;
  const int64_t b = row0 / bw;
  int start = 0, cnt = 0;
  if (b < nb) {
    start = blk[b];
    cnt = blk[b + 1] - start;
  }
  StageMap sm;
  sm.rpi = nthreads / (C + 1);
  sm.tr = tid / (C + 1);
  sm.tcc = tid - sm.tr * (C + 1);
  sm.tc = sm.tcc < C ? sm.tcc : C - 1;
  sm.on = sm.tr < sm.rpi;
  sm.nthreads = nthreads;
  const int c0 = wv * KPW;  // the wave's first chunk
  V acc[KPW];
#pragma unroll
  for (int q = 0; q < KPW; ++q) acc[q] = V{};
  const T* vline0 = vals + static_cast<int64_t>(start) * bw + (row0 - b * bw);

  if (cnt > 0) {
    const int* op = offs + start;  // the entry point requires padding behind the offsets
    // run descriptor: first offset index i, length m (consecutive offsets, at most 3)
    auto run_len = [&](int i) {
      int m = 1;
      if (i + 1 < cnt && op[i + 1] == op[i] + 1) {
        m = 2;
        if (i + 2 < cnt && op[i + 2] == op[i + 1] + 1) m = 3;
      }
      return m;
    };
    int i = 0, m = run_len(0);
    stage_run<T>(lds, sm, X, nrhs, C, row0 + op[0], m, nrows, vline0, bw, tid);
    for (int s = 0;; ++s) {
      // two buffers: run s + 1 is loaded into the other while run s is summed
      lds_byte* cur = lds + (s & 1) * bufb;
      lds_byte* nxt = lds + ((s + 1) & 1) * bufb;
      __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's direct loads of run s
      __syncthreads();  // run s staged by every wave; run s - 1's buffer free
      const int o = op[i];
      const int in = i + m;
      int mn = 0;
      if (in < cnt) {
        mn = run_len(in);
        stage_run<T>(nxt, sm, X, nrhs, C, row0 + op[in], mn, nrows, vline0 + static_cast<int64_t>(in) * bw,
                           bw, tid);
      }
      // run s: offsets o .. o + m - 1, window row of (lane, d) = lane + d
      const __attribute__((address_space(3))) T* sv =
          reinterpret_cast<const __attribute__((address_space(3))) T*>(cur + kWinRows * RS);
      T v[kMaxRun];
      bool ok[kMaxRun];
      bool all = true;
#pragma unroll
      for (int d = 0; d < kMaxRun; ++d) {
        v[d] = d < m ? sv[d * kWave + lane] : T(0);
        const int64_t c = row + o + d;
        ok[d] = d < m && active && v[d] != T(0) && c >= 0 && c < nrows;
        all = all && (d >= m || ok[d]);
      }
      const bool fast = __builtin_amdgcn_ballot_w64(!all) == 0;
      const lds_byte* xb = cur + lane * RS + c0 * 16;
      if (fast) {
#pragma unroll
        for (int q = 0; q < KPW; ++q) {
          if (c0 + q >= C) continue;
#pragma unroll
          for (int d = 0; d < kMaxRun; ++d) {
            if (d >= m) continue;
            const V xv = *reinterpret_cast<const __attribute__((address_space(3))) V*>(xb + d * RS + q * 16);
#pragma unroll
            for (int e = 0; e < E; ++e) acc[q][e] = fma_t<T>(v[d], xv[e], acc[q][e]);
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < KPW; ++q) {
          if (c0 + q >= C) continue;
#pragma unroll
          for (int d = 0; d < kMaxRun; ++d) {
            if (d >= m) continue;
            const V xv = *reinterpret_cast<const __attribute__((address_space(3))) V*>(xb + d * RS + q * 16);
#pragma unroll
            for (int e = 0; e < E; ++e) acc[q][e] = masked_fma(ok[d], v[d], xv[e], acc[q][e]);
          }
        }
      }
      if (in >= cnt) break;
      i = in;
      m = mn;
    }
  }

  if (!active) return;
  T* yr = Y + row * nrhs;
  if constexpr (HAS_COO) {
    // the row's fallback entries: a lower bound in the group's sorted rows
    const int p0 = gptr[work], p1 = gptr[work + 1];
    V cs[KPW];
#pragma unroll
    for (int q = 0; q < KPW; ++q) cs[q] = V{};
    if (p0 < p1) {
      int lo = p0, hi = p1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cr[mid] < row) lo = mid + 1;
        else hi = mid;
      }
      for (int e0 = lo; e0 < p1 && cr[e0] == row; ++e0) {
        const T a = cv[e0];
        const T* xr = X + static_cast<int64_t>(cc[e0]) * nrhs;
#pragma unroll
        for (int q = 0; q < KPW; ++q) {
          if (c0 + q >= C) continue;
          const V xv = *reinterpret_cast<const V*>(xr + (c0 + q) * E);
#pragma unroll
          for (int e = 0; e < E; ++e) cs[q][e] = fma_t<T>(a, xv[e], cs[q][e]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < KPW; ++q) {
      if (c0 + q >= C) continue;
      V* yp = reinterpret_cast<V*>(yr + (c0 + q) * E);
      const V base = cnt > 0 ? acc[q] : *yp;
      __builtin_nontemporal_store(base + cs[q], yp);
    }
  } else {
    if (cnt == 0) return;
#pragma unroll
    for (int q = 0; q < KPW; ++q) {
      if (c0 + q >= C) continue;
      __builtin_nontemporal_store(acc[q], reinterpret_cast<V*>(yr + (c0 + q) * E));
    }
  }
}

template <typename T, int KPW>
void launch_kpw(const T* vals, const int32_t* offs, const int32_t* blk, int32_t nb, int32_t bw, const int32_t* cr,
                const int32_t* cc, const T* cv, const int32_t* gptr, const T* X, T* Y, int32_t nrhs, int64_t nrows,
                int waves, size_t lds_bytes, hipStream_t s) {
  const int64_t wgs = (nrows + kWave - 1) / kWave;
  const int ks = xcd_chunk_shift(wgs);
  const dim3 grid(static_cast<unsigned>(wgs)), block(kWave * waves);
  if (gptr)
    hipLaunchKernelGGL((k_spmm_rm<T, KPW, true>), grid, block, lds_bytes, s, vals, offs, blk, X, Y, nrhs,
                       nrows, nb, bw, cr, cc, cv, gptr, ks);
  else
    hipLaunchKernelGGL((k_spmm_rm<T, KPW, false>), grid, block, lds_bytes, s, vals, offs, blk, X, Y, nrhs,
                       nrows, nb, bw, cr, cc, cv, gptr, ks);
}

template <typename T>
int launch_spmm_rm(const T* vals, const int32_t* offs, const int32_t* blk, int32_t nb, int32_t bw,
                   const int32_t* cr, const int32_t* cc, const T* cv, const int32_t* gptr, const T* X, T* Y,
                   int32_t nrhs, int64_t nrows, hbdia_stream_t stream, const char* name) {
  constexpr int E = Chunk16<T>::n;
  if (nrows < 0 || nb < 0 || bw <= 0 || nrhs < 0)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": negative size or block_width <= 0");
  if (nrows == 0 || nrhs == 0) return HBDIA_SUCCESS;
  if (bw % kWave != 0)
    return fail(HBDIA_ERROR_NOT_SUPPORTED, std::string(name) + ": block_width must be a multiple of 64");
  if (nrhs % E != 0 || nrhs / E > HBDIA_SPMM_RM_MAX_CHUNKS)
    return fail(HBDIA_ERROR_NOT_SUPPORTED, std::string(name) + ": nrhs * sizeof(value) must be a multiple of 16 "
                                                               "bytes and at most 512 bytes");
  if (!X || !Y || !blk || (nb > 0 && (!offs || !vals)))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null pointer");
  const bool with_coo = gptr != nullptr;
  if (with_coo && (!cr || !cc || !cv))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": null COO pointer");
  if (nrows > 0x7fffffffLL) return fail(HBDIA_ERROR_INVALID_ARGUMENT, std::string(name) + ": too many rows");
  const int C = nrhs / E;
  // chunks per wave: a power of two, at most kMaxWaves waves (C <= 32: at most 8)
  int kpw = 1;
  while (kpw * kMaxWaves < C) kpw *= 2;
  const int waves = (C + kpw - 1) / kpw;
  const size_t lds_bytes = 2 * static_cast<size_t>(buffer_bytes<T>(C));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  switch (kpw) {
    case 1: launch_kpw<T, 1>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, Y, nrhs, nrows, waves, lds_bytes, s); break;
    case 2: launch_kpw<T, 2>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, Y, nrhs, nrows, waves, lds_bytes, s); break;
    case 4: launch_kpw<T, 4>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, Y, nrhs, nrows, waves, lds_bytes, s); break;
    default: launch_kpw<T, 8>(vals, offs, blk, nb, bw, cr, cc, cv, gptr, X, Y, nrhs, nrows, waves, lds_bytes, s); break;
  }
  return check_hip(hipGetLastError(), name);
}

}  // namespace

extern "C" {

int hbdia_hbdia_spmm_rm_f64(const double* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const double* coo_values, const int32_t* coo_group_ptr,
                            const double* X, double* Y, int32_t nrhs, int64_t num_rows, hbdia_stream_t stream) {
  return launch_spmm_rm<double>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols, coo_values,
                                coo_group_ptr, X, Y, nrhs, num_rows, stream, "hbdia_hbdia_spmm_rm_f64");
}

int hbdia_hbdia_spmm_rm_f32(const float* values, const int32_t* offsets, const int32_t* block_start,
                            int32_t num_blocks, int32_t block_width, const int32_t* coo_rows,
                            const int32_t* coo_cols, const float* coo_values, const int32_t* coo_group_ptr,
                            const float* X, float* Y, int32_t nrhs, int64_t num_rows, hbdia_stream_t stream) {
  return launch_spmm_rm<float>(values, offsets, block_start, num_blocks, block_width, coo_rows, coo_cols, coo_values,
                               coo_group_ptr, X, Y, nrhs, num_rows, stream, "hbdia_hbdia_spmm_rm_f32");
}

}  // extern "C"
