// hbdia_convert.hip -- COO -> HBDIA conversion on the device (MI355X, gfx950).
//
// Produces exactly what the host conversion produces (reference
// src/Format/HBDIA.cpp:507-718), with sorts and scans in place of the
// reference's per-entry map inserts:
//   1. every entry is keyed by (block, offset); entries whose row lies past the
//      last block share one key that sorts after every block;
//   2. a stable radix sort of (key, entry index) makes the entries of each
//      (block, offset) pair one run, in input order, pairs ascending;
//   3. head flags and an inclusive scan give each sorted entry its run (pair)
//      number, and the run starts give the counts;
//   4. the threshold rule becomes a prefix condition.  The host visits pairs
//      in this order and keeps pair p iff count_p >= threshold or its running
//      fallback total has reached max_coo_entries, adding count_p to the total
//      only when p is not kept.  While that total is below the cap it equals
//      S_p, the sum of the counts of all earlier pairs below the threshold;
//      once it reaches the cap it stays there, and so does S_p.  Hence
//      kept_p = count_p >= threshold || S_p >= max_coo_entries, with S an
//      exclusive scan;
//   5. an exclusive scan of the kept flags ranks the kept pairs: the rank is
//      the pair's position in the flattened offsets and rank * block_width its
//      first value slot (blocks are stored in order, so this is also where the
//      host's running storage position puts it);
//   6. each entry of a kept pair claims its slot with atomicMax of its input
//      index -- the host assigns in input order, so the last entry wins -- and
//      a second pass replaces each claim by that entry's value (0 if none);
//   7. the other entries are compacted in input order (exclusive scan of
//      flags) into the fallback COO; when it is not sorted by row, a stable
//      radix sort by row gives the copy the fused SpMV kernel needs.
// All kernels are grid-stride loops over 256-thread workgroups; the data they
// move is integer or copied bit for bit, so the result is exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>  // before rocPRIM: its texture iterator calls memset unqualified
#include <mutex>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "hbdia_hip.h"
#include "hbdia_internal.hpp"

namespace {

using hbdia_detail::check_hip;
using hbdia_detail::fail;

constexpr int kThreads = 256;

unsigned grid_for(int64_t n) {
  const int64_t blocks = (n + kThreads - 1) / kThreads;
  return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks, int64_t(1) << 20)));
}

int bit_width(uint64_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

#define HBDIA_TRY(expr)                   \
  do {                                    \
    const int rc_ = (expr);               \
    if (rc_ != HBDIA_SUCCESS) return rc_; \
  } while (0)
#define HBDIA_TRY_HIP(call, what) HBDIA_TRY(check_hip((call), (what)))

#define HBDIA_GRID_STRIDE(i, n)                                                  \
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; \
       i < (n); i += static_cast<int64_t>(gridDim.x) * blockDim.x)

// key = (block << obits) | (offset + bias) for rows in a block, nb << obits else.
// Without a global row map offset + bias = col - row + num_rows - 1 lies in
// [0, num_rows + num_cols - 2]; with one the bias is 2^31 (sign flip).
__global__ void k_keys(const int* __restrict__ R, const int* __restrict__ C, const int* __restrict__ G,
                       int64_t n, int bw, int nb, int obits, uint32_t bias, uint64_t* __restrict__ keys,
                       uint32_t* __restrict__ idx) {
  HBDIA_GRID_STRIDE(i, n) {
    const int r = R[i];
    const int b = r / bw;
    uint64_t key = static_cast<uint64_t>(nb) << obits;
    if (b < nb) {
      const int o = C[i] - (G ? G[r] : r);
      key = (static_cast<uint64_t>(b) << obits) | static_cast<uint32_t>(static_cast<uint32_t>(o) + bias);
    }
    keys[i] = key;
    idx[i] = static_cast<uint32_t>(i);
  }
}

__global__ void k_heads(const uint64_t* __restrict__ ks, int64_t n, int* __restrict__ heads) {
  HBDIA_GRID_STRIDE(j, n) heads[j] = (j == 0 || ks[j] != ks[j - 1]) ? 1 : 0;
}

// run_start[run] = its first sorted position; rid is the inclusive scan of heads (1-based runs)
__global__ void k_run_starts(const int* __restrict__ heads, const int* __restrict__ rid, int64_t n,
                             int* __restrict__ run_start) {
  HBDIA_GRID_STRIDE(j, n) {
    if (heads[j]) run_start[rid[j] - 1] = static_cast<int>(j);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) run_start[rid[n - 1]] = static_cast<int>(n);
}

// counts of pairs below the threshold (0 for kept-for-sure pairs and the no-block run)
__global__ void k_pairs_below(const uint64_t* __restrict__ ks, const int* __restrict__ run_start, int64_t runs,
                              int obits, int nb, int thr, int64_t* __restrict__ below) {
  HBDIA_GRID_STRIDE(p, runs) {
    const int s = run_start[p];
    const int cnt = run_start[p + 1] - s;
    const int b = static_cast<int>(ks[s] >> obits);
    below[p] = (b < nb && cnt < thr) ? cnt : 0;
  }
}

__global__ void k_pairs_keep(const uint64_t* __restrict__ ks, const int* __restrict__ run_start,
                             const int64_t* __restrict__ S, int64_t runs, int obits, int nb, int thr,
                             int64_t max_coo, int* __restrict__ kept) {
  HBDIA_GRID_STRIDE(p, runs) {
    const int s = run_start[p];
    const int cnt = run_start[p + 1] - s;
    const int b = static_cast<int>(ks[s] >> obits);
    kept[p] = (b < nb && (cnt >= thr || S[p] >= max_coo)) ? 1 : 0;
  }
}

// kept pairs: their offset at their rank, per-block counts; every in-block pair's
// biased offset into ukeys (for the distinct-offset count); rank := -1 when dropped
__global__ void k_pairs_emit(const uint64_t* __restrict__ ks, const int* __restrict__ run_start,
                             const int* __restrict__ kept, int* __restrict__ rank, int64_t runs, int obits,
                             uint32_t bias, int nb, int* __restrict__ offsets, int* __restrict__ block_sizes,
                             int* __restrict__ block_placed, uint32_t* __restrict__ ukeys) {
  const uint64_t omask = (obits >= 64) ? ~uint64_t(0) : ((uint64_t(1) << obits) - 1);
  HBDIA_GRID_STRIDE(p, runs) {
    const int s = run_start[p];
    const int cnt = run_start[p + 1] - s;
    const uint64_t key = ks[s];
    const int b = static_cast<int>(key >> obits);
    const uint32_t u = static_cast<uint32_t>(key & omask);
    if (b < nb) ukeys[p] = u;
    if (kept[p]) {
      offsets[rank[p]] = static_cast<int>(u - bias);
      atomicAdd(&block_sizes[b], 1);
      atomicAdd(&block_placed[b], cnt);
    } else {
      rank[p] = -1;
    }
  }
}

// every value slot starts with the claim -1 in its low 32 bits
template <typename U>
__global__ void k_claims_init(U* __restrict__ values, int64_t slots) {
  HBDIA_GRID_STRIDE(s, slots) values[s] = static_cast<U>(0xffffffffu);
}

// entries of kept pairs claim their slot (largest input index wins); to_coo[i]
// flags the others, in input order
template <typename U>
__global__ void k_place(const uint32_t* __restrict__ idx, const int* __restrict__ rid,
                        const int* __restrict__ rank, const int* __restrict__ R, int64_t n, int bw,
                        U* __restrict__ values, int* __restrict__ to_coo) {
  constexpr int kWords = sizeof(U) / 4;
  HBDIA_GRID_STRIDE(j, n) {
    const uint32_t i = idx[j];
    const int rk = rank[rid[j] - 1];
    if (rk >= 0) {
      const int64_t slot = static_cast<int64_t>(rk) * bw + R[i] % bw;
      atomicMax(reinterpret_cast<int*>(values) + slot * kWords, static_cast<int>(i));
      to_coo[i] = 0;
    } else {
      to_coo[i] = 1;
    }
  }
}

template <typename U>
__global__ void k_claims_resolve(U* __restrict__ values, const U* __restrict__ V, int64_t slots) {
  HBDIA_GRID_STRIDE(s, slots) {
    const int w = static_cast<int>(static_cast<uint32_t>(values[s]));
    values[s] = w >= 0 ? V[w] : U(0);
  }
}

template <typename U>
__global__ void k_compact(const int* __restrict__ to_coo, const int* __restrict__ pos,
                          const int* __restrict__ R, const int* __restrict__ C, const U* __restrict__ V,
                          int64_t n, int* __restrict__ cr, int* __restrict__ cc, U* __restrict__ cv) {
  HBDIA_GRID_STRIDE(i, n) {
    if (to_coo[i]) {
      const int q = pos[i];
      cr[q] = R[i];
      cc[q] = C[i];
      cv[q] = V[i];
    }
  }
}

__global__ void k_descents(const int* __restrict__ r, int64_t n, int* __restrict__ flag) {
  HBDIA_GRID_STRIDE(q, n) {
    if (q > 0 && r[q] < r[q - 1]) *flag = 1;
  }
}

__global__ void k_iota(uint32_t* __restrict__ v, int64_t n) {
  HBDIA_GRID_STRIDE(i, n) v[i] = static_cast<uint32_t>(i);
}

template <typename U>
__global__ void k_gather(const uint32_t* __restrict__ perm, const int* __restrict__ cc, const U* __restrict__ cv,
                         int64_t n, int* __restrict__ sc, U* __restrict__ sv) {
  HBDIA_GRID_STRIDE(q, n) {
    const uint32_t k = perm[q];
    sc[q] = cc[k];
    sv[q] = cv[k];
  }
}

// One wave covers 64 consecutive blocks = two bitset words: its ballot is the
// pair of words.  Block b qualifies when it has 3k > 0 offsets and each triple
// is consecutive.
__global__ void k_runs3(const int* __restrict__ offsets, const int* __restrict__ block_start, int64_t nb,
                        uint32_t* __restrict__ bits, int* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  for (int64_t b0 = static_cast<int64_t>(blockIdx.x) * blockDim.x; b0 < nb;
       b0 += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t b = b0 + threadIdx.x;
    bool ok = false;
    if (b < nb) {
      const int s = block_start[b], e = block_start[b + 1];
      ok = e > s && (e - s) % 3 == 0;
      for (int t = s; ok && t + 2 < e; t += 3)
        ok = offsets[t + 1] == offsets[t] + 1 && offsets[t + 2] == offsets[t + 1] + 1;
    }
    const uint64_t m = __ballot(ok);
    if ((lane & 31) == 0 && b < nb) bits[b >> 5] = lane ? static_cast<uint32_t>(m >> 32) : static_cast<uint32_t>(m);
    if (lane == 0 && m) atomicAdd(count, static_cast<int>(__popcll(m)));
  }
}

// Temporaries come from a cache of device blocks per device that outlives
// the call: repeated conversions of similar sizes reuse the same blocks
// instead of paying hipMalloc / hipFree (and page mapping) for tens of GB each
// time; hbdia_convert_trim frees the cache.  A call takes blocks out of the
// cache, keeps them (also the ones it no longer needs) until its stream has
// finished, and then puts them all back.
constexpr int kMaxDevices = 64;
struct CachedBlock {
  void* p;
  size_t bytes;
};
std::mutex g_cache_mu;
std::vector<CachedBlock> g_cache[kMaxDevices];

// Smallest cached block of at least `bytes` and at most twice that plus 1 MiB.
bool cache_take(std::vector<CachedBlock>& pool, size_t bytes, CachedBlock* out) {
  size_t best = pool.size();
  for (size_t k = 0; k < pool.size(); ++k) {
    const size_t b = pool[k].bytes;
    if (b >= bytes && b <= 2 * bytes + (size_t(1) << 20) && (best == pool.size() || b < pool[best].bytes)) best = k;
  }
  if (best == pool.size()) return false;
  *out = pool[best];
  pool.erase(pool.begin() + static_cast<std::ptrdiff_t>(best));
  return true;
}

// Device allocations of one conversion: temporaries from the cache, returned
// on destruction after the stream has finished; outputs from hipMalloc, kept
// unless the conversion fails.
class Arena {
 public:
  Arena(int dev, hipStream_t s) : dev_(dev), s_(s) {}
  ~Arena() {
    (void)hipStreamSynchronize(s_);
    std::lock_guard<std::mutex> lock(g_cache_mu);
    for (const CachedBlock& b : live_) g_cache[dev_].push_back(b);
    for (const CachedBlock& b : spare_) g_cache[dev_].push_back(b);
  }
  template <typename P>
  int temp(P** p, size_t count, const char* what) {
    *p = nullptr;
    if (count == 0) return HBDIA_SUCCESS;
    const size_t bytes = count * sizeof(P);
    CachedBlock b{};
    bool hit = cache_take(spare_, bytes, &b);  // stream-ordered reuse within this call
    if (!hit) {
      std::lock_guard<std::mutex> lock(g_cache_mu);
      hit = cache_take(g_cache[dev_], bytes, &b);
    }
    if (!hit) {
      b.bytes = bytes;
      HBDIA_TRY_HIP(hipMalloc(&b.p, bytes), what);
    }
    live_.push_back(b);
    *p = static_cast<P*>(b.p);
    return HBDIA_SUCCESS;
  }
  template <typename P>
  int out(P** p, size_t count, const char* what) {
    *p = nullptr;
    if (count == 0) return HBDIA_SUCCESS;
    void* q = nullptr;
    HBDIA_TRY_HIP(hipMalloc(&q, count * sizeof(P)), what);
    *p = static_cast<P*>(q);
    return HBDIA_SUCCESS;
  }
  // p's contents are no longer needed; later temporaries of this call may reuse it
  void drop(void* p) {
    for (size_t k = 0; k < live_.size(); ++k)
      if (live_[k].p == p) {
        spare_.push_back(live_[k]);
        live_.erase(live_.begin() + static_cast<std::ptrdiff_t>(k));
        return;
      }
  }

 private:
  int dev_;
  hipStream_t s_;
  std::vector<CachedBlock> live_, spare_;
};

// rocPRIM's two-phase calls: query the temporary size, then run; one growing
// buffer of the arena serves them all.
class Scratch {
 public:
  explicit Scratch(Arena& ar) : ar_(ar) {}
  int ensure(size_t bytes) {
    if (bytes <= size_) return HBDIA_SUCCESS;
    if (p_) ar_.drop(p_);
    p_ = nullptr;
    size_ = 0;
    const size_t want = std::max<size_t>(bytes, 256);
    const int rc = ar_.temp(reinterpret_cast<unsigned char**>(&p_), want, "hbdia_convert_coo: temporary storage");
    if (rc == HBDIA_SUCCESS) size_ = want;
    return rc;
  }
  void* get() const { return p_; }

 private:
  Arena& ar_;
  void* p_ = nullptr;
  size_t size_ = 0;
};

template <typename F>
int prim(Scratch& sc, const char* what, F&& call) {
  size_t bytes = 0;
  HBDIA_TRY(check_hip(call(nullptr, bytes), what));
  HBDIA_TRY(sc.ensure(bytes));
  return check_hip(call(sc.get(), bytes), what);
}

template <typename V>
int fetch(V* host, const V* dev, hipStream_t s, const char* what) {
  HBDIA_TRY(check_hip(hipMemcpyAsync(host, dev, sizeof(V), hipMemcpyDeviceToHost, s), what));
  return check_hip(hipStreamSynchronize(s), what);
}

template <typename U>
int convert(const hbdia_convert_args_t& a, hbdia_convert_result_t& res, hipStream_t s) {
  int dev = 0;
  HBDIA_TRY_HIP(hipGetDevice(&dev), "hbdia_convert_coo: device");
  if (dev < 0 || dev >= kMaxDevices) return fail(HBDIA_ERROR_NOT_SUPPORTED, "hbdia_convert_coo: device index");
  Arena ar(dev, s);
  Scratch sc(ar);
  const int64_t n = a.nnz;
  const int bw = a.block_width;
  const int nb = static_cast<int>((static_cast<int64_t>(std::min(a.num_rows, a.num_cols)) + bw - 1) / bw);
  const bool global = a.global_rows != nullptr;
  const int obits = global ? 32 : std::max(1, bit_width(static_cast<uint64_t>(a.num_rows) + a.num_cols - 2));
  const uint32_t bias = global ? 0x80000000u : static_cast<uint32_t>(a.num_rows - 1);
  const int kbits = obits + bit_width(static_cast<uint64_t>(nb));
  const U* V = static_cast<const U*>(a.values);
  const unsigned gn = grid_for(n);
  res.num_blocks = nb;

  // 1-2. keys and the stable sort of (key, input index)
  uint64_t *ka = nullptr, *kb = nullptr;
  uint32_t *ia = nullptr, *ib = nullptr;
  HBDIA_TRY(ar.temp(&ka, n, "hbdia_convert_coo: keys"));
  HBDIA_TRY(ar.temp(&kb, n, "hbdia_convert_coo: keys"));
  HBDIA_TRY(ar.temp(&ia, n, "hbdia_convert_coo: indices"));
  HBDIA_TRY(ar.temp(&ib, n, "hbdia_convert_coo: indices"));
  hipLaunchKernelGGL(k_keys, dim3(gn), dim3(kThreads), 0, s, a.rows, a.cols, a.global_rows, n, bw, nb, obits,
                     bias, ka, ia);
  HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: keys"));
  rocprim::double_buffer<uint64_t> keys(ka, kb);
  rocprim::double_buffer<uint32_t> idx(ia, ib);
  HBDIA_TRY(prim(sc, "hbdia_convert_coo: sort by (block, offset)", [&](void* t, size_t& b) {
    return rocprim::radix_sort_pairs(t, b, keys, idx, static_cast<size_t>(n), 0u, static_cast<unsigned>(kbits), s);
  }));
  const uint64_t* ks = keys.current();
  const uint32_t* is = idx.current();
  // the alternate buffers are free from here on: heads and to_coo in the
  // index buffer, rid and pos in the two halves of the key buffer
  int* heads = reinterpret_cast<int*>(idx.alternate());
  int* rid = reinterpret_cast<int*>(keys.alternate());
  int* pos = rid + n;

  // 3. runs
  hipLaunchKernelGGL(k_heads, dim3(gn), dim3(kThreads), 0, s, ks, n, heads);
  HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: heads"));
  HBDIA_TRY(prim(sc, "hbdia_convert_coo: run numbers", [&](void* t, size_t& b) {
    return rocprim::inclusive_scan(t, b, heads, rid, static_cast<size_t>(n), rocprim::plus<int>(), s);
  }));
  int runs32 = 0;
  uint64_t last_key = 0;
  HBDIA_TRY(check_hip(hipMemcpyAsync(&last_key, ks + n - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s),
                      "hbdia_convert_coo: last key"));
  HBDIA_TRY(fetch(&runs32, rid + n - 1, s, "hbdia_convert_coo: run count"));
  const int64_t runs = runs32;
  const int64_t runs_blk = runs - ((last_key >> obits) >= static_cast<uint64_t>(nb) ? 1 : 0);
  int* run_start = nullptr;
  HBDIA_TRY(ar.temp(&run_start, runs + 1, "hbdia_convert_coo: run starts"));
  hipLaunchKernelGGL(k_run_starts, dim3(gn), dim3(kThreads), 0, s, heads, rid, n, run_start);
  HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: run starts"));

  // 4. the threshold rule
  const unsigned gr = grid_for(runs);
  int64_t *below = nullptr, *S = nullptr;
  int *kept = nullptr, *rank = nullptr;
  HBDIA_TRY(ar.temp(&below, runs, "hbdia_convert_coo: pair counts"));
  HBDIA_TRY(ar.temp(&S, runs, "hbdia_convert_coo: pair counts"));
  HBDIA_TRY(ar.temp(&kept, runs, "hbdia_convert_coo: kept flags"));
  HBDIA_TRY(ar.temp(&rank, runs, "hbdia_convert_coo: ranks"));
  hipLaunchKernelGGL(k_pairs_below, dim3(gr), dim3(kThreads), 0, s, ks, run_start, runs, obits, nb,
                     a.threshold, below);
  HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: pair counts"));
  HBDIA_TRY(prim(sc, "hbdia_convert_coo: fallback totals", [&](void* t, size_t& b) {
    return rocprim::exclusive_scan(t, b, below, S, int64_t(0), static_cast<size_t>(runs), rocprim::plus<int64_t>(),
                                   s);
  }));
  hipLaunchKernelGGL(k_pairs_keep, dim3(gr), dim3(kThreads), 0, s, ks, run_start, S, runs, obits, nb,
                     a.threshold, static_cast<int64_t>(a.max_coo_entries), kept);
  HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: kept flags"));
  ar.drop(below);
  ar.drop(S);

  // 5. ranks, offsets and block pointers
  HBDIA_TRY(prim(sc, "hbdia_convert_coo: ranks", [&](void* t, size_t& b) {
    return rocprim::exclusive_scan(t, b, kept, rank, 0, static_cast<size_t>(runs), rocprim::plus<int>(), s);
  }));
  int last_rank = 0, last_kept = 0;
  HBDIA_TRY(check_hip(hipMemcpyAsync(&last_rank, rank + runs - 1, sizeof(int), hipMemcpyDeviceToHost, s),
                      "hbdia_convert_coo: pair count"));
  HBDIA_TRY(fetch(&last_kept, kept + runs - 1, s, "hbdia_convert_coo: pair count"));
  const int64_t pairs = static_cast<int64_t>(last_rank) + last_kept;
  const int64_t slots = pairs * bw;
  if (slots > static_cast<int64_t>(1e11))
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_convert_coo: excessive memory requirement for HBDIA format");
  res.num_pairs = pairs;
  HBDIA_TRY(ar.out(&res.offsets, pairs + HBDIA_OFFSETS_PAD, "hbdia_convert_coo: offsets"));
  HBDIA_TRY(ar.out(&res.block_sizes, static_cast<size_t>(nb) + 1, "hbdia_convert_coo: block sizes"));
  HBDIA_TRY(ar.out(&res.block_start, static_cast<size_t>(nb) + 1, "hbdia_convert_coo: block starts"));
  HBDIA_TRY(ar.out(&res.block_placed, static_cast<size_t>(nb), "hbdia_convert_coo: block entries"));
  HBDIA_TRY(check_hip(hipMemsetAsync(res.offsets, 0, (pairs + HBDIA_OFFSETS_PAD) * sizeof(int), s),
                      "hbdia_convert_coo: offsets"));
  HBDIA_TRY(check_hip(hipMemsetAsync(res.block_sizes, 0, (static_cast<size_t>(nb) + 1) * sizeof(int), s),
                      "hbdia_convert_coo: block sizes"));
  if (nb > 0)
    HBDIA_TRY(check_hip(hipMemsetAsync(res.block_placed, 0, static_cast<size_t>(nb) * sizeof(int), s),
                        "hbdia_convert_coo: block entries"));
  uint32_t* ukeys = nullptr;
  HBDIA_TRY(ar.temp(&ukeys, std::max<int64_t>(runs_blk, 1), "hbdia_convert_coo: offsets of pairs"));
  hipLaunchKernelGGL(k_pairs_emit, dim3(gr), dim3(kThreads), 0, s, ks, run_start, kept, rank, runs, obits, bias,
                     nb, res.offsets, res.block_sizes, res.block_placed, ukeys);
  HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: pairs"));
  HBDIA_TRY(prim(sc, "hbdia_convert_coo: block starts", [&](void* t, size_t& b) {
    return rocprim::exclusive_scan(t, b, res.block_sizes, res.block_start, 0, static_cast<size_t>(nb) + 1,
                                   rocprim::plus<int>(), s);
  }));

  // 6. placement, last entry in input order wins
  int* to_coo = heads;  // heads are consumed by k_run_starts
  U* values = nullptr;
  HBDIA_TRY(ar.out(&values, static_cast<size_t>(slots), "hbdia_convert_coo: block storage"));
  res.values = values;
  if (slots > 0) {
    hipLaunchKernelGGL(k_claims_init<U>, dim3(grid_for(slots)), dim3(kThreads), 0, s, values, slots);
    HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: claims"));
  }
  hipLaunchKernelGGL(k_place<U>, dim3(gn), dim3(kThreads), 0, s, is, rid, rank, a.rows, n, bw, values, to_coo);
  HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: placement"));
  if (slots > 0) {
    hipLaunchKernelGGL(k_claims_resolve<U>, dim3(grid_for(slots)), dim3(kThreads), 0, s, values, V, slots);
    HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: values"));
  }

  // 7. fallback in input order, and its row-sorted copy
  HBDIA_TRY(prim(sc, "hbdia_convert_coo: fallback positions", [&](void* t, size_t& b) {
    return rocprim::exclusive_scan(t, b, to_coo, pos, 0, static_cast<size_t>(n), rocprim::plus<int>(), s);
  }));
  int last_pos = 0, last_flag = 0;
  HBDIA_TRY(check_hip(hipMemcpyAsync(&last_pos, pos + n - 1, sizeof(int), hipMemcpyDeviceToHost, s),
                      "hbdia_convert_coo: fallback count"));
  HBDIA_TRY(fetch(&last_flag, to_coo + n - 1, s, "hbdia_convert_coo: fallback count"));
  const int64_t ncoo = static_cast<int64_t>(last_pos) + last_flag;
  res.num_coo = ncoo;
  res.coo_sorted = 1;
  if (ncoo > 0) {
    U* cv = nullptr;
    HBDIA_TRY(ar.out(&res.coo_rows, ncoo, "hbdia_convert_coo: fallback"));
    HBDIA_TRY(ar.out(&res.coo_cols, ncoo, "hbdia_convert_coo: fallback"));
    HBDIA_TRY(ar.out(&cv, ncoo, "hbdia_convert_coo: fallback"));
    res.coo_values = cv;
    hipLaunchKernelGGL(k_compact<U>, dim3(gn), dim3(kThreads), 0, s, to_coo, pos, a.rows, a.cols, V, n,
                       res.coo_rows, res.coo_cols, cv);
    HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: fallback"));
    int* flag = nullptr;
    HBDIA_TRY(ar.temp(&flag, 1, "hbdia_convert_coo: order flag"));
    HBDIA_TRY(check_hip(hipMemsetAsync(flag, 0, sizeof(int), s), "hbdia_convert_coo: order flag"));
    hipLaunchKernelGGL(k_descents, dim3(grid_for(ncoo)), dim3(kThreads), 0, s, res.coo_rows, ncoo, flag);
    HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: order check"));
    int unsorted = 0;
    HBDIA_TRY(fetch(&unsorted, flag, s, "hbdia_convert_coo: order flag"));
    res.coo_sorted = unsorted ? 0 : 1;
    if (unsorted && a.want_row_sorted_coo) {
      uint32_t *perm_in = nullptr, *perm = nullptr;
      U* sv = nullptr;
      HBDIA_TRY(ar.temp(&perm_in, ncoo, "hbdia_convert_coo: fallback order"));
      HBDIA_TRY(ar.temp(&perm, ncoo, "hbdia_convert_coo: fallback order"));
      HBDIA_TRY(ar.out(&res.coo_rows_sorted, ncoo, "hbdia_convert_coo: sorted fallback"));
      HBDIA_TRY(ar.out(&res.coo_cols_sorted, ncoo, "hbdia_convert_coo: sorted fallback"));
      HBDIA_TRY(ar.out(&sv, ncoo, "hbdia_convert_coo: sorted fallback"));
      res.coo_values_sorted = sv;
      hipLaunchKernelGGL(k_iota, dim3(grid_for(ncoo)), dim3(kThreads), 0, s, perm_in, ncoo);
      HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: fallback order"));
      const uint32_t* rows_in = reinterpret_cast<const uint32_t*>(res.coo_rows);
      uint32_t* rows_out = reinterpret_cast<uint32_t*>(res.coo_rows_sorted);
      const unsigned rbits = static_cast<unsigned>(std::max(1, bit_width(static_cast<uint64_t>(a.num_rows))));
      HBDIA_TRY(prim(sc, "hbdia_convert_coo: sort fallback by row", [&](void* t, size_t& b) {
        return rocprim::radix_sort_pairs(t, b, rows_in, rows_out, perm_in, perm, static_cast<size_t>(ncoo), 0u,
                                         rbits, s);
      }));
      hipLaunchKernelGGL(k_gather<U>, dim3(grid_for(ncoo)), dim3(kThreads), 0, s, perm, res.coo_cols, cv, ncoo,
                         res.coo_cols_sorted, sv);
      HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: sorted fallback"));
    }
  }

  // distinct offsets over every in-block pair, kept or not
  if (runs_blk > 0) {
    uint32_t* usorted = nullptr;
    unsigned* ucount = nullptr;
    HBDIA_TRY(ar.temp(&usorted, runs_blk, "hbdia_convert_coo: distinct offsets"));
    HBDIA_TRY(ar.temp(&ucount, 1, "hbdia_convert_coo: distinct offsets"));
    HBDIA_TRY(prim(sc, "hbdia_convert_coo: sort offsets", [&](void* t, size_t& b) {
      return rocprim::radix_sort_keys(t, b, ukeys, usorted, static_cast<size_t>(runs_blk), 0u,
                                      static_cast<unsigned>(obits), s);
    }));
    HBDIA_TRY(prim(sc, "hbdia_convert_coo: distinct offsets", [&](void* t, size_t& b) {
      return rocprim::unique(t, b, usorted, ukeys, ucount, static_cast<size_t>(runs_blk),
                             rocprim::equal_to<uint32_t>(), s);
    }));
    unsigned distinct = 0;
    HBDIA_TRY(fetch(&distinct, ucount, s, "hbdia_convert_coo: distinct offsets"));
    res.number_diagonals = static_cast<int>(distinct);
  }

  // runs-of-three hint for the SpMV kernels
  if (a.want_runs3 && (bw % HBDIA_ROW_GROUP == 0 || bw == HBDIA_ROW_GROUP / 2) && nb > 0) {
    const int64_t words = (static_cast<int64_t>(nb) + 31) / 32;
    int* count = nullptr;
    HBDIA_TRY(ar.out(&res.runs3_blocks, words, "hbdia_convert_coo: runs-of-three bits"));
    HBDIA_TRY(ar.temp(&count, 1, "hbdia_convert_coo: runs-of-three count"));
    HBDIA_TRY(check_hip(hipMemsetAsync(res.runs3_blocks, 0, words * sizeof(uint32_t), s),
                        "hbdia_convert_coo: runs-of-three bits"));
    HBDIA_TRY(check_hip(hipMemsetAsync(count, 0, sizeof(int), s), "hbdia_convert_coo: runs-of-three count"));
    hipLaunchKernelGGL(k_runs3, dim3(grid_for(nb)), dim3(kThreads), 0, s, res.offsets, res.block_start,
                       static_cast<int64_t>(nb), res.runs3_blocks, count);
    HBDIA_TRY(check_hip(hipGetLastError(), "hbdia_convert_coo: runs-of-three"));
    int c = 0;
    HBDIA_TRY(fetch(&c, count, s, "hbdia_convert_coo: runs-of-three count"));
    res.runs3_count = c;
    if (c == 0) {
      HBDIA_TRY_HIP(hipFree(res.runs3_blocks), "hbdia_convert_coo: runs-of-three bits");
      res.runs3_blocks = nullptr;
    }
  }
  return check_hip(hipStreamSynchronize(s), "hbdia_convert_coo");
}

}  // namespace

extern "C" {

int hbdia_convert_release(hbdia_convert_result_t* res) {
  if (!res) return HBDIA_SUCCESS;
  void* arrays[] = {res->values,         res->offsets,     res->block_start,     res->block_sizes,
                    res->block_placed,   res->runs3_blocks, res->coo_rows,       res->coo_cols,
                    res->coo_values,     res->coo_rows_sorted, res->coo_cols_sorted, res->coo_values_sorted};
  int rc = HBDIA_SUCCESS;
  for (void* p : arrays)
    if (p && rc == HBDIA_SUCCESS) rc = check_hip(hipFree(p), "hbdia_convert_release");
    else if (p) (void)hipFree(p);
  std::memset(res, 0, sizeof(*res));
  return rc;
}

int hbdia_convert_trim(void) {
  int dev = 0;
  HBDIA_TRY_HIP(hipGetDevice(&dev), "hbdia_convert_trim");
  if (dev < 0 || dev >= kMaxDevices) return HBDIA_SUCCESS;
  std::vector<CachedBlock> blocks;
  {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    blocks.swap(g_cache[dev]);
  }
  int rc = HBDIA_SUCCESS;
  for (const CachedBlock& b : blocks) {
    const int r = check_hip(hipFree(b.p), "hbdia_convert_trim");
    if (rc == HBDIA_SUCCESS) rc = r;
  }
  return rc;
}

int hbdia_convert_coo(const hbdia_convert_args_t* args, hbdia_convert_result_t* out, hbdia_stream_t stream) {
  if (!args || !out) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_convert_coo: null argument");
  std::memset(out, 0, sizeof(*out));
  const hbdia_convert_args_t& a = *args;
  if (a.nnz <= 0 || a.nnz > 0x7fffffffLL)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_convert_coo: nnz must be in [1, 2^31 - 1]");
  if (a.num_rows <= 0 || a.num_cols <= 0 || a.block_width <= 0)
    return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_convert_coo: num_rows, num_cols and block_width must be positive");
  if (!a.rows || !a.cols || !a.values) return fail(HBDIA_ERROR_INVALID_ARGUMENT, "hbdia_convert_coo: null pointer");
  if (a.elem_size != 4 && a.elem_size != 8)
    return fail(HBDIA_ERROR_NOT_SUPPORTED, "hbdia_convert_coo: elem_size must be 4 or 8");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int rc = a.elem_size == 8 ? convert<uint64_t>(a, *out, s) : convert<uint32_t>(a, *out, s);
  if (rc != HBDIA_SUCCESS) {
    const std::string msg = hbdia_detail::g_last_error;
    (void)hipStreamSynchronize(s);
    (void)hbdia_convert_release(out);
    hbdia_detail::g_last_error = msg;
  }
  return rc;
}

}  // extern "C"
