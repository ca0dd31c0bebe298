// HBDIAServe.hpp -- a continuous-batching SpMV server over matrices kept
// resident in HBM.
//
// No reference counterpart: the reference multiplies one matrix per
// hbdiaSpMV call (include/Operations/HBDIASpMV.hpp:8-21) and has no request
// loop.  Here a pool of converted matrices is uploaded once (an HBDIABatch
// pool); clients submit requests (matrix, x, y) from any thread.  A worker
// thread admits queued requests in arrival order while a launch may take more
// jobs and, for requests whose vectors are in host memory, the request arena
// (device memory for their x and y, the part of the HBM budget not used by the
// resident matrices) has room, and issues them as one persistent launch of the
// batch kernel over a job table.  A request whose vectors are already on the
// device is read and written in place: each job carries its own x and y, and
// the kernel writes every row of y, so nothing is copied or cleared for it.
// Requests that arrive while launches are on the GPU are admitted into the
// next launch, which is queued behind them; copies in (x), the kernel and
// copies out (y) run on three streams, so one launch's copies overlap
// another's kernel.  Every request's y equals, bit for bit, hbdiaSpMV on its
// matrix with y starting at zero, alone or in any mix.
//
// With batching off the server is the plain form: one request at a time, each
// through the single-matrix kernel, copies and kernel in sequence.
#ifndef HBDIA_SERVE_HPP
#define HBDIA_SERVE_HPP

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "HBDIABatch.hpp"

namespace hbdia {

// First-fit allocator over [0, capacity) units, every block aligned to
// `align` units; free blocks are kept coalesced.
class VectorArena {
 public:
  explicit VectorArena(int64_t capacity, int64_t align = 64);
  int64_t allocate(int64_t n);  // offset of a block of >= n units, or -1
  void release(int64_t offset);  // an offset returned by allocate
  int64_t capacity() const { return capacity_; }
  int64_t used() const { return used_total_; }
  int64_t largestFree() const;
  int64_t freeBlocks() const { return static_cast<int64_t>(free_.size()); }

 private:
  int64_t capacity_, align_, used_total_ = 0;
  std::map<int64_t, int64_t> free_;             // offset -> size
  std::unordered_map<int64_t, int64_t> taken_;  // offset -> size
};

// Elements of each of the server's x and y arenas for an HBM budget and the
// resident matrices' bytes: half of what remains per vector, a multiple of 64,
// and at most kMaxServeArenaElems (one allocation per vector of at most 16 GiB
// in FP64; a request's vectors are addressed through 64-bit job addresses, its
// rows as int32 within its matrix).  Throws when the budget leaves no room.
constexpr int64_t kMaxServeArenaElems = (int64_t{1} << 31) - 64;
int64_t serveArenaElems(int64_t budgetBytes, int64_t residentBytes, int64_t elemBytes);

}  // namespace hbdia

struct ServeConfig {
  int64_t hbmBudgetBytes = 0;   // resident matrices + request arena; 0: resident bytes + 1 GiB
  int maxJobsPerLaunch = 256;   // requests in one launch
  int maxLaunchesInFlight = 2;  // launches queued on the device at once
  int workgroups = 0;           // persistent grid; 0: every resident workgroup
  int tilesPerStep = 4;         // interleaved schedule run length
  bool batching = true;         // false: the plain form, one request at a time
};

struct RequestTimes {
  int64_t submitNs = 0, admitNs = 0, completeNs = 0;  // steady clock
};

template <typename T>
class HBDIAServer {
 public:
  HBDIAServer(const std::vector<const HBDIA<T>*>& matrices, const ServeConfig& config);
  ~HBDIAServer();  // completes every submitted request, then stops
  HBDIAServer(const HBDIAServer&) = delete;
  HBDIAServer& operator=(const HBDIAServer&) = delete;

  // y = A_matrix x.  x, y: host memory (pinned for overlapped copies) or, with
  // deviceVectors, device memory; both stay valid until the request
  // completes.  Device vectors are used in place by the batched server, so
  // requests in flight together must not share a y.  Throws if the matrix's
  // vectors can never fit the arena (host vectors, or the plain form).
  int64_t submit(int matrix, const T* x, T* y, bool deviceVectors = false);
  void wait(int64_t id);  // rethrows a failure of the worker
  bool done(int64_t id) const;
  RequestTimes times(int64_t id) const;
  void drain();  // waits for every request submitted so far

  int numMatrices() const { return pool_->numMatrices(); }
  int numRows(int m) const { return pool_->numRows(m); }
  int64_t residentBytes() const { return pool_->storedBytes(); }
  int64_t arenaBytes() const { return arenaElems_ * 2 * static_cast<int64_t>(sizeof(T)); }
  int64_t launches() const;
  int64_t completed() const;
  int64_t largestLaunch() const;  // most jobs in one launch so far
  const HBDIABatch<T>& pool() const { return *pool_; }

 private:
  struct Request {
    int matrix = 0;
    const T* x = nullptr;
    T* y = nullptr;
    bool device = false;
    int64_t slot = -1;
    int state = 0;  // 0 queued, 1 admitted, 2 done, 3 failed
    RequestTimes t;
  };
  struct Launch {
    std::vector<int64_t> ids;
    // per request: the event after which its y is final.  A request whose y is
    // copied out has its own (from eventPool_); one written in place shares the
    // launch's kernel event evComp_[ring].
    std::vector<hbdia_event_t> done;
    size_t retired = 0;               // requests of the launch already completed, in order
    int ring = 0;
  };

  void loop();
  void issue(std::vector<int64_t>& ids, int ring, std::vector<hbdia_event_t>& done);
  hbdia_event_t takeEvent();
  void fail(std::exception_ptr e);

  ServeConfig cfg_;
  std::unique_ptr<HBDIABatch<T>> pool_;
  int64_t arenaElems_ = 0;
  hbdia::VectorArena arena_{1};
  T* xArena_ = nullptr;
  T* yArena_ = nullptr;
  int workgroups_ = 1;
  hbdia_stream_t sIn_ = nullptr, sComp_ = nullptr, sOut_ = nullptr;
  std::vector<hbdia_event_t> evIn_, evComp_;
  std::vector<hbdia_event_t> eventPool_;  // per-request completion events, reused
  std::vector<hbdia_serve_job_t*> hostJobs_;  // pinned, one table per ring slot
  std::vector<hbdia_serve_job_t*> devJobs_;
  std::vector<int> ringFree_;

  mutable std::mutex mu_;
  std::condition_variable cvSubmit_, cvDone_;
  std::deque<Request> requests_;  // index = id
  std::deque<int64_t> pending_;
  std::deque<Launch> inflight_;
  int64_t launches_ = 0, completed_ = 0, largest_ = 0;
  bool stop_ = false;
  std::exception_ptr error_;
  std::thread worker_;
};

// Open-loop load: `requests` arrivals with exponential gaps of mean
// 1 / rate seconds (rate <= 0: all at once), each on a matrix drawn uniformly
// from the pool and one of its x variants (xs[m * variants + v]); results go
// to ys[m], shared by the requests of a matrix: a load generator for timing,
// whose outputs overlapping requests overwrite.  Latency of a request =
// completion - submission.  Returns per-request latencies (us) and the wall
// time from the first submission to the last completion.
struct LoadReport {
  std::vector<double> latencyUs;
  double wallSeconds = 0.0;
};

template <typename T>
LoadReport runOpenLoop(HBDIAServer<T>& server, double rate, int64_t requests, uint64_t seed,
                       const std::vector<const T*>& xs, int variants, const std::vector<T*>& ys, bool deviceVectors);

#endif  // HBDIA_SERVE_HPP
