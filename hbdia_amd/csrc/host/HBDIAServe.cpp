// HBDIAServe.cpp -- continuous-batching SpMV server and its open-loop driver.
#include "Operations/HBDIAServe.hpp"

#include <algorithm>
#include <chrono>
#include <random>
#include <stdexcept>
#include <string>

#include "types.hpp"

namespace hbdia {

VectorArena::VectorArena(int64_t capacity, int64_t align) : capacity_(capacity), align_(align) {
  if (capacity < 0 || align < 1) throw std::invalid_argument("VectorArena: capacity >= 0 and align >= 1");
  if (capacity_ > 0) free_[0] = capacity_;
}

int64_t VectorArena::allocate(int64_t n) {
  if (n <= 0) throw std::invalid_argument("VectorArena::allocate: size must be positive");
  const int64_t need = (n + align_ - 1) / align_ * align_;
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < need) continue;
    const int64_t off = it->first, size = it->second;
    free_.erase(it);
    if (size > need) free_[off + need] = size - need;
    taken_[off] = need;
    used_total_ += need;
    return off;
  }
  return -1;
}

void VectorArena::release(int64_t offset) {
  auto t = taken_.find(offset);
  if (t == taken_.end()) throw std::invalid_argument("VectorArena::release: not an allocated block");
  int64_t off = offset, size = t->second;
  taken_.erase(t);
  used_total_ -= size;
  auto next = free_.lower_bound(off);
  if (next != free_.end() && next->first == off + size) {  // merge with the following block
    size += next->second;
    next = free_.erase(next);
  }
  if (next != free_.begin()) {  // merge with the preceding block
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      prev->second += size;
      return;
    }
  }
  free_[off] = size;
}

int64_t serveArenaElems(int64_t budgetBytes, int64_t residentBytes, int64_t elemBytes) {
  if (elemBytes <= 0 || residentBytes < 0) throw std::invalid_argument("serveArenaElems: invalid sizes");
  if (budgetBytes <= residentBytes)
    throw std::invalid_argument("HBDIAServer: the HBM budget does not cover the resident matrices");
  const int64_t elems = std::min((budgetBytes - residentBytes) / (2 * elemBytes), kMaxServeArenaElems) / 64 * 64;
  if (elems < 64) throw std::invalid_argument("HBDIAServer: no room for request vectors in the budget");
  return elems;
}

int64_t VectorArena::largestFree() const {
  int64_t best = 0;
  for (const auto& f : free_) best = std::max(best, f.second);
  return best;
}

}  // namespace hbdia

namespace {

int64_t nowNs() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

template <typename T>
int serveLaunch(const hbdia_batch_t* pool, const hbdia_serve_job_t* jobs, int32_t n, int64_t tiles, int32_t wg,
                int32_t tps, hbdia_stream_t s);
template <>
int serveLaunch<double>(const hbdia_batch_t* pool, const hbdia_serve_job_t* jobs, int32_t n, int64_t tiles,
                        int32_t wg, int32_t tps, hbdia_stream_t s) {
  return hbdia_serve_spmv_f64(pool, jobs, n, tiles, wg, tps, s);
}
template <>
int serveLaunch<float>(const hbdia_batch_t* pool, const hbdia_serve_job_t* jobs, int32_t n, int64_t tiles, int32_t wg,
                       int32_t tps, hbdia_stream_t s) {
  return hbdia_serve_spmv_f32(pool, jobs, n, tiles, wg, tps, s);
}

uint64_t devAddr(const void* p) { return static_cast<uint64_t>(reinterpret_cast<uintptr_t>(p)); }

constexpr int kQueued = 0, kAdmitted = 1, kDone = 2, kFailed = 3;

}  // namespace

template <typename T>
HBDIAServer<T>::HBDIAServer(const std::vector<const HBDIA<T>*>& matrices, const ServeConfig& config) : cfg_(config) {
  if (matrices.empty()) throw std::invalid_argument("HBDIAServer: no matrices");
  if (cfg_.maxJobsPerLaunch < 1 || cfg_.maxLaunchesInFlight < 1 || cfg_.tilesPerStep < 1 || cfg_.workgroups < 0)
    throw std::invalid_argument("HBDIAServer: invalid configuration");
  if (!cfg_.batching) {
    cfg_.maxJobsPerLaunch = 1;
    cfg_.maxLaunchesInFlight = 1;
  }
  pool_ = std::make_unique<HBDIABatch<T>>(matrices);
  const int64_t resident = pool_->storedBytes();
  const int64_t budget = cfg_.hbmBudgetBytes > 0 ? cfg_.hbmBudgetBytes : resident + (int64_t{1} << 30);
  arenaElems_ = hbdia::serveArenaElems(budget, resident, static_cast<int64_t>(sizeof(T)));
  arena_ = hbdia::VectorArena(arenaElems_, 64);
  workgroups_ = cfg_.workgroups > 0 ? cfg_.workgroups : pool_->residentWorkgroups(HBDIA_BATCH_INTERLEAVED);
  try {
    void* p = nullptr;
    CHECK_HBDIA(hbdia_malloc(&p, static_cast<size_t>(arenaElems_) * sizeof(T)));
    xArena_ = static_cast<T*>(p);
    CHECK_HBDIA(hbdia_malloc(&p, static_cast<size_t>(arenaElems_) * sizeof(T)));
    yArena_ = static_cast<T*>(p);
    CHECK_HBDIA(hbdia_stream_create(&sIn_));
    CHECK_HBDIA(hbdia_stream_create(&sComp_));
    CHECK_HBDIA(hbdia_stream_create(&sOut_));
    const int rings = cfg_.maxLaunchesInFlight;
    evIn_.assign(rings, nullptr);
    evComp_.assign(rings, nullptr);

    hostJobs_.assign(rings, nullptr);
    devJobs_.assign(rings, nullptr);
    for (int r = 0; r < rings; ++r) {
      CHECK_HBDIA(hbdia_event_create(&evIn_[r]));
      CHECK_HBDIA(hbdia_event_create(&evComp_[r]));
      const size_t bytes = static_cast<size_t>(cfg_.maxJobsPerLaunch) * sizeof(hbdia_serve_job_t);
      CHECK_HBDIA(hbdia_malloc_host(&p, bytes));
      hostJobs_[r] = static_cast<hbdia_serve_job_t*>(p);
      CHECK_HBDIA(hbdia_malloc(&p, bytes));
      devJobs_[r] = static_cast<hbdia_serve_job_t*>(p);
      ringFree_.push_back(r);
    }
  } catch (...) {
    for (auto* q : hostJobs_) if (q) (void)hbdia_free_host(q);
    for (auto* q : devJobs_) if (q) (void)hbdia_free(q);
    for (auto* v : {&evIn_, &evComp_})
      for (auto e : *v) if (e) (void)hbdia_event_destroy(e);
    for (auto s : {sIn_, sComp_, sOut_}) if (s) (void)hbdia_stream_destroy(s);
    if (xArena_) (void)hbdia_free(xArena_);
    if (yArena_) (void)hbdia_free(yArena_);
    throw;
  }
  worker_ = std::thread([this] { loop(); });
}

template <typename T>
HBDIAServer<T>::~HBDIAServer() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cvSubmit_.notify_all();
  if (worker_.joinable()) worker_.join();
  for (auto s : {sIn_, sComp_, sOut_}) (void)hbdia_stream_synchronize(s);
  for (auto* q : hostJobs_) (void)hbdia_free_host(q);
  for (auto* q : devJobs_) (void)hbdia_free(q);
  for (auto& L : inflight_)
    for (size_t i = L.retired; i < L.done.size(); ++i)
      if (L.done[i] != evComp_[L.ring]) (void)hbdia_event_destroy(L.done[i]);
  for (auto* v : {&evIn_, &evComp_, &eventPool_})
    for (auto e : *v) (void)hbdia_event_destroy(e);
  for (auto s : {sIn_, sComp_, sOut_}) (void)hbdia_stream_destroy(s);
  (void)hbdia_free(xArena_);
  (void)hbdia_free(yArena_);
}

template <typename T>
int64_t HBDIAServer<T>::submit(int matrix, const T* x, T* y, bool deviceVectors) {
  if (matrix < 0 || matrix >= numMatrices()) throw std::out_of_range("HBDIAServer::submit: no such matrix");
  if (!x || !y) throw std::invalid_argument("HBDIAServer::submit: null vector");
  const int64_t n = pool_->numRows(matrix);
  if ((!deviceVectors || !cfg_.batching) && (n + 63) / 64 * 64 > arenaElems_)
    throw std::invalid_argument("HBDIAServer::submit: the matrix's vectors exceed the request arena");
  std::lock_guard<std::mutex> g(mu_);
  if (error_) std::rethrow_exception(error_);
  if (stop_) throw std::runtime_error("HBDIAServer::submit: server stopping");
  Request r;
  r.matrix = matrix;
  r.x = x;
  r.y = y;
  r.device = deviceVectors;
  r.t.submitNs = nowNs();
  const int64_t id = static_cast<int64_t>(requests_.size());
  requests_.push_back(r);
  pending_.push_back(id);
  cvSubmit_.notify_one();
  return id;
}

template <typename T>
void HBDIAServer<T>::wait(int64_t id) {
  std::unique_lock<std::mutex> g(mu_);
  if (id < 0 || id >= static_cast<int64_t>(requests_.size())) throw std::out_of_range("HBDIAServer::wait: no such id");
  cvDone_.wait(g, [&] { return requests_[id].state >= kDone; });
  if (requests_[id].state == kFailed) {
    if (error_) std::rethrow_exception(error_);
    throw std::runtime_error("HBDIAServer: request failed");
  }
}

template <typename T>
bool HBDIAServer<T>::done(int64_t id) const {
  std::lock_guard<std::mutex> g(mu_);
  return id >= 0 && id < static_cast<int64_t>(requests_.size()) && requests_[id].state >= kDone;
}

template <typename T>
RequestTimes HBDIAServer<T>::times(int64_t id) const {
  std::lock_guard<std::mutex> g(mu_);
  if (id < 0 || id >= static_cast<int64_t>(requests_.size())) throw std::out_of_range("HBDIAServer::times: no such id");
  return requests_[id].t;
}

template <typename T>
void HBDIAServer<T>::drain() {
  // admission is in arrival order and launches retire in issue order, so
  // requests complete in id order: the first `completed_` ids are done
  std::unique_lock<std::mutex> g(mu_);
  const int64_t count = static_cast<int64_t>(requests_.size());
  cvDone_.wait(g, [&] { return error_ || completed_ >= count; });
  if (error_) std::rethrow_exception(error_);
}

template <typename T>
int64_t HBDIAServer<T>::launches() const {
  std::lock_guard<std::mutex> g(mu_);
  return launches_;
}
template <typename T>
int64_t HBDIAServer<T>::completed() const {
  std::lock_guard<std::mutex> g(mu_);
  return completed_;
}
template <typename T>
int64_t HBDIAServer<T>::largestLaunch() const {
  std::lock_guard<std::mutex> g(mu_);
  return largest_;
}

template <typename T>
void HBDIAServer<T>::fail(std::exception_ptr e) {
  std::lock_guard<std::mutex> g(mu_);
  if (!error_) error_ = e;
  for (auto& r : requests_)
    if (r.state < kDone) r.state = kFailed;
  pending_.clear();
  cvDone_.notify_all();
}

template <typename T>
hbdia_event_t HBDIAServer<T>::takeEvent() {
  if (!eventPool_.empty()) {
    hbdia_event_t e = eventPool_.back();
    eventPool_.pop_back();
    return e;
  }
  hbdia_event_t e = nullptr;
  CHECK_HBDIA(hbdia_event_create(&e));
  return e;
}

template <typename T>
void HBDIAServer<T>::issue(std::vector<int64_t>& ids, int ring, std::vector<hbdia_event_t>& done) {
  // the job table and the copies in.  Batched, a request whose vectors are on
  // the device is read and written in place (no arena slot); a host request's
  // x goes into its arena slot.  The serve kernel writes every row of y, so
  // nothing is cleared.  The plain form copies every request through the
  // arena and clears y for the single-matrix kernel.
  hbdia_serve_job_t* hj = hostJobs_[ring];
  int64_t work = 0;
  std::vector<Request> rs;
  {
    std::lock_guard<std::mutex> g(mu_);
    for (int64_t id : ids) rs.push_back(requests_[id]);
  }
  for (size_t k = 0; k < rs.size(); ++k) {
    const Request& r = rs[k];
    hj[k].work_begin = work;
    hj[k].tile_begin = pool_->desc(r.matrix).first_tile;
    hj[k].x = r.slot >= 0 ? devAddr(xArena_ + r.slot) : devAddr(r.x);
    hj[k].y = r.slot >= 0 ? devAddr(yArena_ + r.slot) : devAddr(r.y);
    work += pool_->tilesOf(r.matrix);
  }
  const size_t T_ = sizeof(T);
  if (cfg_.batching)
    CHECK_HBDIA(hbdia_memcpy_h2d(devJobs_[ring], hj, rs.size() * sizeof(hbdia_serve_job_t), sIn_));
  for (const Request& r : rs) {
    if (r.slot < 0) continue;
    const size_t bytes = static_cast<size_t>(pool_->numRows(r.matrix)) * T_;
    if (r.device) CHECK_HBDIA(hbdia_memcpy_d2d(xArena_ + r.slot, r.x, bytes, sIn_));
    else CHECK_HBDIA(hbdia_memcpy_h2d(xArena_ + r.slot, r.x, bytes, sIn_));
    if (!cfg_.batching) CHECK_HBDIA(hbdia_memset(yArena_ + r.slot, 0, bytes, sIn_));
  }
  CHECK_HBDIA(hbdia_event_record(evIn_[ring], sIn_));
  CHECK_HBDIA(hbdia_stream_wait_event(sComp_, evIn_[ring]));
  if (cfg_.batching) {
    CHECK_HBDIA(serveLaunch<T>(&pool_->deviceBatch(), devJobs_[ring], static_cast<int32_t>(rs.size()), work,
                               workgroups_, cfg_.tilesPerStep, sComp_));
  } else {
    for (const Request& r : rs) pool_->spmvMatrix(r.matrix, xArena_ + r.slot, yArena_ + r.slot, sComp_);
  }
  CHECK_HBDIA(hbdia_event_record(evComp_[ring], sComp_));
  CHECK_HBDIA(hbdia_stream_wait_event(sOut_, evComp_[ring]));
  // each request completes when its own y is final: one copied out does not
  // wait for the copies of the requests after it in the launch, and one
  // written in place completes with the kernel
  done.clear();
  for (const Request& r : rs) {
    if (r.slot < 0) {
      done.push_back(evComp_[ring]);
      continue;
    }
    const size_t bytes = static_cast<size_t>(pool_->numRows(r.matrix)) * T_;
    if (r.device) CHECK_HBDIA(hbdia_memcpy_d2d(r.y, yArena_ + r.slot, bytes, sOut_));
    else CHECK_HBDIA(hbdia_memcpy_d2h(r.y, yArena_ + r.slot, bytes, sOut_));
    done.push_back(takeEvent());
    CHECK_HBDIA(hbdia_event_record(done.back(), sOut_));
  }
}

template <typename T>
void HBDIAServer<T>::loop() {
  std::vector<int64_t> next;  // admitted, waiting for a launch
  try {
    while (true) {
      bool progress = false;
      // retire finished launches, oldest first
      // (a launch's requests complete in order: their events are recorded in
      // order on one stream)
      while (!inflight_.empty()) {
        Launch& L = inflight_.front();
        bool any = false;
        while (L.retired < L.ids.size()) {
          int fin = 0;
          CHECK_HBDIA(hbdia_event_query(L.done[L.retired], &fin));
          if (!fin) break;
          const int64_t t = nowNs();
          {
            std::lock_guard<std::mutex> g(mu_);
            Request& r = requests_[L.ids[L.retired]];
            r.state = kDone;
            r.t.completeNs = t;
            if (r.slot >= 0) arena_.release(r.slot);
            ++completed_;
          }
          if (L.done[L.retired] != evComp_[L.ring]) eventPool_.push_back(L.done[L.retired]);
          ++L.retired;
          any = true;
        }
        if (any) {
          cvDone_.notify_all();
          progress = true;
        }
        if (L.retired < L.ids.size()) break;
        ringFree_.push_back(L.ring);
        inflight_.pop_front();
      }
      // admit queued requests in arrival order while the launch and the arena have room
      {
        std::lock_guard<std::mutex> g(mu_);
        while (!pending_.empty() && static_cast<int>(next.size()) < cfg_.maxJobsPerLaunch) {
          Request& r = requests_[pending_.front()];
          int64_t slot = -1;  // batched device requests need no arena slot
          if (!cfg_.batching || !r.device) {
            slot = arena_.allocate(pool_->numRows(r.matrix));
            if (slot < 0) break;  // strict order: later requests wait behind this one
          }
          r.slot = slot;
          r.state = kAdmitted;
          r.t.admitNs = nowNs();
          next.push_back(pending_.front());
          pending_.pop_front();
          progress = true;
        }
      }
      // launch them when a ring slot is free
      if (!next.empty() && !ringFree_.empty()) {
        const int ring = ringFree_.back();
        ringFree_.pop_back();
        Launch L;
        L.ring = ring;
        issue(next, ring, L.done);
        {
          std::lock_guard<std::mutex> g(mu_);
          ++launches_;
          largest_ = std::max<int64_t>(largest_, static_cast<int64_t>(next.size()));
        }
        L.ids = next;
        inflight_.push_back(std::move(L));
        next.clear();
        progress = true;
      }
      if (progress) continue;
      std::unique_lock<std::mutex> g(mu_);
      if (stop_ && pending_.empty() && next.empty() && inflight_.empty()) break;
      if (inflight_.empty() && next.empty()) {
        cvSubmit_.wait_for(g, std::chrono::milliseconds(2), [&] { return stop_ || !pending_.empty(); });
      } else {
        g.unlock();
        std::this_thread::yield();  // a launch is on the device: poll its event
      }
    }
  } catch (...) {
    fail(std::current_exception());
  }
}

template <typename T>
LoadReport runOpenLoop(HBDIAServer<T>& server, double rate, int64_t requests, uint64_t seed,
                       const std::vector<const T*>& xs, int variants, const std::vector<T*>& ys, bool deviceVectors) {
  const int M = server.numMatrices();
  if (variants < 1 || static_cast<int>(xs.size()) != M * variants || static_cast<int>(ys.size()) != M)
    throw std::invalid_argument("runOpenLoop: xs needs matrices * variants entries and ys one per matrix");
  if (requests < 0) throw std::invalid_argument("runOpenLoop: negative request count");
  std::mt19937_64 rng(seed);
  std::exponential_distribution<double> gap(rate > 0 ? rate : 1.0);
  std::uniform_int_distribution<int> pickM(0, M - 1), pickV(0, variants - 1);
  std::vector<int64_t> ids;
  ids.reserve(static_cast<size_t>(requests));
  std::vector<int64_t> due(static_cast<size_t>(requests));
  std::vector<int> mat(static_cast<size_t>(requests)), var(static_cast<size_t>(requests));
  double t = 0.0;
  for (int64_t i = 0; i < requests; ++i) {
    if (rate > 0) t += gap(rng);
    due[i] = static_cast<int64_t>(t * 1e9);
    mat[i] = pickM(rng);
    var[i] = pickV(rng);
  }
  const int64_t t0 = nowNs();
  for (int64_t i = 0; i < requests; ++i) {
    while (nowNs() - t0 < due[i]) std::this_thread::yield();
    ids.push_back(server.submit(mat[i], xs[static_cast<size_t>(mat[i]) * variants + var[i]], ys[mat[i]],
                                deviceVectors));
  }
  LoadReport rep;
  int64_t first = INT64_MAX, last = 0;
  for (int64_t id : ids) {
    server.wait(id);
    const RequestTimes tt = server.times(id);
    rep.latencyUs.push_back(static_cast<double>(tt.completeNs - tt.submitNs) * 1e-3);
    first = std::min(first, tt.submitNs);
    last = std::max(last, tt.completeNs);
  }
  rep.wallSeconds = ids.empty() ? 0.0 : static_cast<double>(last - first) * 1e-9;
  return rep;
}

template class HBDIAServer<double>;
template class HBDIAServer<float>;
template LoadReport runOpenLoop<double>(HBDIAServer<double>&, double, int64_t, uint64_t,
                                        const std::vector<const double*>&, int, const std::vector<double*>&, bool);
template LoadReport runOpenLoop<float>(HBDIAServer<float>&, double, int64_t, uint64_t, const std::vector<const float*>&,
                                       int, const std::vector<float*>&, bool);
