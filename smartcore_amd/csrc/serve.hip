// kNN serving loop with continuous batching (sc_server_*).
//
// smartcore has no serving component; this runtime keeps one index resident
// in HBM and answers LinearKNNSearch::find requests (linear_search.rs:52-84)
// that arrive asynchronously, each carrying 1..many query rows.
//
//   submit     copies the request's rows to host memory and queues it; it
//              never blocks on the device.
//   admission  a request is admitted only while the HBM its batch will need
//              (its rows, outputs and the search workspace, estimated up
//              front) fits the server's budget; others wait in FIFO order.
//              A request larger than the whole budget is refused.
//   batching   the scheduler thread groups admitted requests with the same
//              (metric, p, mode, k) into one device batch of up to
//              max_batch_queries rows as soon as one of max_inflight slots
//              is free, so requests that arrive while batches are running
//              join the next batch (continuous batching).
//   overlap    every slot owns a stream (default priority), pinned staging
//              and device buffers; batches on different slots run
//              concurrently on the device.
//   completion a completion thread waits for each slot's event, scatters the
//              rows back to their requests, records latency, frees the slot
//              and the budget, and wakes waiters.
// A request's result does not depend on which other requests share its
// batch: every query's search is independent and exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"

namespace sc {
int search_batch_device(const sc_index* index, const void* d_padded_queries, size_t q, size_t k,
                        sc_metric metric, uint16_t p, sc_mode mode, int64_t* d_idx,
                        double* d_dist, uint32_t* d_count, hipStream_t s);
}

using Clock = std::chrono::steady_clock;
using namespace sc;

namespace {

struct Request {
  uint64_t id = 0;
  std::vector<char> rows;  // q x d of the index dtype
  size_t q = 0, k = 0;
  sc_metric metric = SC_EUCLIDIAN;
  uint16_t p = 2;
  sc_mode mode = SC_PLAIN_FP64;
  size_t bytes = 0;  // admission estimate
  Clock::time_point submitted;
  Clock::time_point finished;
  bool admitted = false;
  bool done = false;
  int status = SC_OK;
  std::string error;
  std::vector<int64_t> idx;
  std::vector<double> dist;
  std::vector<uint32_t> count;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  size_t cap_q = 0, cap_k = 0;
  void* h_q = nullptr;      // pinned staging
  int64_t* h_idx = nullptr;
  double* h_dist = nullptr;
  uint32_t* h_cnt = nullptr;
  void* d_q = nullptr;
  int64_t* d_idx = nullptr;
  double* d_dist = nullptr;
  uint32_t* d_cnt = nullptr;
  std::vector<std::shared_ptr<Request>> reqs;
  size_t batch_q = 0, batch_k = 0;
  size_t bytes = 0;
  int status = SC_OK;
  std::string error;
};

}  // namespace

struct sc_server {
  const sc_index* index = nullptr;
  sc_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_sched;   // work or a free slot
  std::condition_variable cv_done;    // a request finished
  std::condition_variable cv_slot;    // completion thread: a slot was launched
  std::deque<std::shared_ptr<Request>> pending;
  std::map<uint64_t, std::shared_ptr<Request>> all;
  std::vector<Slot> slots;
  std::deque<int> launched;  // slot ids in launch order
  size_t budget = 0, in_use = 0, peak = 0;
  uint64_t next_id = 1;
  uint64_t completed = 0, batches = 0, refused = 0, queries_done = 0;
  bool stop = false;
  std::thread sched, compl_;
};

namespace {

size_t estimate_bytes(const sc_index* ix, size_t q, size_t k, sc_mode mode) {
  const size_t es = elem_size(ix->dtype);
  size_t b = q * ix->ld * es + q * k * (sizeof(int64_t) + sizeof(double)) + q * sizeof(uint32_t);
  // search workspace: candidate lists (~ nseg x cap entries per query) and
  // the mixed form's BF16 query copy
  const double L = static_cast<double>(std::max<size_t>(ix->n, 1));
  const size_t cap = static_cast<size_t>(4.0 * k * (1.0 + std::log(std::max(1.0, L / k)))) + 128;
  b += q * 8 * cap * 12;
  if (mode == SC_MIXED) b += q * (ix->ld + 16) * 2 + q * 16;
  return b;
}

int slot_reserve(Slot& s, const sc_index* ix, size_t q, size_t k) {
  if (q <= s.cap_q && k <= s.cap_k) return SC_OK;
  const size_t nq = std::max(q, s.cap_q), nk = std::max(k, s.cap_k);
  const size_t es = elem_size(ix->dtype);
  if (s.h_q) {
    (void)hipHostFree(s.h_q);
    (void)hipHostFree(s.h_idx);
    (void)hipHostFree(s.h_dist);
    (void)hipHostFree(s.h_cnt);
    (void)hipFree(s.d_q);
    (void)hipFree(s.d_idx);
    (void)hipFree(s.d_dist);
    (void)hipFree(s.d_cnt);
  }
  SC_HIP(hipHostMalloc(&s.h_q, nq * ix->ld * es));
  SC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_idx), nq * nk * sizeof(int64_t)));
  SC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_dist), nq * nk * sizeof(double)));
  SC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_cnt), nq * sizeof(uint32_t)));
  SC_HIP(hipMalloc(&s.d_q, nq * ix->ld * es));
  SC_HIP(hipMalloc(reinterpret_cast<void**>(&s.d_idx), nq * nk * sizeof(int64_t)));
  SC_HIP(hipMalloc(reinterpret_cast<void**>(&s.d_dist), nq * nk * sizeof(double)));
  SC_HIP(hipMalloc(reinterpret_cast<void**>(&s.d_cnt), nq * sizeof(uint32_t)));
  s.cap_q = nq;
  s.cap_k = nk;
  return SC_OK;
}

// Packs the batch's rows (zero padded to the index stride), runs the search
// on the slot's stream and queues the copies back.  Called without the lock.
int launch_batch(sc_server* sv, Slot& s) {
  const sc_index* ix = sv->index;
  const size_t es = elem_size(ix->dtype);
  SC_HIP(hipSetDevice(ix->device));
  if (int rc = slot_reserve(s, ix, s.batch_q, s.batch_k)) return rc;
  char* hq = static_cast<char*>(s.h_q);
  size_t row = 0;
  std::memset(hq, 0, s.batch_q * ix->ld * es);
  for (auto& r : s.reqs) {
    for (size_t i = 0; i < r->q; ++i)
      std::memcpy(hq + (row + i) * ix->ld * es, r->rows.data() + i * ix->d * es, ix->d * es);
    row += r->q;
  }
  SC_HIP(hipMemcpyAsync(s.d_q, hq, s.batch_q * ix->ld * es, hipMemcpyHostToDevice, s.stream));
  const auto& r0 = *s.reqs.front();
  if (int rc = search_batch_device(ix, s.d_q, s.batch_q, s.batch_k, r0.metric, r0.p, r0.mode,
                                   s.d_idx, s.d_dist, s.d_cnt, s.stream))
    return rc;
  SC_HIP(hipMemcpyAsync(s.h_idx, s.d_idx, s.batch_q * s.batch_k * sizeof(int64_t),
                        hipMemcpyDeviceToHost, s.stream));
  SC_HIP(hipMemcpyAsync(s.h_dist, s.d_dist, s.batch_q * s.batch_k * sizeof(double),
                        hipMemcpyDeviceToHost, s.stream));
  SC_HIP(hipMemcpyAsync(s.h_cnt, s.d_cnt, s.batch_q * sizeof(uint32_t), hipMemcpyDeviceToHost,
                        s.stream));
  SC_HIP(hipEventRecord(s.done, s.stream));
  return SC_OK;
}

void scheduler_loop(sc_server* sv) {
  (void)hipSetDevice(sv->index->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    int free_slot = -1;
    sv->cv_sched.wait(lk, [&] {
      if (sv->stop) return true;
      free_slot = -1;
      for (size_t i = 0; i < sv->slots.size(); ++i)
        if (!sv->slots[i].busy) {
          free_slot = static_cast<int>(i);
          break;
        }
      return free_slot >= 0 && !sv->pending.empty();
    });
    if (sv->stop) return;
    // a short window lets concurrent arrivals join a batch that would
    // otherwise hold a single small request
    if (sv->cfg.batch_window_us > 0 && sv->pending.size() == 1 &&
        sv->pending.front()->q < sv->cfg.max_batch_queries) {
      sv->cv_sched.wait_for(lk, std::chrono::microseconds(sv->cfg.batch_window_us),
                            [&] { return sv->stop || sv->pending.size() > 1; });
      if (sv->stop) return;
    }
    // form a batch: FIFO, same (metric, p, mode, k), admitted by budget
    Slot& s = sv->slots[free_slot];
    s.reqs.clear();
    s.batch_q = 0;
    s.bytes = 0;
    const auto head = sv->pending.front();
    for (auto it = sv->pending.begin(); it != sv->pending.end();) {
      auto& r = *it;
      const bool same = r->metric == head->metric && r->p == head->p && r->mode == head->mode &&
                        r->k == head->k;
      if (!same) {
        ++it;
        continue;
      }
      if (!s.reqs.empty() && s.batch_q + r->q > sv->cfg.max_batch_queries) break;
      if (sv->in_use + s.bytes + r->bytes > sv->budget) break;  // wait for memory
      s.reqs.push_back(r);
      s.batch_q += r->q;
      s.bytes += r->bytes;
      r->admitted = true;
      it = sv->pending.erase(it);
    }
    if (s.reqs.empty()) {
      // the head does not fit yet: wait for a completion to free memory
      sv->cv_sched.wait(lk, [&] {
        return sv->stop || sv->in_use + sv->pending.front()->bytes <= sv->budget;
      });
      continue;
    }
    s.batch_k = head->k;
    s.busy = true;
    sv->in_use += s.bytes;
    sv->peak = std::max(sv->peak, sv->in_use);
    sv->batches += 1;
    lk.unlock();
    const int rc = launch_batch(sv, s);
    lk.lock();
    s.status = rc;
    if (rc != SC_OK) {
      const size_t n = sc_last_error(nullptr, 0);
      s.error.assign(n, '\0');
      if (n) sc_last_error(&s.error[0], n + 1);
    }
    sv->launched.push_back(free_slot);
    sv->cv_slot.notify_one();
  }
}

void completion_loop(sc_server* sv) {
  (void)hipSetDevice(sv->index->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    sv->cv_slot.wait(lk, [&] { return sv->stop || !sv->launched.empty(); });
    if (sv->launched.empty() && sv->stop) return;
    const int si = sv->launched.front();
    sv->launched.pop_front();
    Slot& s = sv->slots[si];
    lk.unlock();
    int rc = s.status;
    if (rc == SC_OK) {
      const hipError_t e = hipEventSynchronize(s.done);
      if (e != hipSuccess) {
        rc = SC_ERR_DEVICE;
        s.error = std::string("HIP error ") + hipGetErrorString(e);
      }
    }
    const auto now = Clock::now();
    size_t row = 0;
    for (auto& r : s.reqs) {
      r->status = rc;
      r->error = s.error;
      if (rc == SC_OK) {
        r->idx.assign(s.h_idx + row * s.batch_k, s.h_idx + (row + r->q) * s.batch_k);
        r->dist.assign(s.h_dist + row * s.batch_k, s.h_dist + (row + r->q) * s.batch_k);
        r->count.assign(s.h_cnt + row, s.h_cnt + row + r->q);
      }
      r->finished = now;
      row += r->q;
    }
    lk.lock();
    for (auto& r : s.reqs) {
      r->done = true;
      sv->completed += 1;
      sv->queries_done += r->q;
    }
    s.reqs.clear();
    sv->in_use -= s.bytes;
    s.busy = false;
    sv->cv_done.notify_all();
    sv->cv_sched.notify_one();
  }
}

}  // namespace

namespace sc {
// The search of one batch of padded device rows, ordered on stream s.
int search_batch_device(const sc_index* index, const void* d_padded_queries, size_t q, size_t k,
                        sc_metric metric, uint16_t p, sc_mode mode, int64_t* d_idx,
                        double* d_dist, uint32_t* d_count, hipStream_t s) {
  if (mode == SC_MIXED && index->dtype == SC_F32 && metric == SC_EUCLIDIAN)
    return knn_mixed_device(index, d_padded_queries, q, k, d_idx, d_dist, d_count, s, nullptr);
  if (mode == SC_MIXED && metric == SC_MANHATTAN)
    return knn_l1_mixed_device(index, d_padded_queries, q, k, d_idx, d_dist, d_count, s, nullptr);
  return knn_exact_device(index, d_padded_queries, q, k, metric, p, d_idx, d_dist, d_count, s,
                          nullptr);
}
}  // namespace sc

using namespace sc;

extern "C" {

int sc_server_create(const sc_index* index, const sc_server_config* cfg, sc_server** out) {
  clear_error();
  if (!index || !out) return fail(SC_PARAMETERS_ERROR, "null argument");
  *out = nullptr;
  SC_HIP(hipSetDevice(index->device));
  auto sv = std::make_unique<sc_server>();
  sv->index = index;
  sc_server_config c{};
  if (cfg) c = *cfg;
  if (c.max_batch_queries == 0) c.max_batch_queries = 4096;
  if (c.max_inflight == 0) c.max_inflight = 3;
  if (c.hbm_budget_bytes == 0) {
    size_t fr = 0, tot = 0;
    SC_HIP(hipMemGetInfo(&fr, &tot));
    c.hbm_budget_bytes = fr / 10 * 8;
  }
  sv->cfg = c;
  sv->budget = c.hbm_budget_bytes;
  sv->slots.resize(c.max_inflight);
  for (auto& s : sv->slots) {
    hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
      scratch_register(s.stream);
      e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
      for (auto& t : sv->slots) {
        if (t.done) (void)hipEventDestroy(t.done);
        if (t.stream) {
          scratch_release(t.stream);
          (void)hipStreamDestroy(t.stream);
        }
      }
      return hip_fail(e, "server stream");
    }
  }
  sc_server* raw = sv.release();
  raw->sched = std::thread(scheduler_loop, raw);
  raw->compl_ = std::thread(completion_loop, raw);
  *out = raw;
  return SC_OK;
}

int sc_server_submit(sc_server* sv, const void* queries, size_t q, size_t k, sc_metric metric,
                     uint16_t p, sc_mode mode, uint64_t* ticket) {
  clear_error();
  if (!sv || !ticket || (!queries && q > 0)) return fail(SC_PARAMETERS_ERROR, "null argument");
  const sc_index* ix = sv->index;
  // the checks LinearKNNSearch::find makes (linear_search.rs:53-58)
  if (k < 1 || k > ix->n) return fail(SC_FIND_FAILED, "k should be >= 1 and <= length(data)");
  if (metric == SC_MINKOWSKI && p < 1) return fail(SC_ERR_PANIC, "p must be at least 1");
  if (mode == SC_MIXED && metric == SC_MINKOWSKI)
    return fail(SC_ERR_UNSUPPORTED, "SC_MIXED supports the Euclidian and Manhattan metrics only");
  if (q == 0 || q > sv->cfg.max_batch_queries)
    return fail(SC_PARAMETERS_ERROR, "request must carry 1..max_batch_queries rows");
  auto r = std::make_shared<Request>();
  const size_t es = elem_size(ix->dtype);
  r->rows.assign(static_cast<const char*>(queries), static_cast<const char*>(queries) + q * ix->d * es);
  r->q = q;
  r->k = k;
  r->metric = metric;
  r->p = p;
  r->mode = mode;
  r->bytes = estimate_bytes(ix, q, k, mode);
  r->submitted = Clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  if (r->bytes > sv->budget) {
    sv->refused += 1;
    return fail(SC_ERR_OUT_OF_MEMORY, "request exceeds the server's HBM budget");
  }
  r->id = sv->next_id++;
  sv->all[r->id] = r;
  sv->pending.push_back(r);
  *ticket = r->id;
  sv->cv_sched.notify_one();
  return SC_OK;
}

int sc_server_wait(sc_server* sv, uint64_t ticket, int64_t* idx, double* dist, uint32_t* count,
                   double* latency_ms) {
  clear_error();
  if (!sv) return fail(SC_PARAMETERS_ERROR, "null server");
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(ticket);
    if (it == sv->all.end()) return fail(SC_PARAMETERS_ERROR, "unknown ticket");
    r = it->second;
    sv->cv_done.wait(lk, [&] { return r->done; });
    sv->all.erase(it);
  }
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
  if (r->status != SC_OK) return fail(r->status, r->error);
  if (idx) std::copy(r->idx.begin(), r->idx.end(), idx);
  if (dist) std::copy(r->dist.begin(), r->dist.end(), dist);
  if (count) std::copy(r->count.begin(), r->count.end(), count);
  return SC_OK;
}

int sc_server_stats(sc_server* sv, sc_server_info* out) {
  if (!sv || !out) return SC_PARAMETERS_ERROR;
  std::lock_guard<std::mutex> lk(sv->mu);
  out->completed = sv->completed;
  out->batches = sv->batches;
  out->refused = sv->refused;
  out->queries = sv->queries_done;
  out->pending = sv->pending.size();
  out->hbm_in_use = sv->in_use;
  out->hbm_peak = sv->peak;
  out->hbm_budget = sv->budget;
  return SC_OK;
}

int sc_server_destroy(sc_server* sv) {
  if (!sv) return SC_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_sched.notify_all();
  sv->cv_slot.notify_all();
  if (sv->sched.joinable()) sv->sched.join();
  // requests never launched fail; their waiters wake up
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (auto& r : sv->pending) {
      r->status = SC_INVALID_STATE;
      r->error = "server destroyed before the request ran";
      r->finished = Clock::now();
      r->done = true;
    }
    sv->pending.clear();
    sv->cv_done.notify_all();
  }
  // drain what was launched
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] {
      for (auto& s : sv->slots)
        if (s.busy) return false;
      return true;
    });
  }
  sv->cv_slot.notify_all();
  if (sv->compl_.joinable()) sv->compl_.join();
  (void)hipSetDevice(sv->index->device);
  for (auto& s : sv->slots) {
    if (s.stream) {
      scratch_release(s.stream);
      (void)hipStreamSynchronize(s.stream);
    }
    if (s.h_q) {
      (void)hipHostFree(s.h_q);
      (void)hipHostFree(s.h_idx);
      (void)hipHostFree(s.h_dist);
      (void)hipHostFree(s.h_cnt);
      (void)hipFree(s.d_q);
      (void)hipFree(s.d_idx);
      (void)hipFree(s.d_dist);
      (void)hipFree(s.d_cnt);
    }
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
  }
  delete sv;
  return SC_OK;
}

}  // extern "C"
