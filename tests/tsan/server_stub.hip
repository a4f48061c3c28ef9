// Host ThreadSanitizer harness for the serving scheduler (csrc/serve.hip).
//
// The HIP runtime does not come up inside a ThreadSanitizer process, so the
// scheduler's host code is exercised on the CPU: this file supplies the HIP
// calls the server makes (host memory for device memory, copies done at
// once, an event as a release/acquire flag, which is the ordering the real
// event gives the completion thread) and a search that writes a
// deterministic answer per query row after a short sleep (behind the
// server's own batch dispatch).  Both this file
// and serve.hip are compiled host-only with -Xarch_host -fsanitize=thread by
// tests/tsan/build.py; the test runs the result and fails on any report.
//
// Four submitter threads push requests of mixed metric, mode, k and size at
// one server with three slots and a small admission budget (so requests
// also wait for memory), while a fifth thread polls sc_server_stats; every
// answer is compared with the same deterministic answer computed directly.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

// ---- the HIP calls of serve.hip, on the host ----
namespace {
struct FakeEvent {
  std::atomic<int> done{0};
};
}  // namespace

extern "C" {
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) {
  *p = std::malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t) {
  std::memcpy(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* free_b, size_t* total) {
  *free_b = size_t(1) << 30;
  *total = size_t(1) << 31;
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "stub error"; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(new char[16]);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete[] reinterpret_cast<char*>(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
  *e = reinterpret_cast<hipEvent_t>(new FakeEvent());
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete reinterpret_cast<FakeEvent*>(e);
  return hipSuccess;
}
// the copies and the search already ran in the recording thread; the event
// publishes their results to the thread that waits on it
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  reinterpret_cast<FakeEvent*>(e)->done.store(1, std::memory_order_release);
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  while (reinterpret_cast<FakeEvent*>(e)->done.load(std::memory_order_acquire) == 0)
    std::this_thread::yield();
  return hipSuccess;
}
}  // extern "C"

// ---- the library pieces serve.hip links against ----
namespace {
thread_local std::string t_err;

// The answer of one query row: k pseudo-neighbours from the row's bits.
void answer(const float* row, size_t d, size_t k, uint64_t n, int metric, int64_t* idx,
            double* dist, uint32_t* cnt) {
  uint64_t h = 1469598103934665603ull ^ static_cast<uint64_t>(metric);
  double s = 0.0;
  for (size_t j = 0; j < d; ++j) {
    uint32_t b;
    std::memcpy(&b, row + j, 4);
    h = (h ^ b) * 1099511628211ull;
    s += row[j];
  }
  for (size_t j = 0; j < k; ++j) {
    idx[j] = static_cast<int64_t>((h + 7919 * j) % n);
    dist[j] = s + static_cast<double>(j);
  }
  *cnt = static_cast<uint32_t>(k);
}
}  // namespace

namespace sc {
int fail(int code, const std::string& msg) {
  t_err = msg;
  return code;
}
void clear_error() { t_err.clear(); }
int hip_fail(hipError_t, const char* what) { return fail(SC_ERR_DEVICE, what); }
void scratch_register(hipStream_t) {}
void scratch_release(hipStream_t) {}
// the searches serve.hip's batch dispatch calls, on the host
int stub_search(const sc_index* index, const void* d_queries, size_t q, size_t k, int metric,
                int64_t* d_idx, double* d_dist, uint32_t* d_count) {
  // the device's time, so batches overlap the scheduler's work
  std::this_thread::sleep_for(std::chrono::microseconds(200 + 2 * q));
  const float* qs = static_cast<const float*>(d_queries);
  for (size_t i = 0; i < q; ++i)
    answer(qs + i * index->ld, index->d, k, index->n, metric, d_idx + i * k, d_dist + i * k,
           d_count + i);
  return SC_OK;
}
int knn_exact_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                     sc_metric metric, int, int64_t* d_idx, double* d_dist, uint32_t* d_count,
                     hipStream_t, sc_stats*) {
  return stub_search(index, d_queries, q, k, metric, d_idx, d_dist, d_count);
}
int knn_l1_mixed_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                        int64_t* d_idx, double* d_dist, uint32_t* d_count, hipStream_t,
                        sc_stats*) {
  return stub_search(index, d_queries, q, k, SC_MANHATTAN, d_idx, d_dist, d_count);
}
int knn_mixed_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                     int64_t* d_idx, double* d_dist, uint32_t* d_count, hipStream_t, sc_stats*) {
  return stub_search(index, d_queries, q, k, SC_EUCLIDIAN, d_idx, d_dist, d_count);
}
}  // namespace sc

extern "C" size_t sc_last_error(char* buf, size_t len) {
  if (buf && len) {
    const size_t n = std::min(len - 1, t_err.size());
    std::memcpy(buf, t_err.data(), n);
    buf[n] = '\0';
  }
  return t_err.size();
}

namespace {
struct Job {
  std::vector<float> rows;
  size_t q = 0, k = 0;
  sc_metric metric = SC_EUCLIDIAN;
  sc_mode mode = SC_PLAIN_FP64;
  uint64_t ticket = 0;
  std::vector<int64_t> idx;
  std::vector<double> dist;
  std::vector<uint32_t> cnt;
  int rc = -1;
};
}  // namespace

int main() {
  sc_index ix{};
  ix.dtype = SC_F32;
  ix.n = 50000;
  ix.d = 12;
  ix.ld = 16;
  ix.device = 0;
  std::mutex ix_mu;
  ix.mu = &ix_mu;

  sc_server_config cfg{};
  cfg.max_batch_queries = 512;
  cfg.max_inflight = 3;
  cfg.hbm_budget_bytes = size_t(40) << 20;
  cfg.batch_window_us = 100;
  sc_server* sv = nullptr;
  if (sc_server_create(&ix, &cfg, &sv) != 0) return 3;

  const int kThreads = 4, kPerThread = 60;
  std::vector<std::vector<Job>> jobs(kThreads);
  for (int t = 0; t < kThreads; ++t) {
    std::mt19937_64 r(100 + t);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    for (int j = 0; j < kPerThread; ++j) {
      Job jb;
      jb.q = 1 + r() % 400;
      jb.k = 1 + r() % 20;
      jb.metric = (r() % 5 == 0) ? SC_MANHATTAN : SC_EUCLIDIAN;
      jb.mode = (r() % 2) ? SC_MIXED : SC_PLAIN_FP64;
      jb.rows.resize(jb.q * ix.d);
      for (auto& v : jb.rows) v = nd(r);
      jobs[t].push_back(std::move(jb));
    }
  }
  std::atomic<bool> done{false};
  std::thread poller([&] {
    sc_server_info info{};
    while (!done.load()) {
      (void)sc_server_stats(sv, &info);
      std::this_thread::yield();
    }
  });
  std::vector<std::thread> th;
  for (int t = 0; t < kThreads; ++t)
    th.emplace_back([&, t] {
      // submit all, then wait all: a thread's requests overlap each other
      for (auto& jb : jobs[t])
        jb.rc = sc_server_submit(sv, jb.rows.data(), jb.q, jb.k, jb.metric, 2, jb.mode, &jb.ticket);
      for (auto& jb : jobs[t]) {
        if (jb.rc != 0) continue;
        jb.idx.resize(jb.q * jb.k);
        jb.dist.resize(jb.q * jb.k);
        jb.cnt.resize(jb.q);
        double lat = 0.0;
        jb.rc = sc_server_wait(sv, jb.ticket, jb.idx.data(), jb.dist.data(), jb.cnt.data(), &lat);
      }
    });
  for (auto& t : th) t.join();
  done.store(true);
  poller.join();
  sc_server_info info{};
  (void)sc_server_stats(sv, &info);
  if (sc_server_destroy(sv) != 0) return 4;

  size_t bad = 0, failed = 0, total = 0;
  for (auto& per : jobs)
    for (auto& jb : per) {
      ++total;
      if (jb.rc != 0) {
        ++failed;
        continue;
      }
      std::vector<int64_t> idx(jb.k);
      std::vector<double> dist(jb.k);
      for (size_t i = 0; i < jb.q; ++i) {
        uint32_t c = 0;
        answer(jb.rows.data() + i * ix.d, ix.d, jb.k, ix.n, jb.metric, idx.data(), dist.data(), &c);
        if (c != jb.cnt[i] || std::memcmp(idx.data(), &jb.idx[i * jb.k], jb.k * 8) != 0 ||
            std::memcmp(dist.data(), &jb.dist[i * jb.k], jb.k * 8) != 0) {
          ++bad;
          break;
        }
      }
    }
  std::printf("requests %zu, failed %zu, differing %zu, batches %llu\n", total, failed, bad,
              static_cast<unsigned long long>(info.batches));
  return (bad == 0 && failed == 0) ? 0 : 1;
}
