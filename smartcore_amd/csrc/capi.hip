// C ABI surface of libsmartcore_amd: index lifetime, kNN search entry
// points, error and stats reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace sc {

namespace {
thread_local std::string t_last_error;
thread_local sc_stats t_stats{};

// One non-blocking stream per host thread and device, at the default
// priority; its scratch is cached.
hipStream_t thread_stream(int device) {
  static constexpr int kMaxDev = 64;
  thread_local hipStream_t streams[kMaxDev] = {};
  if (device < 0 || device >= kMaxDev) return nullptr;
  if (!streams[device]) {
    if (hipStreamCreateWithFlags(&streams[device], hipStreamNonBlocking) != hipSuccess)
      streams[device] = nullptr;
    else
      scratch_register(streams[device]);
  }
  return streams[device];
}

struct ScratchCache {
  struct PerStream {
    std::multimap<size_t, void*> blocks;  // by size
    size_t bytes = 0;
  };
  std::mutex mu;
  std::unordered_map<hipStream_t, PerStream> streams;
};
// never destroyed: threads may return blocks during process exit
ScratchCache& scratch_cache() {
  static ScratchCache* c = new ScratchCache();
  return *c;
}
constexpr size_t kScratchKeepPerStream = size_t(8) << 30;

// Scratch of a call is stream-ordered (hipMallocAsync / hipFreeAsync from
// the device's default pool).  The pool keeps up to kPoolKeepBytes of freed
// scratch across synchronisations instead of handing it back at each one,
// so repeated searches do not map their candidate buffers afresh.
constexpr uint64_t kPoolKeepBytes = uint64_t(16) << 30;

void configure_pool(int device) {
  static constexpr int kMaxDev = 64;
  static std::once_flag once[kMaxDev];
  if (device < 0 || device >= kMaxDev) return;
  std::call_once(once[device], [device] {
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) != hipSuccess || !pool) {
      (void)hipGetLastError();
      return;
    }
    uint64_t keep = kPoolKeepBytes;
    if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess)
      (void)hipGetLastError();
  });
}

__global__ void pad_rows_kernel(const char* __restrict__ src, char* __restrict__ dst, size_t n,
                                size_t row_bytes, size_t ld_bytes) {
  const size_t total = n * ld_bytes;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t r = e / ld_bytes, c = e % ld_bytes;
    dst[e] = c < row_bytes ? src[r * row_bytes + c] : 0;
  }
}
}  // namespace

int fail(int code, const std::string& msg) {
  t_last_error = msg;
  return code;
}

void clear_error() { t_last_error.clear(); }

sc_stats& last_stats() { return t_stats; }

int hip_fail(hipError_t e, const char* what) {
  // HIP keeps the last failure of any call for hipGetLastError; clear it so
  // that the launch checks of the next call do not report this one again
  (void)hipGetLastError();
  const int code = (e == hipErrorOutOfMemory) ? SC_ERR_OUT_OF_MEMORY : SC_ERR_DEVICE;
  return fail(code, std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
}

namespace {

constexpr size_t kStageBytes = size_t{16} << 20;  // per pinned buffer
constexpr int kCopyThreads = 8;

// The buffers are kept for the life of the process (like the per-thread
// streams): freeing them from a thread-exit destructor could run after the
// HIP runtime has shut down.
struct Stager {
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  hipError_t init() {
    for (int i = 0; i < 2; ++i) {
      if (!buf[i]) {
        hipError_t e = hipHostMalloc(&buf[i], kStageBytes, hipHostMallocDefault);
        if (e != hipSuccess) {
          buf[i] = nullptr;
          return e;
        }
      }
      if (!ev[i]) {
        hipError_t e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
        if (e != hipSuccess) {
          ev[i] = nullptr;
          return e;
        }
      }
    }
    return hipSuccess;
  }
};

// One stager per host thread (its buffers are reused by the thread's calls).
Stager& thread_stager() {
  thread_local Stager st;
  return st;
}

// Host copy workers shared by every host thread, started once: run(fn)
// calls fn(t, kCopyThreads) for every t, t = 0 on the caller, and returns
// when all have returned; one job at a time.  The pool is never destroyed
// (its workers wait for work until the process ends).
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* pool = new CopyPool();
    return *pool;
  }
  void run(const std::function<void(int, int)>& fn) {
    std::lock_guard<std::mutex> one(use_mu_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      pending_ = kCopyThreads - 1;
      ++gen_;
    }
    cv_.notify_all();
    fn(0, kCopyThreads);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  CopyPool() {
    for (int t = 1; t < kCopyThreads; ++t) std::thread([this, t] { loop(t); }).detach();
  }
  void loop(int t) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int, int)>* job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        job = job_;
      }
      (*job)(t, kCopyThreads);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::mutex use_mu_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int, int)>* job_ = nullptr;
  uint64_t gen_ = 0;
  int pending_ = 0;
};

// memcpy split over the copy workers for large spans.
void parallel_copy(void* dst, const void* src, size_t bytes) {
  if (bytes < (size_t{4} << 20)) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const size_t part = (bytes / kCopyThreads + 4095) / 4096 * 4096;
  CopyPool::get().run([&](int t, int) {
    const size_t off = part * static_cast<size_t>(t);
    if (off >= bytes) return;
    std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off,
                std::min(part, bytes - off));
  });
}

}  // namespace

int upload_staged(void* dev, const void* host, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return SC_OK;
  Stager& st = thread_stager();
  SC_HIP(st.init());
  int b = 0;
  for (size_t off = 0; off < bytes; off += kStageBytes, b ^= 1) {
    const size_t len = std::min(kStageBytes, bytes - off);
    if (st.used[b]) SC_HIP(hipEventSynchronize(st.ev[b]));
    parallel_copy(st.buf[b], static_cast<const char*>(host) + off, len);
    SC_HIP(hipMemcpyAsync(static_cast<char*>(dev) + off, st.buf[b], len, hipMemcpyHostToDevice,
                          stream));
    SC_HIP(hipEventRecord(st.ev[b], stream));
    st.used[b] = true;
  }
  return SC_OK;
}

int download_staged(void* host, const void* dev, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return SC_OK;
  Stager& st = thread_stager();
  SC_HIP(st.init());
  // chunk i is copied out of its buffer while chunk i + 1 is in flight
  const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
  auto issue = [&](size_t c) -> hipError_t {
    const size_t off = c * kStageBytes;
    const size_t len = std::min(kStageBytes, bytes - off);
    const int b = static_cast<int>(c & 1);
    hipError_t e = hipMemcpyAsync(st.buf[b], static_cast<const char*>(dev) + off, len,
                                  hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return e;
    st.used[b] = true;
    return hipEventRecord(st.ev[b], stream);
  };
  SC_HIP(issue(0));
  for (size_t c = 0; c < nch; ++c) {
    if (c + 1 < nch) SC_HIP(issue(c + 1));
    const int b = static_cast<int>(c & 1);
    SC_HIP(hipEventSynchronize(st.ev[b]));
    const size_t off = c * kStageBytes;
    parallel_copy(static_cast<char*>(host) + off, st.buf[b], std::min(kStageBytes, bytes - off));
  }
  return SC_OK;
}

namespace {
// n rows of rb bytes to a stride of lb bytes, the padding zeroed (word
// loops: rows here are a few dozen bytes, too short for memcpy calls)
void copy_rows(char* __restrict__ dst, const char* __restrict__ src, size_t n, size_t rb, size_t lb) {
  if (rb == lb) {
    std::memcpy(dst, src, n * rb);
    return;
  }
  if (rb % 4 == 0 && lb % 4 == 0) {
    const size_t rw = rb / 4, lw = lb / 4;
    for (size_t r = 0; r < n; ++r) {
      uint32_t* __restrict__ d = reinterpret_cast<uint32_t*>(dst + r * lb);
      const char* sr = src + r * rb;
      for (size_t w = 0; w < rw; ++w) {
        uint32_t v;
        std::memcpy(&v, sr + 4 * w, 4);
        d[w] = v;
      }
      for (size_t w = rw; w < lw; ++w) d[w] = 0u;
    }
    return;
  }
  for (size_t r = 0; r < n; ++r) {
    std::memcpy(dst + r * lb, src + r * rb, rb);
    std::memset(dst + r * lb + rb, 0, lb - rb);
  }
}
}  // namespace

int upload_staged_segments(void* dev, const StageSeg* segs, size_t count, hipStream_t stream) {
  Stager& st = thread_stager();
  SC_HIP(st.init());
  size_t si = 0, r = 0;  // next row to stage
  size_t dev_off = 0;
  int b = 0;
  struct Piece {
    const char* src;
    size_t n, rb, lb, at;  // rows, row bytes, padded bytes, offset in the chunk
  };
  std::vector<Piece> pieces;
  while (si < count) {
    // whole rows, in order across segments, up to one staging buffer
    pieces.clear();
    size_t fill = 0;
    while (si < count) {
      const StageSeg& g = segs[si];
      if (r >= g.n) {
        ++si;
        r = 0;
        continue;
      }
      const size_t room = (kStageBytes - fill) / g.ld_bytes;
      if (room == 0) break;
      const size_t take = std::min(room, g.n - r);
      pieces.push_back(Piece{static_cast<const char*>(g.host) + r * g.row_bytes, take, g.row_bytes,
                             g.ld_bytes, fill});
      fill += take * g.ld_bytes;
      r += take;
    }
    if (fill == 0) break;
    if (st.used[b]) SC_HIP(hipEventSynchronize(st.ev[b]));
    char* dst = static_cast<char*>(st.buf[b]);
    auto work = [&](size_t t, size_t nt) {
      // piece rows split evenly over nt workers
      for (const Piece& pc : pieces) {
        const size_t a = pc.n * t / nt, e = pc.n * (t + 1) / nt;
        if (e > a) copy_rows(dst + pc.at + a * pc.lb, pc.src + a * pc.rb, e - a, pc.rb, pc.lb);
      }
    };
    if (fill < (size_t{4} << 20))
      work(0, 1);
    else
      CopyPool::get().run([&](int t, int nt) { work(static_cast<size_t>(t), static_cast<size_t>(nt)); });
    SC_HIP(hipMemcpyAsync(static_cast<char*>(dev) + dev_off, st.buf[b], fill, hipMemcpyHostToDevice,
                          stream));
    SC_HIP(hipEventRecord(st.ev[b], stream));
    st.used[b] = true;
    dev_off += fill;
    b ^= 1;
  }
  return SC_OK;
}

int upload_staged_rows(void* dev, const void* host, size_t n, size_t row_bytes, size_t ld_bytes,
                       hipStream_t stream) {
  if (n == 0) return SC_OK;
  if (row_bytes == ld_bytes) return upload_staged(dev, host, n * row_bytes, stream);
  Stager& st = thread_stager();
  SC_HIP(st.init());
  // rows are widened to the padded stride while they are copied into the
  // pinned buffer, so the DMA moves the final layout
  const size_t rows_per_chunk = std::max<size_t>(kStageBytes / ld_bytes, 1);
  int b = 0;
  for (size_t r0 = 0; r0 < n; r0 += rows_per_chunk, b ^= 1) {
    const size_t nr = std::min(rows_per_chunk, n - r0);
    if (st.used[b]) SC_HIP(hipEventSynchronize(st.ev[b]));
    char* dst = static_cast<char*>(st.buf[b]);
    const char* src = static_cast<const char*>(host) + r0 * row_bytes;
    auto pad_rows = [=](size_t a, size_t e) {
      copy_rows(dst + a * ld_bytes, src + a * row_bytes, e - a, row_bytes, ld_bytes);
    };
    if (nr * ld_bytes < (size_t{4} << 20)) {
      pad_rows(0, nr);
    } else {
      const size_t part = (nr + kCopyThreads - 1) / kCopyThreads;
      CopyPool::get().run([&](int t, int) {
        pad_rows(std::min(nr, part * t), std::min(nr, part * (t + 1)));
      });
    }
    SC_HIP(hipMemcpyAsync(static_cast<char*>(dev) + r0 * ld_bytes, st.buf[b], nr * ld_bytes,
                          hipMemcpyHostToDevice, stream));
    SC_HIP(hipEventRecord(st.ev[b], stream));
    st.used[b] = true;
  }
  return SC_OK;
}

int upload_padded(const void* host, sc_dtype dtype, size_t n, size_t d, size_t ld, void* dev,
                  hipStream_t stream) {
  const size_t es = elem_size(dtype);
  if (n == 0) return SC_OK;
  if (d == 0) {
    SC_HIP(hipMemsetAsync(dev, 0, n * ld * es, stream));
    return SC_OK;
  }
  return upload_staged_rows(dev, host, n, d * es, ld * es, stream);
}

int copy_padded_device(const void* src, sc_dtype dtype, size_t n, size_t d, size_t ld, void* dev,
                       hipStream_t stream) {
  const size_t es = elem_size(dtype);
  if (n == 0) return SC_OK;
  if (ld == d) {
    SC_HIP(hipMemcpyAsync(dev, src, n * d * es, hipMemcpyDeviceToDevice, stream));
    return SC_OK;
  }
  const size_t total = n * ld * es;
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(pad_rows_kernel, dim3(blocks), dim3(256), 0, stream,
                     static_cast<const char*>(src), static_cast<char*>(dev), n, d * es, ld * es);
  SC_HIP(hipGetLastError());
  return SC_OK;
}

namespace {
struct ProfRecord {
  std::string name;
  hipEvent_t start;
  hipEvent_t stop;
};
std::mutex g_prof_mu;
std::atomic<bool> g_prof_on{false};
std::vector<ProfRecord> g_prof;
}  // namespace

ProfScope::ProfScope(const char* name, hipStream_t stream) : name_(name), stream_(stream) {
  if (!g_prof_on.load(std::memory_order_relaxed)) return;
  // Launches recorded into a graph are timed by the scope around the graph
  // launch: events recorded during a capture cannot be timed.
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  if (hipEventCreate(&start_) != hipSuccess) {
    start_ = nullptr;
    return;
  }
  (void)hipEventRecord(start_, stream_);
}

ProfScope::~ProfScope() {
  if (!start_) return;
  hipEvent_t stop = nullptr;
  if (hipEventCreate(&stop) != hipSuccess) {
    (void)hipEventDestroy(start_);
    return;
  }
  (void)hipEventRecord(stop, stream_);
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.push_back(ProfRecord{name_, start_, stop});
}

bool profiling_enabled() { return g_prof_on.load(std::memory_order_relaxed); }

void scratch_register(hipStream_t s) {
  ScratchCache& c = scratch_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  c.streams[s];
}

void scratch_release(hipStream_t s) {
  ScratchCache& c = scratch_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.streams.find(s);
  if (it == c.streams.end()) return;
  for (auto& b : it->second.blocks) (void)hipFreeAsync(b.second, s);
  c.streams.erase(it);
}

void scratch_trim(hipStream_t s) {
  ScratchCache& c = scratch_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.streams.find(s);
  if (it == c.streams.end()) return;
  for (auto& b : it->second.blocks) (void)hipFreeAsync(b.second, s);
  it->second.blocks.clear();
  it->second.bytes = 0;
}

size_t scratch_round(size_t want) {
  // powers of two up to 64 MiB, then whole 64 MiB
  constexpr size_t kBig = size_t(64) << 20;
  if (want >= kBig) return (want + kBig - 1) / kBig * kBig;
  size_t r = 256;
  while (r < want) r <<= 1;
  return r;
}

namespace {
bool capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}
}  // namespace

void* scratch_take(hipStream_t s, size_t want, size_t* cap) {
  ScratchCache& c = scratch_cache();
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.streams.find(s);
    if (it == c.streams.end()) return nullptr;
    auto& ps = it->second;
    // the smallest block that fits and is not far larger than needed
    auto b = ps.blocks.lower_bound(want);
    if (b == ps.blocks.end() || b->first > 2 * scratch_round(want)) return nullptr;
    if (capturing(s)) return nullptr;
    void* p = b->second;
    *cap = b->first;
    ps.bytes -= b->first;
    ps.blocks.erase(b);
    return p;
  }
}

bool scratch_give(hipStream_t s, void* p, size_t cap) {
  ScratchCache& c = scratch_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  auto it = c.streams.find(s);
  if (it == c.streams.end() || capturing(s)) return false;
  auto& ps = it->second;
  if (ps.bytes + cap > kScratchKeepPerStream) return false;
  ps.blocks.emplace(cap, p);
  ps.bytes += cap;
  return true;
}

void prof_clear_locked() {
  for (auto& r : g_prof) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  g_prof.clear();
}

hipStream_t stream_for(const sc_index* index) { return thread_stream(index->device); }
hipStream_t stream_for_device(int device) { return thread_stream(device); }

// Created with a CU mask that enables every CU: such a stream gets a hardware
// queue of its own.  A stream that shares its queue with kernel streams sees
// every event it records (the copy done, a staging buffer free) complete
// only after the kernels queued ahead of it there, which halved the rate of
// staged uploads next to running kernels.
hipStream_t copy_stream_for_device(int device) {
  static constexpr int kMaxDev = 64;
  thread_local hipStream_t streams[kMaxDev] = {};
  if (device < 0 || device >= kMaxDev) return nullptr;
  if (!streams[device]) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
      return nullptr;
    std::vector<uint32_t> mask((static_cast<size_t>(cus) + 31) / 32, 0u);
    for (int c = 0; c < cus; ++c) mask[c / 32] |= 1u << (c % 32);
    if (hipExtStreamCreateWithCUMask(&streams[device], static_cast<uint32_t>(mask.size()),
                                     mask.data()) != hipSuccess)
      streams[device] = nullptr;
  }
  return streams[device];
}

}  // namespace sc

using namespace sc;

namespace {

int check_rows(sc_dtype dtype, size_t n, size_t d) {
  if (dtype != SC_F64 && dtype != SC_F32) return fail(SC_ERR_UNSUPPORTED, "dtype must be f64 or f32");
  if (n >= (size_t(1) << 32)) return fail(SC_ERR_UNSUPPORTED, "at most 2^32-1 rows per index");
  (void)d;
  return SC_OK;
}

int index_create_impl(const void* rows, bool on_device, sc_dtype dtype, size_t n, size_t d,
                      sc_index** out) {
  clear_error();
  if (!out) return fail(SC_PARAMETERS_ERROR, "null output handle");
  *out = nullptr;
  if (int rc = check_rows(dtype, n, d)) return rc;
  int dev = 0;
  SC_HIP(hipGetDevice(&dev));
  sc_index* ix = new sc_index{};
  ix->dtype = dtype;
  ix->n = n;
  ix->d = d;
  ix->ld = padded_dim(d);
  ix->device = dev;
  ix->mu = new std::mutex();
  configure_pool(dev);
  hipStream_t s = thread_stream(dev);
  const size_t bytes = std::max<size_t>(n * ix->ld * elem_size(dtype), 16);
  hipError_t e = hipMalloc(&ix->rows, bytes);
  if (e != hipSuccess) {
    delete static_cast<std::mutex*>(ix->mu);
    delete ix;
    return hip_fail(e, "hipMalloc(index rows)");
  }
  int rc = on_device ? copy_padded_device(rows, dtype, n, d, ix->ld, ix->rows, s)
                     : upload_padded(rows, dtype, n, d, ix->ld, ix->rows, s);
  if (rc == SC_OK) {
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "upload");
  }
  if (rc != SC_OK) {
    (void)hipFree(ix->rows);
    delete static_cast<std::mutex*>(ix->mu);
    delete ix;
    return rc;
  }
  *out = ix;
  return SC_OK;
}

int validate_search(const sc_index* index, size_t k, sc_metric metric, uint16_t p, sc_mode mode) {
  if (!index) return fail(SC_PARAMETERS_ERROR, "null index");
  // linear_search.rs:53-58
  if (k < 1 || k > index->n)
    return fail(SC_FIND_FAILED, "k should be >= 1 and <= length(data)");
  if (metric != SC_EUCLIDIAN && metric != SC_MANHATTAN && metric != SC_MINKOWSKI)
    return fail(SC_ERR_UNSUPPORTED, "unknown metric");
  // minkowski.rs:57-59
  if (metric == SC_MINKOWSKI && p < 1) return fail(SC_ERR_PANIC, "p must be at least 1");
  if (mode != SC_PLAIN_FP64 && mode != SC_MIXED) return fail(SC_ERR_UNSUPPORTED, "unknown mode");
  if (mode == SC_MIXED && metric == SC_MINKOWSKI)
    return fail(SC_ERR_UNSUPPORTED, "SC_MIXED supports the Euclidian and Manhattan metrics only");
  return SC_OK;
}

int search_device_impl(const sc_index* index, const void* d_padded_queries, size_t q, size_t k,
                       sc_metric metric, uint16_t p, sc_mode mode, int64_t* d_idx, double* d_dist,
                       uint32_t* d_count, hipStream_t s, bool with_stats) {
  sc_stats& st = last_stats();
  st = sc_stats{};
  sc_stats* sp = with_stats ? &st : nullptr;
  if (mode == SC_MIXED && index->dtype == SC_F32 && metric == SC_EUCLIDIAN)
    return knn_mixed_device(index, d_padded_queries, q, k, d_idx, d_dist, d_count, s, sp);
  if (mode == SC_MIXED && metric == SC_MANHATTAN)
    return knn_l1_mixed_device(index, d_padded_queries, q, k, d_idx, d_dist, d_count, s, sp);
  return knn_exact_device(index, d_padded_queries, q, k, metric, p, d_idx, d_dist, d_count, s, sp);
}

}  // namespace

extern "C" {

int sc_index_create(const void* rows, sc_dtype dtype, size_t n, size_t d, sc_index** out) {
  if (!rows && n > 0) return fail(SC_PARAMETERS_ERROR, "null rows");
  return index_create_impl(rows, false, dtype, n, d, out);
}

int sc_index_create_device(const void* d_rows, sc_dtype dtype, size_t n, size_t d,
                           sc_index** out) {
  if (!d_rows && n > 0) return fail(SC_PARAMETERS_ERROR, "null rows");
  return index_create_impl(d_rows, true, dtype, n, d, out);
}

int sc_index_destroy(sc_index* index) {
  clear_error();
  if (!index) return SC_OK;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(index->device);
  if (index->rows) (void)hipFree(index->rows);
  if (index->rows_bf16) (void)hipFree(index->rows_bf16);
  if (index->norms2) (void)hipFree(index->norms2);
  if (index->rows_q16) (void)hipFree(index->rows_q16);
  if (index->err_q16) (void)hipFree(index->err_q16);
  (void)hipSetDevice(cur);
  delete static_cast<std::mutex*>(index->mu);
  delete index;
  return SC_OK;
}

size_t sc_index_rows(const sc_index* index) { return index ? index->n : 0; }
size_t sc_index_dim(const sc_index* index) { return index ? index->d : 0; }

int sc_knn_search(const sc_index* index, const void* queries, size_t q, size_t k,
                  sc_metric metric, uint16_t p, sc_mode mode, int64_t* idx, double* dist,
                  uint32_t* count) {
  clear_error();
  if (int rc = validate_search(index, k, metric, p, mode)) return rc;
  if (q == 0) {
    last_stats() = sc_stats{};
    return SC_OK;
  }
  if (!queries || !idx || !dist || !count) return fail(SC_PARAMETERS_ERROR, "null buffer");
  SC_HIP(hipSetDevice(index->device));
  hipStream_t s = stream_for(index);
  const size_t es = elem_size(index->dtype);
  DevBuf dq, di, dd, dc;
  SC_HIP(dq.alloc(q * index->ld * es, s));
  SC_HIP(di.alloc(q * k * sizeof(int64_t), s));
  SC_HIP(dd.alloc(q * k * sizeof(double), s));
  SC_HIP(dc.alloc(q * sizeof(uint32_t), s));
  if (int rc = upload_padded(queries, index->dtype, q, index->d, index->ld, dq.p, s)) return rc;
  if (int rc = search_device_impl(index, dq.p, q, k, metric, p, mode, di.as<int64_t>(),
                                  dd.as<double>(), dc.as<uint32_t>(), s, true))
    return rc;
  if (int rc = download_staged(idx, di.p, q * k * sizeof(int64_t), s)) return rc;
  if (int rc = download_staged(dist, dd.p, q * k * sizeof(double), s)) return rc;
  if (int rc = download_staged(count, dc.p, q * sizeof(uint32_t), s)) return rc;
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

int sc_knn_search_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                         sc_metric metric, uint16_t p, sc_mode mode, int64_t* d_idx,
                         double* d_dist, uint32_t* d_count, void* stream) {
  clear_error();
  if (int rc = validate_search(index, k, metric, p, mode)) return rc;
  if (q == 0) {
    last_stats() = sc_stats{};
    return SC_OK;
  }
  if (!d_queries || !d_idx || !d_dist || !d_count) return fail(SC_PARAMETERS_ERROR, "null buffer");
  SC_HIP(hipSetDevice(index->device));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (index->ld == index->d) {
    return search_device_impl(index, d_queries, q, k, metric, p, mode, d_idx, d_dist, d_count, s,
                              false);
  }
  DevBuf dq;
  SC_HIP(dq.alloc(q * index->ld * elem_size(index->dtype), s));
  if (int rc = copy_padded_device(d_queries, index->dtype, q, index->d, index->ld, dq.p, s))
    return rc;
  return search_device_impl(index, dq.p, q, k, metric, p, mode, d_idx, d_dist, d_count, s, false);
}

int sc_knn_search_batch(const sc_index* index, const sc_knn_block* blocks, size_t count,
                        sc_metric metric, uint16_t p, sc_mode mode) {
  clear_error();
  if (count == 0) return SC_OK;
  if (!index || !blocks) return fail(SC_PARAMETERS_ERROR, "null argument");
  for (size_t b = 0; b < count; ++b) {
    if (int rc = validate_search(index, blocks[b].k, metric, p, mode)) return rc;
    if (blocks[b].q > 0 && (!blocks[b].queries || !blocks[b].idx || !blocks[b].dist ||
                            !blocks[b].count))
      return fail(SC_PARAMETERS_ERROR, "block " + std::to_string(b) + " has a null buffer");
  }
  SC_HIP(hipSetDevice(index->device));
  hipStream_t s = stream_for(index);
  const size_t es = elem_size(index->dtype);
  // blocks with the same k are searched together: one device search over
  // their concatenated rows (each query is searched independently of the
  // rows beside it, so every block gets what sc_knn_search gives it)
  std::vector<size_t> order(count);
  for (size_t b = 0; b < count; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return blocks[a].k < blocks[b].k; });
  sc_stats total{};
  for (size_t g0 = 0; g0 < count;) {
    size_t g1 = g0;
    size_t rows = 0;
    const size_t k = blocks[order[g0]].k;
    while (g1 < count && blocks[order[g1]].k == k) rows += blocks[order[g1++]].q;
    if (rows > 0) {
      DevBuf dq, di, dd, dc;
      SC_HIP(dq.alloc(rows * index->ld * es, s));
      SC_HIP(di.alloc(rows * k * sizeof(int64_t), s));
      SC_HIP(dd.alloc(rows * k * sizeof(double), s));
      SC_HIP(dc.alloc(rows * sizeof(uint32_t), s));
      size_t at = 0;
      for (size_t g = g0; g < g1; ++g) {
        const sc_knn_block& bl = blocks[order[g]];
        if (int rc = upload_padded(bl.queries, index->dtype, bl.q, index->d, index->ld,
                                   dq.as<char>() + at * index->ld * es, s))
          return rc;
        at += bl.q;
      }
      if (int rc = search_device_impl(index, dq.p, rows, k, metric, p, mode, di.as<int64_t>(),
                                      dd.as<double>(), dc.as<uint32_t>(), s, true))
        return rc;
      const sc_stats& st = last_stats();
      total.queries += st.queries;
      total.candidates += st.candidates;
      total.overflow_queries += st.overflow_queries;
      total.shortlist_rescored += st.shortlist_rescored;
      total.uncertified += st.uncertified;
      at = 0;
      for (size_t g = g0; g < g1; ++g) {
        const sc_knn_block& bl = blocks[order[g]];
        if (bl.q > 0) {
          if (int rc = download_staged(bl.idx, di.as<int64_t>() + at * k, bl.q * k * sizeof(int64_t), s))
            return rc;
          if (int rc = download_staged(bl.dist, dd.as<double>() + at * k, bl.q * k * sizeof(double), s))
            return rc;
          if (int rc = download_staged(bl.count, dc.as<uint32_t>() + at, bl.q * sizeof(uint32_t), s))
            return rc;
        }
        at += bl.q;
      }
    }
    g0 = g1;
  }
  SC_HIP(hipStreamSynchronize(s));
  last_stats() = total;
  return SC_OK;
}

size_t sc_last_error(char* buf, size_t len) {
  const std::string& m = t_last_error;
  if (buf && len > 0) {
    const size_t c = std::min(len - 1, m.size());
    std::memcpy(buf, m.data(), c);
    buf[c] = '\0';
  }
  return m.size();
}

int sc_last_stats(sc_stats* out) {
  if (!out) return SC_PARAMETERS_ERROR;
  *out = last_stats();
  return SC_OK;
}

int sc_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (on) prof_clear_locked();
  g_prof_on.store(on != 0);
  return SC_OK;
}

int sc_profile_read(const char* kernel, double* total_ms, uint64_t* launches) {
  clear_error();
  if (!kernel || !total_ms || !launches) return fail(SC_PARAMETERS_ERROR, "null argument");
  std::lock_guard<std::mutex> lk(g_prof_mu);
  double sum = 0.0;
  uint64_t cnt = 0;
  for (auto& r : g_prof) {
    if (r.name != kernel) continue;
    SC_HIP(hipEventSynchronize(r.stop));
    float ms = 0.0f;
    SC_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    sum += ms;
    ++cnt;
  }
  *total_ms = sum;
  *launches = cnt;
  return SC_OK;
}

}  // extern "C"
