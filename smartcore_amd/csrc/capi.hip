// C ABI surface of libsmartcore_amd: index lifetime, kNN search entry
// points, error and stats reporting.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace sc {

namespace {
thread_local std::string t_last_error;
thread_local sc_stats t_stats{};

// One non-blocking stream per host thread and device, at the default
// priority.
hipStream_t thread_stream(int device) {
  static constexpr int kMaxDev = 64;
  thread_local hipStream_t streams[kMaxDev] = {};
  if (device < 0 || device >= kMaxDev) return nullptr;
  if (!streams[device]) {
    if (hipStreamCreateWithFlags(&streams[device], hipStreamNonBlocking) != hipSuccess)
      streams[device] = nullptr;
  }
  return streams[device];
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
  const int code = (e == hipErrorOutOfMemory) ? SC_ERR_OUT_OF_MEMORY : SC_ERR_DEVICE;
  return fail(code, std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
}

int upload_padded(const void* host, sc_dtype dtype, size_t n, size_t d, size_t ld, void* dev,
                  hipStream_t stream) {
  const size_t es = elem_size(dtype);
  if (n == 0) return SC_OK;
  if (ld == d) {
    SC_HIP(hipMemcpyAsync(dev, host, n * d * es, hipMemcpyHostToDevice, stream));
    return SC_OK;
  }
  if (d == 0) {
    SC_HIP(hipMemsetAsync(dev, 0, n * ld * es, stream));
    return SC_OK;
  }
  // one contiguous transfer, then the padding on the device (a strided
  // copy from pageable memory moves row by row)
  DevBuf tmp;
  SC_HIP(tmp.alloc(n * d * es, stream));
  SC_HIP(hipMemcpyAsync(tmp.p, host, n * d * es, hipMemcpyHostToDevice, stream));
  return copy_padded_device(tmp.p, dtype, n, d, ld, dev, stream);
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
  // launches recorded into a graph are timed by the scope around the graph
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

void prof_clear_locked() {
  for (auto& r : g_prof) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  g_prof.clear();
}

hipStream_t stream_for(const sc_index* index) { return thread_stream(index->device); }
hipStream_t stream_for_device(int device) { return thread_stream(device); }

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
  if (mode == SC_MIXED && metric != SC_EUCLIDIAN)
    return fail(SC_ERR_UNSUPPORTED, "SC_MIXED supports the Euclidian metric only");
  return SC_OK;
}

int search_device_impl(const sc_index* index, const void* d_padded_queries, size_t q, size_t k,
                       sc_metric metric, uint16_t p, sc_mode mode, int64_t* d_idx, double* d_dist,
                       uint32_t* d_count, hipStream_t s, bool with_stats) {
  sc_stats& st = last_stats();
  st = sc_stats{};
  sc_stats* sp = with_stats ? &st : nullptr;
  if (mode == SC_MIXED && index->dtype == SC_F32)
    return knn_mixed_device(index, d_padded_queries, q, k, d_idx, d_dist, d_count, s, sp);
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
  SC_HIP(hipMemcpyAsync(idx, di.p, q * k * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(dist, dd.p, q * k * sizeof(double), hipMemcpyDeviceToHost, s));
  SC_HIP(hipMemcpyAsync(count, dc.p, q * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
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
