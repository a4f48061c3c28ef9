// Shared host/device helpers of libsmartcore_amd: status plumbing, HIP
// checks, the device-resident index, and per-call workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <new>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/smartcore_amd.h"

namespace sc {
// A vector allocator whose value-less construct leaves the element
// uninitialised: buffers that a device copy fills at once are not zeroed
// first (tens of milliseconds for the neighbour lists of a large radius
// search).
template <typename T>
struct DefaultInitAllocator : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = DefaultInitAllocator<U>;
  };
  DefaultInitAllocator() = default;
  template <typename U>
  DefaultInitAllocator(const DefaultInitAllocator<U>&) noexcept {}
  template <typename U>
  void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new (static_cast<void*>(p)) U;
  }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
}  // namespace sc

// Result of a radius search: per query, the matching rows in index order
// (CSR offsets over the queries).
struct sc_radius_result {
  size_t q = 0;
  std::vector<uint64_t> offsets;
  std::vector<int64_t, sc::DefaultInitAllocator<int64_t>> idx;
  std::vector<double, sc::DefaultInitAllocator<double>> dist;  // empty when not asked for
};

namespace sc {

// Row stride (elements) used for every row block in HBM: rows are padded
// with zeros up to a multiple of kRowPad so that kernels stage whole
// dimension chunks.  A zero pad adds exact zeros to every distance sum
// (r = 0 - 0 = 0, |0| = 0, 0^p = 0) and therefore leaves it bit-identical.
constexpr int kRowPad = 16;

// A pointer read from memory (a field of a problem struct) has no known
// address space, so its loads compile to FLAT instructions.  Those count in
// LGKM_CNT as well as VM_CNT, so every wait for an LDS result would also
// wait for them (a prefetch would stall the next LDS read).  Device buffers
// are global memory: this says so.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* to_global(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

inline size_t padded_dim(size_t d) {
  const size_t r = (d + kRowPad - 1) / kRowPad * kRowPad;
  return r ? r : kRowPad;
}

struct Status {
  int code = SC_OK;
  std::string msg;
};

// Sets the calling thread's last error and returns code.
int fail(int code, const std::string& msg);
// Clears the calling thread's last error.
void clear_error();
// Thread-local counters of the last search.
sc_stats& last_stats();

// Converts a HIP error to SC_ERR_DEVICE (or SC_ERR_OUT_OF_MEMORY).
int hip_fail(hipError_t e, const char* what);

#define SC_HIP(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return ::sc::hip_fail(_e, #expr); \
  } while (0)

inline size_t elem_size(sc_dtype t) { return t == SC_F64 ? 8 : 4; }

}  // namespace sc

struct sc_index {
  sc_dtype dtype;
  size_t n;
  size_t d;
  size_t ld;        // padded row stride in elements
  int device;       // HIP device ordinal the rows live on
  void* rows;       // n x ld, dtype, zero padded
  // Lazily built derived data for the mixed-precision form (guarded by mu).
  void* rows_bf16;  // n x ldb bf16 (hi part)
  float* norms2;    // n squared norms (FP32 of the exact FP64 value)
  size_t ldb;
  // Lazily built 16-bit quantised copy for the mixed Manhattan form
  // (guarded by mu): rows_q16 holds n x ld u16 codes (two per dword),
  // err_q16 per row ceil(sum_i |y_i - (q16_lo + code_i * q16_step)| / q16_step).
  // q16_state: 0 not built, 1 built, 2 outside the form's domain.
  void* rows_q16;
  uint32_t* err_q16;
  double q16_lo, q16_step;
  int q16_state;
  void* mu;         // std::mutex*
};

namespace sc {

// Scratch blocks kept per library-owned stream (the calling thread's stream
// and the server's slot streams).  A block freed on such a stream goes back
// to that stream's cache and serves a later call on the same stream: reuse
// is ordered by the stream itself, with no allocator call.  Other streams
// (a caller's, a graph being captured) allocate and free stream-ordered.
// A stream keeps at most a few GiB; beyond that blocks are freed.
void scratch_register(hipStream_t s);
// Frees the stream's cached blocks (stream-ordered) and forgets the stream.
void scratch_release(hipStream_t s);
// A cached block of at least `want` bytes for stream s (its size in *cap),
// or null.
void* scratch_take(hipStream_t s, size_t want, size_t* cap);
// Returns the block to s's cache; false when s keeps no cache.
bool scratch_give(hipStream_t s, void* p, size_t cap);
// Allocation size for a request (rounded so blocks serve similar requests).
size_t scratch_round(size_t want);
// Frees every block cached for s (after an allocation failure).
void scratch_trim(hipStream_t s);

// Stream-ordered scratch allocation that is released when the object dies.
struct DevBuf {
  void* p = nullptr;
  hipStream_t s = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p && !scratch_give(s, p, cap)) (void)hipFreeAsync(p, s);
    p = nullptr;
  }
  hipError_t alloc(size_t bytes, hipStream_t stream) {
    release();
    s = stream;
    const size_t want = bytes ? bytes : 16;
    if ((p = scratch_take(stream, want, &cap)) != nullptr) return hipSuccess;
    cap = scratch_round(want);
    hipError_t e = hipMallocAsync(&p, cap, stream);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      scratch_trim(stream);
      e = hipMallocAsync(&p, cap, stream);
    }
    if (e != hipSuccess) p = nullptr;
    return e;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Host <-> device copies of pageable caller memory through two pinned
// staging buffers of the calling thread: the host copies chunk i into one
// buffer (split over a few threads) while the DMA engine moves chunk i - 1
// from the other.  upload_staged returns once the caller's memory has been
// read (the last DMA may still be in flight, from pinned memory);
// download_staged returns once `host` holds the data.
int upload_staged(void* dev, const void* host, size_t bytes, hipStream_t stream);
// n rows of row_bytes to the device at a stride of ld_bytes (the padding is
// zeroed on the host side of the copy).
int upload_staged_rows(void* dev, const void* host, size_t n, size_t row_bytes, size_t ld_bytes,
                       hipStream_t stream);
int download_staged(void* host, const void* dev, size_t bytes, hipStream_t stream);
// Several host row blocks, each widened to its own stride, placed back to
// back on the device from `dev`: rows of consecutive blocks share staging
// buffers, so many small blocks move as a few large copies.
struct StageSeg {
  const void* host;
  size_t n, row_bytes, ld_bytes;
};
int upload_staged_segments(void* dev, const StageSeg* segs, size_t count, hipStream_t stream);

// Copies `n` host rows of `d` elements into a zero-padded device block with
// stride ld (async on `stream`).
int upload_padded(const void* host, sc_dtype dtype, size_t n, size_t d, size_t ld, void* dev,
                  hipStream_t stream);
// Same for device-resident rows (row stride d).
int copy_padded_device(const void* src, sc_dtype dtype, size_t n, size_t d, size_t ld, void* dev,
                       hipStream_t stream);

// Kernel timing for the tracing hooks (sc_profile_*): while enabled, a
// ProfScope records a HIP event pair on `stream` around the launches made in
// its lifetime and files them under `name`.
class ProfScope {
 public:
  ProfScope(const char* name, hipStream_t stream);
  ~ProfScope();
  ProfScope(const ProfScope&) = delete;
  ProfScope& operator=(const ProfScope&) = delete;

 private:
  const char* name_;
  hipStream_t stream_;
  hipEvent_t start_ = nullptr;
};

// Whether the tracing hooks are on (sc_profile_enable).
bool profiling_enabled();

// The calling host thread's stream for the index's device.
hipStream_t stream_for(const sc_index* index);
hipStream_t stream_for_device(int device);
// A second stream of the calling host thread for the device, for host-to-
// device copies that overlap the kernels of the first (no scratch cache).
hipStream_t copy_stream_for_device(int device);

// ---- kernels entry points shared between translation units ----

// Exact (reference arithmetic) kNN over an index; queries are padded device
// rows (stride index->ld).  Outputs device arrays of q*k / q.
int knn_exact_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                     sc_metric metric, int p, int64_t* d_idx, double* d_dist, uint32_t* d_count,
                     hipStream_t stream, sc_stats* stats);

// SC_MIXED Manhattan: bounded f32 scan, candidates rescored in the
// reference arithmetic (f32 rows; f64 rows take the exact form).
int knn_l1_mixed_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                        int64_t* d_idx, double* d_dist, uint32_t* d_count, hipStream_t stream,
                        sc_stats* stats);

int knn_mixed_device(const sc_index* index, const void* d_queries, size_t q, size_t k,
                     int64_t* d_idx, double* d_dist, uint32_t* d_count, hipStream_t stream,
                     sc_stats* stats);

// find_radius (linear_search.rs:89-108) for q padded device queries (stride
// index->ld), all at once: res gets each query's rows in index order.
// want_dist false: res->dist stays empty (callers that need the rows only).
int radius_search_device(const sc_index* index, const void* d_queries, int64_t q,
                         sc_metric metric, int p, double radius, sc_radius_result* res,
                         hipStream_t stream, bool want_dist = true);

}  // namespace sc
