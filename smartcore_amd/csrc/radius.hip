// Brute-force radius search (LinearKNNSearch::find_radius,
// linear_search.rs:89-108): every row with distance <= radius, in index order.
//
// Two passes over the same tiled distance block as the kNN scan: `count`
// tallies matches per (query, segment); an exclusive scan turns the tallies
// into CSR offsets; `fill` recomputes the block and writes each query's
// matches of a tile in index order at its running offset.  When every query
// block scans all rows (one segment: large query sets, such as DBSCAN's
// search of every row), one `capped` pass instead writes each query's
// matches, in index order, to a slot of kCap entries while counting all of
// them; a compaction kernel moves the slots to CSR order, and the queries
// with more than kCap matches are searched again as a subset with the two
// passes.  The lists are the same either way.  A row is tested
// with a conservative bound on the sum first and then exactly on the
// finalised distance (sqrt / identity / powf(1/p)), so the decision is the
// reference's `d <= radius`.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "metric.hpp"
#include "tile.hpp"

namespace sc {
namespace {

constexpr int QT = kTileQ;
constexpr int RT = kTileR;
constexpr int NT = kTileThreads;

// Kernel modes: count matches per (query, segment); fill them at CSR
// offsets; or (one segment) write them to a kCap-entry slot per query while
// counting all.
constexpr int kCount = 0, kFill = 1, kCapped = 2;
constexpr int64_t kCap = 256;
// Largest capped buffer (index and distance slots) one search may take.
constexpr double kCapBytes = 1024.0 * 1024.0 * 1024.0;

// DIST: the matches' distances are written too (their LDS buffer, 32 KB,
// would otherwise cost occupancy).
template <typename T, int M, int MODE, bool DIST>
__global__ __launch_bounds__(NT) void radius_kernel(const T* __restrict__ refs, int64_t n, int ld, int d,
                                                    const T* __restrict__ queries, int64_t q,
                                                    double p, double radius, double sum_bound,
                                                    int64_t seg_len, int nseg,
                                                    uint64_t* __restrict__ counts,
                                                    int64_t* __restrict__ out_idx,
                                                    double* __restrict__ out_dist, int64_t cap) {
  __shared__ TileSmem<T> sm;
  __shared__ double dbuf[DIST ? QT : 1][DIST ? RT : 1];
  __shared__ uint32_t mask[QT][RT / 32];
  const int tid = threadIdx.x;
  const int tq = tid >> 5;
  const int tr = tid & 31;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * QT;
  const int seg = blockIdx.y;
  const int64_t r_begin = static_cast<int64_t>(seg) * seg_len;
  const int64_t r_end = min(n, r_begin + seg_len);
  for (int e = tid; e < QT * (RT / 32); e += NT) (&mask[0][0])[e] = 0u;
  const int64_t my_q = q0 + tid;
  uint64_t run = 0;
  if (MODE == kFill && tid < QT && my_q < q) run = counts[static_cast<size_t>(my_q) * nseg + seg];

  for (int64_t t0 = r_begin; t0 < r_end; t0 += RT) {
    double acc[4][4];
    tile_sums<T, M>(sm, acc, refs, t0, r_end, queries, q0, q, ld, d, p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int qi = tq * 4 + i;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int rl = tr + 32 * j;
        if (q0 + qi < q && t0 + rl < r_end && acc[i][j] <= sum_bound) {
          const double dv = finalize<M>(acc[i][j], p);
          if (dv <= radius) {
            if constexpr (DIST) dbuf[qi][rl] = dv;
            atomicOr(&mask[qi][rl >> 5], 1u << (rl & 31));
          }
        }
      }
    }
    __syncthreads();
    if (tid < QT && my_q < q) {
      for (int w = 0; w < RT / 32; ++w) {
        uint32_t bits = mask[tid][w];
        mask[tid][w] = 0u;
        if (MODE == kCount) {
          run += static_cast<uint64_t>(__popc(bits));
          continue;
        }
        while (bits) {
          const int b = __builtin_ctz(bits);
          bits &= bits - 1;
          const int rl = w * 32 + b;
          if (MODE == kFill) {
            out_idx[run] = t0 + rl;
            if constexpr (DIST) out_dist[run] = dbuf[tid][rl];
          } else if (run < static_cast<uint64_t>(cap)) {
            const size_t o = static_cast<size_t>(my_q) * static_cast<size_t>(cap) + run;
            out_idx[o] = t0 + rl;
            if constexpr (DIST) out_dist[o] = dbuf[tid][rl];
          }
          ++run;
        }
      }
    }
  }
  if (MODE != kFill && tid < QT && my_q < q) counts[static_cast<size_t>(my_q) * nseg + seg] = run;
}

// Slots of the capped pass to CSR order: query i with at most cap matches
// moves them to offsets[i] (one wavefront per query); queries with more are
// left to their own search.
__global__ void radius_compact_kernel(const int64_t* __restrict__ cidx,
                                      const double* __restrict__ cdist, int64_t cap,
                                      const uint64_t* __restrict__ counts,
                                      const uint64_t* __restrict__ offsets, int64_t q,
                                      int64_t* __restrict__ out_idx, double* __restrict__ out_dist) {
  const int lane = threadIdx.x & 63;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (i >= q) return;
  const uint64_t c = counts[i];
  if (c > static_cast<uint64_t>(cap)) return;
  const uint64_t o = offsets[i];
  const size_t base = static_cast<size_t>(i) * static_cast<size_t>(cap);
  for (uint64_t j = lane; j < c; j += 64) {
    out_idx[o + j] = cidx[base + j];
    if (out_dist) out_dist[o + j] = cdist[base + j];
  }
}

// Rows list[0 .. m) of src (stride ld), gathered into dst.
template <typename T>
__global__ void gather_rows_kernel(const T* __restrict__ src, int ld,
                                   const int64_t* __restrict__ list, int64_t m,
                                   T* __restrict__ dst) {
  const int64_t total = m * ld;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t r = e / ld;
    dst[e] = src[list[r] * ld + (e - r * ld)];
  }
}

template <typename T, int M>
int run_radius(const sc_index* ix, const T* dq, int64_t q, double p, double radius,
               sc_radius_result* res, bool want_dist, bool allow_cap, hipStream_t s);

// The one-segment search: capped pass, compaction, and the queries past the
// cap searched again (count and fill passes) and merged in.
template <typename T, int M>
int run_radius_capped(const sc_index* ix, const T* dq, int64_t q, double p, double radius,
                      double bound, sc_radius_result* res, bool want_dist, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(ix->n);
  const int ld = static_cast<int>(ix->ld);
  const int64_t nqt = (q + QT - 1) / QT;
  const int64_t tiles = std::max<int64_t>(1, (n + RT - 1) / RT);
  DevBuf cnt, ci, cd;
  SC_HIP(cnt.alloc(sizeof(uint64_t) * q, s));
  SC_HIP(ci.alloc(sizeof(int64_t) * q * kCap, s));
  if (want_dist) SC_HIP(cd.alloc(sizeof(double) * q * kCap, s));
  {
    ProfScope ps("radius_capped", s);
    if (want_dist)
      hipLaunchKernelGGL((radius_kernel<T, M, kCapped, true>), dim3(static_cast<unsigned>(nqt), 1),
                         dim3(NT), 0, s, static_cast<const T*>(ix->rows), n, ld,
                         static_cast<int>(ix->d), dq, q, p, radius, bound, tiles * RT, 1,
                         cnt.as<uint64_t>(), ci.as<int64_t>(), cd.as<double>(), kCap);
    else
      hipLaunchKernelGGL((radius_kernel<T, M, kCapped, false>), dim3(static_cast<unsigned>(nqt), 1),
                         dim3(NT), 0, s, static_cast<const T*>(ix->rows), n, ld,
                         static_cast<int>(ix->d), dq, q, p, radius, bound, tiles * RT, 1,
                         cnt.as<uint64_t>(), ci.as<int64_t>(), nullptr, kCap);
    SC_HIP(hipGetLastError());
  }
  std::vector<uint64_t> hc(static_cast<size_t>(q));
  SC_HIP(hipMemcpyAsync(hc.data(), cnt.p, sizeof(uint64_t) * hc.size(), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  res->q = static_cast<size_t>(q);
  res->offsets.assign(static_cast<size_t>(q) + 1, 0);
  uint64_t total = 0;
  std::vector<int64_t> over;  // queries with more than kCap matches
  for (int64_t i = 0; i < q; ++i) {
    res->offsets[i] = total;
    total += hc[i];
    if (hc[i] > static_cast<uint64_t>(kCap)) over.push_back(i);
  }
  res->offsets[q] = total;
  res->idx.resize(total);
  if (want_dist) res->dist.resize(total);
  if (total == 0) return SC_OK;
  {
    DevBuf off, oi, od;
    SC_HIP(off.alloc(sizeof(uint64_t) * (q + 1), s));
    SC_HIP(hipMemcpyAsync(off.p, res->offsets.data(), sizeof(uint64_t) * (q + 1),
                          hipMemcpyHostToDevice, s));
    SC_HIP(oi.alloc(sizeof(int64_t) * total, s));
    if (want_dist) SC_HIP(od.alloc(sizeof(double) * total, s));
    constexpr int kWaves = 4;
    hipLaunchKernelGGL(radius_compact_kernel, dim3(static_cast<unsigned>((q + kWaves - 1) / kWaves)),
                       dim3(64 * kWaves), 0, s, ci.as<int64_t>(),
                       want_dist ? cd.as<double>() : nullptr, kCap, cnt.as<uint64_t>(),
                       off.as<uint64_t>(), q, oi.as<int64_t>(), want_dist ? od.as<double>() : nullptr);
    SC_HIP(hipGetLastError());
    if (int rc = download_staged(res->idx.data(), oi.p, sizeof(int64_t) * total, s)) return rc;
    if (want_dist)
      if (int rc = download_staged(res->dist.data(), od.p, sizeof(double) * total, s)) return rc;
  }
  if (!over.empty()) {
    const int64_t m = static_cast<int64_t>(over.size());
    DevBuf dl, dsub;
    SC_HIP(dl.alloc(sizeof(int64_t) * m, s));
    SC_HIP(hipMemcpyAsync(dl.p, over.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, s));
    SC_HIP(dsub.alloc(sizeof(T) * m * ld, s));
    const int64_t g = std::min<int64_t>(4096, (m * ld + 255) / 256);
    hipLaunchKernelGGL((gather_rows_kernel<T>), dim3(static_cast<unsigned>(std::max<int64_t>(1, g))),
                       dim3(256), 0, s, dq, ld, dl.as<int64_t>(), m, dsub.as<T>());
    SC_HIP(hipGetLastError());
    sc_radius_result sub;
    if (int rc = run_radius<T, M>(ix, dsub.as<T>(), m, p, radius, &sub, want_dist, false, s))
      return rc;
    for (int64_t j = 0; j < m; ++j) {
      const int64_t i = over[j];
      const uint64_t b = sub.offsets[j], c = sub.offsets[j + 1] - b;
      if (c != hc[i]) return fail(SC_ERR_DEVICE, "radius search: a query's match count differs between passes");
      std::copy(sub.idx.begin() + b, sub.idx.begin() + b + c, res->idx.begin() + res->offsets[i]);
      if (want_dist)
        std::copy(sub.dist.begin() + b, sub.dist.begin() + b + c, res->dist.begin() + res->offsets[i]);
    }
  }
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

template <typename T, int M>
int run_radius(const sc_index* ix, const T* dq, int64_t q, double p, double radius,
               sc_radius_result* res, bool want_dist, bool allow_cap, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(ix->n);
  const int ld = static_cast<int>(ix->ld);
  const int64_t nqt = (q + QT - 1) / QT;
  const int64_t tiles = std::max<int64_t>(1, (n + RT - 1) / RT);
  int64_t nseg = std::min<int64_t>(std::max<int64_t>(1, (4096 + nqt - 1) / nqt), tiles);
  nseg = std::min<int64_t>(nseg, 65535);
  const int64_t seg_tiles = (tiles + nseg - 1) / nseg;
  nseg = (tiles + seg_tiles - 1) / seg_tiles;
  const int64_t seg_len = seg_tiles * RT;
  // Conservative bound on the sum for the exact test d <= radius.
  double bound;
  if (M == kEuclid)
    bound = radius * radius * (1.0 + 1e-12) + 1e-300;
  else if (M == kManhattan)
    bound = radius;
  else
    bound = std::pow(radius, p) * (1.0 + 1e-9) + 1e-300;
  if (!std::isfinite(bound)) bound = INFINITY;
  if (allow_cap && nseg == 1 &&
      static_cast<double>(q) * static_cast<double>(kCap) * (want_dist ? 16.0 : 8.0) <= kCapBytes)
    return run_radius_capped<T, M>(ix, dq, q, p, radius, bound, res, want_dist, s);

  DevBuf cnt;
  SC_HIP(cnt.alloc(sizeof(uint64_t) * q * nseg, s));
  dim3 grid(static_cast<unsigned>(nqt), static_cast<unsigned>(nseg));
  hipLaunchKernelGGL((radius_kernel<T, M, kCount, false>), grid, dim3(NT), 0, s, static_cast<const T*>(ix->rows),
                     n, ld, static_cast<int>(ix->d), dq, q, p, radius, bound, seg_len, static_cast<int>(nseg),
                     cnt.as<uint64_t>(), nullptr, nullptr, int64_t{0});
  SC_HIP(hipGetLastError());
  std::vector<uint64_t> hc(static_cast<size_t>(q * nseg));
  SC_HIP(hipMemcpyAsync(hc.data(), cnt.p, sizeof(uint64_t) * hc.size(), hipMemcpyDeviceToHost, s));
  SC_HIP(hipStreamSynchronize(s));
  // exclusive scan in (query, segment) order = CSR offsets in index order
  res->q = static_cast<size_t>(q);
  res->offsets.assign(static_cast<size_t>(q) + 1, 0);
  uint64_t total = 0;
  for (int64_t i = 0; i < q; ++i) {
    res->offsets[i] = total;
    for (int64_t sg = 0; sg < nseg; ++sg) {
      const uint64_t c = hc[i * nseg + sg];
      hc[i * nseg + sg] = total;
      total += c;
    }
  }
  res->offsets[q] = total;
  res->idx.resize(total);
  if (want_dist) res->dist.resize(total);
  if (total == 0) return SC_OK;
  DevBuf oi, od;
  SC_HIP(oi.alloc(sizeof(int64_t) * total, s));
  if (want_dist) SC_HIP(od.alloc(sizeof(double) * total, s));
  SC_HIP(hipMemcpyAsync(cnt.p, hc.data(), sizeof(uint64_t) * hc.size(), hipMemcpyHostToDevice, s));
  if (want_dist)
    hipLaunchKernelGGL((radius_kernel<T, M, kFill, true>), grid, dim3(NT), 0, s,
                       static_cast<const T*>(ix->rows), n, ld, static_cast<int>(ix->d), dq, q, p,
                       radius, bound, seg_len, static_cast<int>(nseg), cnt.as<uint64_t>(),
                       oi.as<int64_t>(), od.as<double>(), int64_t{0});
  else
    hipLaunchKernelGGL((radius_kernel<T, M, kFill, false>), grid, dim3(NT), 0, s,
                       static_cast<const T*>(ix->rows), n, ld, static_cast<int>(ix->d), dq, q, p,
                       radius, bound, seg_len, static_cast<int>(nseg), cnt.as<uint64_t>(),
                       oi.as<int64_t>(), nullptr, int64_t{0});
  SC_HIP(hipGetLastError());
  // staged through the pinned buffers (parallel host copies)
  if (int rc = download_staged(res->idx.data(), oi.p, sizeof(int64_t) * total, s)) return rc;
  if (want_dist)
    if (int rc = download_staged(res->dist.data(), od.p, sizeof(double) * total, s)) return rc;
  SC_HIP(hipStreamSynchronize(s));
  return SC_OK;
}

template <typename T>
int radius_dispatch(const sc_index* ix, const void* dq, int64_t q, sc_metric metric, int p,
                    double radius, sc_radius_result* res, bool want_dist, hipStream_t s) {
  const T* qs = static_cast<const T*>(dq);
  switch (metric) {
    case SC_EUCLIDIAN:
      return run_radius<T, kEuclid>(ix, qs, q, 2.0, radius, res, want_dist, true, s);
    case SC_MANHATTAN:
      return run_radius<T, kManhattan>(ix, qs, q, 1.0, radius, res, want_dist, true, s);
    case SC_MINKOWSKI:
      return run_radius<T, kMinkowski>(ix, qs, q, static_cast<double>(p), radius, res, want_dist,
                                         true, s);
  }
  return fail(SC_ERR_UNSUPPORTED, "unknown metric");
}

}  // namespace

int radius_search_device(const sc_index* index, const void* d_queries, int64_t q,
                         sc_metric metric, int p, double radius, sc_radius_result* res,
                         hipStream_t s, bool want_dist) {
  res->q = static_cast<size_t>(q);
  res->offsets.assign(static_cast<size_t>(q) + 1, 0);
  res->idx.clear();
  res->dist.clear();
  if (q == 0 || index->n == 0) return SC_OK;
  return index->dtype == SC_F64
             ? radius_dispatch<double>(index, d_queries, q, metric, p, radius, res, want_dist, s)
             : radius_dispatch<float>(index, d_queries, q, metric, p, radius, res, want_dist, s);
}

}  // namespace sc

using namespace sc;

extern "C" {

int sc_radius_search(const sc_index* index, const void* queries, size_t q, double radius,
                     sc_metric metric, uint16_t p, sc_radius_result** out) {
  clear_error();
  if (!index || !out) return fail(SC_PARAMETERS_ERROR, "null argument");
  *out = nullptr;
  // linear_search.rs:90-95
  if (!(radius > 0.0)) return fail(SC_FIND_FAILED, "radius should be > 0");
  if (metric != SC_EUCLIDIAN && metric != SC_MANHATTAN && metric != SC_MINKOWSKI)
    return fail(SC_ERR_UNSUPPORTED, "unknown metric");
  if (metric == SC_MINKOWSKI && p < 1) return fail(SC_ERR_PANIC, "p must be at least 1");
  auto* res = new sc_radius_result();
  res->q = q;
  res->offsets.assign(q + 1, 0);
  if (q == 0 || index->n == 0) {
    *out = res;
    return SC_OK;
  }
  if (!queries) {
    delete res;
    return fail(SC_PARAMETERS_ERROR, "null queries");
  }
  hipError_t e = hipSetDevice(index->device);
  if (e != hipSuccess) {
    delete res;
    return hip_fail(e, "hipSetDevice");
  }
  hipStream_t s = stream_for(index);
  int rc;
  {
    DevBuf dq;
    e = dq.alloc(q * index->ld * elem_size(index->dtype), s);
    if (e != hipSuccess) {
      delete res;
      return hip_fail(e, "alloc queries");
    }
    rc = upload_padded(queries, index->dtype, q, index->d, index->ld, dq.p, s);
    if (rc == SC_OK)
      rc = radius_search_device(index, dq.p, static_cast<int64_t>(q), metric, p, radius, res, s);
  }
  if (rc != SC_OK) {
    delete res;
    return rc;
  }
  *out = res;
  return SC_OK;
}

size_t sc_radius_result_total(const sc_radius_result* r) { return r ? r->idx.size() : 0; }

int sc_radius_result_copy(const sc_radius_result* r, uint64_t* offsets, int64_t* idx,
                          double* dist) {
  clear_error();
  if (!r) return fail(SC_PARAMETERS_ERROR, "null result");
  if (offsets) std::copy(r->offsets.begin(), r->offsets.end(), offsets);
  if (idx) std::copy(r->idx.begin(), r->idx.end(), idx);
  if (dist) std::copy(r->dist.begin(), r->dist.end(), dist);
  return SC_OK;
}

void sc_radius_result_free(sc_radius_result* r) { delete r; }

}  // extern "C"
